"""CPU checks of the algebraic rewrites the pair kernels use in place of the
reference's expressions (the GPU parity tests check the kernels themselves).

* The cubic-spline gradient (Functions.cu:76-88): the reference branches on
  q < 1 / q < 2; the kernels use 3 (u^2 - t^2) with u = clamp(1 - q, 0, 1)
  and t = clamp(1 - q/2, 0, 1), which needs no branch.
* The Monaghan viscosity without the beta term (InteractionAlt.cu:18-150):
  alpha C_ij mu / rho_bar with C_ij = (C_i + C_j)/2, mu = h v/e and
  rho_bar = (rho_i + rho_j)/2 equals v h alpha (C_i + C_j) / (e (rho_i + rho_j)).
"""
import numpy as np


def dwdq_reference(q):
    # Functions.cu:76-88, the bracket of dW/dq (without the 1/(pi h^4) factor)
    q = np.asarray(q, np.float64)
    return np.where(q < 1.0, -3.0 * q + 2.25 * q * q,
                    np.where(q < 2.0, -0.75 * (2.0 - q) ** 2, 0.0))


def dwdq_branch_free(q):
    u = np.clip(1.0 - q, 0.0, 1.0)
    t = np.clip(1.0 - 0.5 * q, 0.0, 1.0)
    return 3.0 * (u - t) * (u + t)


def test_spline_gradient_branch_free_form():
    q = np.concatenate([np.linspace(0.0, 2.6, 20001), [1.0, 2.0, 1e-12, 2.0 - 1e-12]])
    ref, new = dwdq_reference(q), dwdq_branch_free(q)
    assert np.all(new[q >= 2.0] == 0.0)
    assert new[q == 0.0][0] == 0.0
    scale = np.maximum(np.abs(ref), 1e-300)
    assert np.max(np.abs(new - ref) / scale * (np.abs(ref) > 1e-6)) < 1e-12
    assert np.max(np.abs(new - ref)) < 1e-14


def test_spline_gradient_branch_free_form_fp32():
    # the FP32 form against the FP64 reference: a few ulp of the bracket
    q = np.linspace(0.05, 1.995, 4001).astype(np.float32)
    one, half = np.float32(1.0), np.float32(0.5)
    u = np.clip(one - q, 0, 1).astype(np.float32)
    t = np.clip(one - half * q, 0, 1).astype(np.float32)
    new = np.float32(3.0) * ((u - t) * (u + t))
    ref = dwdq_reference(q.astype(np.float64))
    assert np.max(np.abs(new - ref)) < 4 * np.finfo(np.float32).eps * 0.75


def test_viscosity_without_beta_matches_reference_form():
    rng = np.random.default_rng(3)
    n = 1000
    h = rng.uniform(1e-4, 1e-3, n)
    ci, cj = rng.uniform(3e3, 6e3, n), rng.uniform(3e3, 6e3, n)
    ri, rj = rng.uniform(2e3, 9e3, n), rng.uniform(2e3, 9e3, n)
    r2 = (rng.uniform(0.5, 2.0, n) * h) ** 2
    v = -rng.uniform(0.0, 5.0, n) * h   # approaching pairs: v_ij . x_ij < 0
    alpha = 1.0
    e = r2 + 0.01 * h * h
    mu = h * v / e
    ref = alpha * 0.5 * (ci + cj) * mu / (0.5 * (ri + rj))
    new = v * (h * alpha * (ci + cj)) / (e * (ri + rj))
    assert np.max(np.abs(new - ref) / np.abs(ref)) < 1e-13
