"""Shared helpers of the parity tests: load one scenario into the oracle and
into the HIP domain, perturb it with seeded random states, compare fields."""
import numpy as np

from weldformgpu_amd import _abi, scenarios

ORACLE_NAME = {"x_snapshot": "xs"}

# fields compared after the mechanical kernels
MECH_FIELDS = ["x", "v", "a", "u", "rho", "drho", "p", "sigma", "shearstress", "strrate",
               "rotrate", "sigma_y", "pl_strain", "sigma_eq", "eff_strain_rate", "int_energy"]


def to_oracle(po, sc):
    st = po.State(sc.n)
    for k, v in sc.fields.items():
        st[ORACLE_NAME.get(k, k)] = v
    st["xs"] = sc.fields["x"]
    if sc.meshes:
        st.add_contact(sc.meshes)
    return st


def to_gpu(sc, lib=None):
    from weldformgpu_amd.domain import Domain
    d = Domain(sc.cfg, lib)
    scenarios.load_domain(d, sc)
    return d


def randomize(sc, seed=1, v_scale=1.0, s_scale=300e6, shear=True):
    """Seeded random velocity U(-v,v) and deviatoric stress U(-s,s) (symmetric
    6-vectors), sigma = -p I + s with p from the EOS of rho."""
    rng = np.random.default_rng(seed)
    n = sc.n
    sc.fields["v"] = rng.uniform(-v_scale, v_scale, (n, 3))
    if shear:
        ss = rng.uniform(-s_scale, s_scale, (n, 6))
        ss[:, 2] = -(ss[:, 0] + ss[:, 1])
        sc.fields["shearstress"] = ss
        sig = ss.copy()
        p = rng.uniform(-s_scale, s_scale, n)
        sig[:, :3] -= p[:, None]
        sc.fields["sigma"] = sig
    return sc


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0.0:
        return num
    return num / den


def max_rel(a, b, floor):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def small_c1(scale=0.4):
    return scenarios.c1_bilinear(scale=scale)


def sorted_rows(counts, offsets, lst):
    return [np.sort(lst[offsets[i]:offsets[i] + counts[i]]) for i in range(len(counts))]
