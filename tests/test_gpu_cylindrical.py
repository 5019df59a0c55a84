"""The cylindrical grid (Domain::AddCylUniformLength, Domain.cpp:1071-1193)
and the axisymmetric acceleration clamp (ApplyAxiSymmBCKernel,
Domain_d.cu:665-695, applied after the acceleration pass,
SolverFraser.cu:229-231) on the GPU against the FP64 oracle.

  - the clamp entry point alone, on seeded accelerations over the whole
    circle: it only writes zeros where the condition holds, so it is
    bit-exact;
  - a bilinear steel slice (the pi/16 slice of the outer half of a 15 mm
    radius, 3 mm high, the reference's grid for a sliceAngle other than
    exactly 2 pi, with the lattice's particle mass) compressed at 1 m/s for
    60 steps with the clamp on, and the whole 3 mm-radius circle on the same
    grid without it, in FP64 (relative L2 <= 1e-9, the step tolerance) and
    mixed precision (<= 1e-4).  Over those steps the clamp changes the
    oracle's slice by far more than the tolerance, so a GPU that skipped it
    would fail.  The circle's axis particles have 78 neighbours, over the
    reference's cap of 70: every one is kept, as in the oracle.
"""
import math

import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import rel_l2, to_gpu, to_oracle

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "sigma", "pl_strain"]


def _case(kind):
    mat = {"model": _abi.WF_MAT_BILINEAR, "sy0": 300e6, "Et": 1e9}
    if kind == "slice":
        sc = scenarios.cylinder_compression(15e-3, 3e-3, 0.15e-3, rho=7850.0, E=200e9, nu=0.3,
                                            material=mat, v_top=-1.0,
                                            slice_angle=2.0000001 * math.pi)
        # the grid's own mass is the whole cylinder's over the slice's count,
        # 43 times a particle's share, and that run diverges within five
        # steps (the deck test covers it); the lattice mass rho (2r)^3 keeps
        # the slice stable over the compared steps
        sc.fields["m"][:] = 7850.0 * (0.3e-3) ** 3
        return sc
    return scenarios.cylinder_compression(3e-3, 3e-3, 0.15e-3, rho=7850.0, E=200e9, nu=0.3,
                                          material=mat, v_top=-1.0, slice_angle=2.0 * math.pi)


def test_clamp_entry_point_bit_exact(oracle, wflib):
    # the whole circle: x spans -R..R, so the x > h/2 test splits the grid
    sc = _case("full")
    assert sc.cfg.axisymm == 0
    n = sc.n
    rng = np.random.default_rng(11)
    a = rng.normal(0.0, 1.0, (n, 3)) * rng.choice([1e-4, 5e-3, 1.0, 1.0], (n, 1))
    x = sc.fields["x"]
    print(f"[case] axisymm clamp entry point n={n}", flush=True)
    st = to_oracle(oracle, sc)
    st["a"] = a
    st.apply_axisymm_bc()
    with to_gpu(sc, wflib) as d:
        d.set("a", a)
        d.ApplyAxiSymmBC()
        got = d.get("a")
    hit = (np.hypot(a[:, 0], a[:, 1]) > 1e-3) & (x[:, 0] > sc.cfg.h_glob / 2) & (a[:, 0] > 1e-2)
    assert 0.05 * n < hit.sum() < 0.5 * n
    assert np.array_equal(got, st["a"])


@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-9), (_abi.WF_MIXED, 1e-4)])
@pytest.mark.parametrize("kind", ["slice", "full"])
def test_cylindrical_grid_steps_match_oracle(oracle, wflib, kind, precision, tol):
    sc = _case(kind)
    nsteps = 60
    print(f"[case] cylindrical grid {kind} prec={precision} n={sc.n} axisymm={sc.cfg.axisymm}",
          flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.step(sc.cfg, sc.bcs, nsteps)
    if kind == "slice":
        free = to_oracle(oracle, sc)
        cfg0 = _case(kind).cfg
        cfg0.axisymm = 0
        free.nsearch(cfg0)
        free.step(cfg0, sc.bcs, nsteps)
        effect = rel_l2(free["v"], st["v"])
        print(f"  clamp effect on v: {effect:.2e}", flush=True)
        assert effect > 1e3 * tol
    sc.cfg.precision = precision
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        got = {k: d.get(k) for k in FIELDS}
    for k in FIELDS:
        if k == "pl_strain" and not st[k].any():
            assert not got[k].any()
            continue
        err = rel_l2(got[k], st[k])
        print(f"  {k}: {err:.2e}", flush=True)
        assert err < tol, k
