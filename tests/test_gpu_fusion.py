"""A step's conduction pass fused with the next step's density pass.

Inside one wf_step call the two share one traversal of the neighbour lists
whenever no rebuild runs between them.  A 12-step call (fused) is compared
with 12 single-step calls (never fused: a call's last step has no next step
in the call), in both precisions, with the stale-snapshot quirk (no rebuild
runs) and without it (rebuilds every 5 steps once yielding, where the fusion
must stop).  The FP64 run with rebuilds is also compared with the oracle at
the step tolerance 1e-9.  FP64 fused vs unfused: relative L2 <= 1e-13; mixed
<= 1e-6 (the two paths evaluate the same FP32 pair terms in the same order;
the bound leaves room for contraction differences between the kernels).
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import randomize, rel_l2, to_gpu, to_oracle

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "rho", "sigma", "pl_strain", "T", "int_energy"]
NSTEPS = 12


def _case(prec, stale):
    sc = randomize(scenarios.c3_johnson_cook(scale=0.2), seed=11, v_scale=0.5, s_scale=200e6)
    sc.cfg.precision = prec
    if not stale:
        sc.cfg.quirks &= ~_abi.WF_Q_STALE_NS_POS
    return sc


def _run(sc, wflib, single):
    with to_gpu(sc, wflib) as d:
        d.set_timing(True)
        if single:
            for _ in range(NSTEPS):
                d.step(1)
        else:
            d.step(NSTEPS)
        fused = d.kernel_time("cond_dens")[1]
        dens = d.kernel_time("dens")[1]
        out = {k: d.get(k) for k in FIELDS}
        stats = d.stats()
    return out, fused, dens, stats


@pytest.mark.parametrize("stale", [True, False])
@pytest.mark.parametrize("prec", [_abi.WF_FP64, _abi.WF_MIXED])
def test_fused_call_matches_single_steps(wflib, prec, stale):
    print(f"[case] fusion prec={prec} stale={stale}", flush=True)
    a, fused, dens, sa = _run(_case(prec, stale), wflib, single=False)
    b, fused1, dens1, sb = _run(_case(prec, stale), wflib, single=True)
    assert fused1 == 0 and dens1 == NSTEPS
    assert fused + dens == NSTEPS  # one density pass per step
    assert sa["newton_calls"] > 0 and sa["is_yielding"]
    if stale:
        assert fused == NSTEPS - 1
    else:
        assert 0 < fused < NSTEPS - 1  # no fusion across the rebuilds
    tol = 1e-13 if prec == _abi.WF_FP64 else 1e-6
    for k in FIELDS:
        err = rel_l2(a[k], b[k])
        print(f"  {k}: {err:.2e}")
        assert err <= tol, k


def test_fused_steps_with_rebuilds_match_oracle(oracle, wflib):
    sc = _case(_abi.WF_FP64, stale=False)
    print("[case] fusion with rebuilds vs oracle", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.step(sc.cfg, sc.bcs, NSTEPS)
    with to_gpu(sc, wflib) as d:
        d.set_timing(True)
        d.step(NSTEPS)
        assert d.kernel_time("cond_dens")[1] > 0
        assert d.stats()["ns_builds"] >= 2
        for k in FIELDS:
            assert rel_l2(d.get(k), st[k]) < 1e-9, k
        assert np.all(np.isfinite(d.get("T")))


@pytest.mark.parametrize("stale", [True, False])
def test_accel_integrate_fusion_is_bit_identical(wflib, stale):
    """Mixed precision: the acceleration pass that integrates each particle as
    soon as its sum is done (cfg.fuse_integrate, the default) against the
    separate integration pass (fuse_integrate = 0).  Both round the same
    arithmetic in the same order (integrate_particle, no FMA contraction), so
    every field is bit-identical; only the kernel ids the timing reports
    differ."""
    runs = {}
    for fuse in (1, 0):
        sc = _case(_abi.WF_MIXED, stale)
        sc.cfg.fuse_integrate = fuse
        print(f"[case] accel+integrate fusion {fuse} stale={stale}", flush=True)
        with to_gpu(sc, wflib) as d:
            d.set_timing(True)
            d.step(NSTEPS)
            runs[fuse] = ({k: d.get(k) for k in FIELDS + ["x", "a"]},
                          d.kernel_time("accel_integrate")[1], d.kernel_time("integrate")[1],
                          d.stats())
    (a, fa, ia, sa), (b, fb, ib, sb) = runs[1], runs[0]
    assert fa == NSTEPS and ia == 0
    assert fb == 0 and ib == NSTEPS
    assert sa["newton_calls"] > 0 and sa["ns_builds"] == sb["ns_builds"]
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("prec", [_abi.WF_FP64, _abi.WF_MIXED])
def test_fused_call_with_plastic_zone_trigger(wflib, prec):
    """Under the plastic-zone trigger the fused pass is queued before the host
    has read whether the next step starts with a rebuild.  Its density half
    reads the trigger from device memory and writes nothing when a rebuild is
    due; the host reads the same value behind the pass and runs the density
    pass after the rebuild.  A 12-step call against 12 single-step calls
    (skin 0.01 dx: the oracle rebuilds at steps 4, 6 and 9 of this case):
    the same rebuilds and fields (FP64 1e-13, mixed 1e-6), a fused pass in
    every step but the last, and a separate density pass in the first step
    and after each rebuild."""
    def case():
        sc = _case(prec, stale=False)
        sc.cfg.rebuild_policy = _abi.WF_REBUILD_PLASTIC_ZONE
        sc.cfg.ns_skin = 0.01 * 2 * sc.meta["r"]
        sc.cfg.ep_hot = 1e-6
        return sc
    print(f"[case] fusion under the plastic-zone trigger prec={prec}", flush=True)
    a, fused, dens, sa = _run(case(), wflib, single=False)
    b, fused1, dens1, sb = _run(case(), wflib, single=True)
    rebuilds = sa["ns_builds"] - 1  # the initial build aside
    print(f"  rebuilds {rebuilds}, fused {fused}, density passes {dens}")
    assert sa["ns_builds"] == sb["ns_builds"] and rebuilds >= 2
    assert fused1 == 0 and dens1 == NSTEPS
    assert fused == NSTEPS - 1 and dens == 1 + rebuilds
    tol = 1e-13 if prec == _abi.WF_FP64 else 1e-6
    for k in FIELDS:
        err = rel_l2(a[k], b[k])
        print(f"  {k}: {err:.2e}")
        assert err <= tol, k
