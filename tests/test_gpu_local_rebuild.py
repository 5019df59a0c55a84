"""The local re-list policy (WF_REBUILD_PLASTIC_ZONE_LOCAL, an extension) on
the GPU: the plastic-zone trigger, but a rebuild re-lists in place only the
tiles within two search cells of a particle with eps_p > ep_hot, with no
reorder, and a full rebuild every ns_global_every re-lists.

Pairs between 2h and 2h + skin contribute exactly nothing, so lists that
hold every pair within 2h give the physics of any other such lists; the
runs are checked against the oracle under the global plastic-zone policy
(FP64 relative L2 <= 1e-9, mixed <= 1e-4, the step tolerances), and the
GPU lists at the end of a run against a brute-force 2h search of the
current positions (every pair within 2h listed).  The friction-stir case is
the small plate of test_gpu_friction_stir with the tool turning fast enough
that the trigger fires every few steps."""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import rel_l2, to_gpu, to_oracle

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "sigma", "pl_strain", "T"]


def _case(global_every=32):
    sc = scenarios.c5_friction_stir(scale=0.15, r=0.15e-3, n_theta=32, n_rs=4, n_zp=4, n_rc=2,
                                    thick=3.0, mu=0.01, offset=(0.137, 0.291))
    sc.meshes[0]["omega"] = (0.0, 0.0, 1e5)
    dx = sc.meta["dx"]
    sc.cfg.ns_skin = 0.02 * dx
    sc.cfg.ep_hot = 1e-4
    sc.cfg.ns_global_every = global_every
    return sc


def _lists_hold_every_pair(oracle, d, h):
    x = d.get("x")
    cnt, off, lst, _ = oracle.neighbors(x, 2 * h)
    gc, go, gl = d.neighbors()
    missing = 0
    for i in range(len(cnt)):
        want = set(lst[off[i]:off[i] + cnt[i]].tolist())
        have = set(gl[go[i]:go[i] + gc[i]].tolist())
        missing += len(want - have)
    return missing, int(cnt.sum()), int(gc.sum())


@pytest.mark.parametrize("precision,tol,every", [(_abi.WF_FP64, 1e-9, 32),
                                                 (_abi.WF_MIXED, 1e-4, 32),
                                                 (_abi.WF_FP64, 1e-9, 3)])
def test_local_relists_match_oracle(oracle, wflib, precision, tol, every):
    sc = _case(every)
    nsteps = 40
    print(f"[case] local re-lists prec={precision} every={every} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.step(sc.cfg, sc.bcs, nsteps)
    assert st.run.ns_builds >= 3  # the oracle's global policy rebuilt in the run
    sc.cfg.precision = precision
    sc.cfg.rebuild_policy = _abi.WF_REBUILD_PLASTIC_ZONE_LOCAL
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        s = d.stats()
        got = {k: d.get(k) for k in FIELDS}
        missing, exact, held = _lists_hold_every_pair(oracle, d, sc.cfg.h_glob)
    nt = (sc.n + 127) // 128
    print(f"  re-lists {s['ns_local_builds']}, tiles re-listed {s['ns_local_tiles']} "
          f"(of {nt} per list), full builds {s['ns_builds']}, oracle builds "
          f"{st.run.ns_builds}; pairs within 2h {exact}, listed {held}", flush=True)
    assert s["ns_local_builds"] >= 2
    # a re-list touches only part of the plate
    assert s["ns_local_tiles"] < 0.8 * nt * s["ns_local_builds"]
    if every == 3:
        assert s["ns_builds"] >= 2  # the full rebuild after every third re-list
    assert missing == 0
    for k in FIELDS:
        ref = st[k] - 300.0 if k == "T" else st[k]
        val = got[k] - 300.0 if k == "T" else got[k]
        err = rel_l2(val, ref)
        print(f"  {k}: {err:.2e}")
        assert err < tol, k
