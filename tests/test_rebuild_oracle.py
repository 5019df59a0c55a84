"""The plastic-zone rebuild policy (WF_REBUILD_PLASTIC_ZONE, an extension)
in the oracle, on the CPU.

Lists are built from the current positions with radius 2h + skin and the
build positions are kept; once yielding, a rebuild runs when a particle with
eps_p > ep_hot has moved more than skin/2 since the build.  Pinned by brute
force (the lists), the trigger's definition on hand-placed displacements,
and an exactness property: pairs beyond 2h contribute exactly zero to every
pair sum, so an elastic run in which no pair crosses 2h gives bit-identical
fields with 2h lists and with 2h + skin lists.
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import randomize, to_oracle


def _pz(cfg, skin, ep_hot=0.05):
    cfg.rebuild_policy = _abi.WF_REBUILD_PLASTIC_ZONE
    cfg.ns_skin = skin
    cfg.ep_hot = ep_hot
    return cfg


def test_lists_have_the_skin_radius_and_keep_build_positions(oracle):
    sc = scenarios.c1_bilinear(scale=0.3)
    dx = 0.5e-3
    _pz(sc.cfg, 0.3 * dx)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    cnt, off, lst, _ = oracle.neighbors(sc.fields["x"], 2 * sc.cfg.h_glob + 0.3 * dx, brute=True)
    assert np.array_equal(st["nb_count"], cnt)
    oc, oo, ol = st.neighbor_rows()
    assert np.array_equal(oo, off) and np.array_equal(ol, lst)
    assert cnt.max() > 56  # the skin reaches the next lattice shell
    assert np.array_equal(st["xb"], sc.fields["x"])


def test_trigger_counts_only_hot_particles_beyond_half_the_skin(oracle):
    sc = scenarios.c1_bilinear(scale=0.3)
    skin = 1e-5
    _pz(sc.cfg, skin, ep_hot=0.05)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    x0 = st["x"].copy()
    st["x"] = x0 + np.array([0.6 * skin, 0.0, 0.0])  # every particle moves 0.6 skin
    assert not st.plastic_zone_moved(sc.cfg)  # nobody is hot
    ep = np.zeros(sc.n)
    ep[7] = 0.051
    st["pl_strain"] = ep
    assert st.plastic_zone_moved(sc.cfg)  # one hot particle moved > skin/2
    st["x"] = x0
    st["x"][7, 2] += 0.49 * skin
    st["x"][8, 2] += 3.0 * skin  # cold: ignored
    assert not st.plastic_zone_moved(sc.cfg)
    st["x"][7, 2] += 0.02 * skin
    assert st.plastic_zone_moved(sc.cfg)


def test_skin_pairs_contribute_exactly_zero(oracle):
    """Elastic C1 (no yielding, so no rebuild under either policy), random
    velocities small enough that no pair crosses 2h in 10 steps: the fields
    are bit-identical with 2h lists and with 2h + skin lists."""
    base = randomize(scenarios.c1_bilinear(scale=0.3), seed=4, v_scale=0.05, s_scale=20e6)
    base.cfg.quirks &= ~_abi.WF_Q_STALE_NS_POS
    skin = scenarios.c1_bilinear(scale=0.3)
    skin.fields = {k: np.copy(v) for k, v in base.fields.items()}
    skin.cfg = base.cfg.__class__.from_buffer_copy(base.cfg)
    _pz(skin.cfg, 0.2 * 0.5e-3)
    a, b = to_oracle(oracle, base), to_oracle(oracle, skin)
    a.nsearch(base.cfg)
    b.nsearch(skin.cfg)
    assert b["nb_count"].sum() > a["nb_count"].sum()
    a.step(base.cfg, base.bcs, 10)
    b.step(skin.cfg, skin.bcs, 10)
    assert not a.run.is_yielding and not b.run.is_yielding
    for k in ["x", "v", "a", "rho", "sigma", "shearstress", "strrate", "rotrate"]:
        assert np.array_equal(a[k], b[k]), k
