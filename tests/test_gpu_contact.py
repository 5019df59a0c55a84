"""Rigid-mesh contact on the GPU against the FP64 oracle: mesh motion and
geometry, free-surface detection, contact candidate lists, penalty +
friction forces (both laws), the sliding-friction heat (an extension), and
whole steps of a plate compressing a Johnson-Cook cylinder (FP64 and mixed
precision) with that heat feeding the conduction pass.

Tolerances: mesh geometry 1e-14 absolute on O(1e-2) coordinates; surface
IDs and candidate lists bit-exact (in both precisions; the mixed surface pass
sums FP32 terms, its normals within 1e-6); single-call accelerations and contact
forces relative L2 <= 1e-12; FP64 steps relative L2 <= 1e-9 (the pair-loop
step tolerance); mixed-precision steps <= 1e-4 against the FP64 oracle.
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import rel_l2, to_gpu, to_oracle

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-9


def _case(name, sc):
    print(f"[case] {name}: n={sc.n} triangles={sum(len(m['elnode']) for m in sc.meshes)}",
          flush=True)


def _plate_case(law=_abi.WF_CONTACT_REFERENCE, mu=0.3, dens=6, sink=0.4, seed=None,
                scale=0.2, r=0.3e-3, thermal=True, v_scale=0.3):
    """Small JC cylinder with the plate already `sink` h into the top layer
    (offset in x, y so no particle sits on a triangle edge)."""
    sc = scenarios.c4_contact(scale=scale, r=r, dens=dens, law=law, mu=mu, offset=0.137,
                              thermal=thermal)
    sc.meshes[0]["nodes"][:, 2] -= sink * sc.meta["h"]
    if seed is not None:
        rng = np.random.default_rng(seed)
        sc.fields["v"] = rng.uniform(-v_scale, v_scale, (sc.n, 3))
    return sc


def test_mesh_motion_and_geometry(oracle, wflib):
    sc = _plate_case()
    m = sc.meshes[0]
    m["omega"] = (0.0, 0.0, 40.0)
    m["center"] = (1e-3, -2e-3, 0.0)
    _case("mesh-motion", sc)
    st = to_oracle(oracle, sc)
    dt = sc.cfg.dt
    with to_gpu(sc, wflib) as d:
        g0 = d.get_mesh(0)
        assert np.array_equal(g0["normals"], st.mesh_arrays["nrm"])
        assert np.array_equal(g0["pplane"], st.mesh_arrays["pp"])
        for _ in range(10):
            d.MeshUpdate(dt)
            st.mesh_move(dt)
        g = d.get_mesh(0)
    assert np.max(np.abs(g["nodes"] - st.mesh_arrays["x"])) < 1e-14
    assert np.max(np.abs(g["node_v"] - st.mesh_arrays["v"])) < 1e-14
    assert np.max(np.abs(g["centroids"] - st.mesh_arrays["c"])) < 1e-14
    assert np.max(np.abs(g["normals"] - st.mesh_arrays["nrm"])) < 1e-14
    assert np.max(np.abs(g["pplane"] - st.mesh_arrays["pp"])) < 1e-14


def test_surface_ids_and_contact_lists(oracle, wflib):
    sc = _plate_case()
    _case("surface+lists", sc)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_surface(sc.cfg)
    st.contact_nb(sc.cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalculateSurface()
        ID, nrm = d.get("ID"), d.get("normal")
        d.CalcContactNb()
        cnt, lst = d.contact_neighbors()
        s = d.stats()
    assert np.array_equal(ID, st["ID"])
    assert (ID == 1).sum() > 0 and np.all(ID[sc.fields["ID"] == 2] == 2)
    assert rel_l2(nrm, st.contact_arrays["normal"]) < 1e-12
    assert np.array_equal(cnt, st.contact_arrays["cnb_count"])
    assert np.array_equal(lst, st.contact_arrays["cnb_list"])
    assert cnt.max() > 0 and s["cont_over_cap"] == st.ct.cont_over_cap == 0


def test_surface_ids_mixed_precision(oracle, wflib):
    sc = _plate_case()
    sc.cfg.precision = _abi.WF_MIXED
    _case("surface mixed", sc)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_surface(sc.cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalculateSurface()
        ID, nrm = d.get("ID"), d.get("normal")
    assert np.array_equal(ID, st["ID"])
    assert (ID == 1).sum() > 0
    assert rel_l2(nrm, st.contact_arrays["normal"]) < 1e-6


def test_contact_list_overflow_keeps_smallest_indices(oracle, wflib):
    sc = _plate_case(dens=24, scale=0.15)
    _case("overflow", sc)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_surface(sc.cfg)
    st.contact_nb(sc.cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalculateSurface()
        d.CalcContactNb()
        cnt, lst = d.contact_neighbors()
        s = d.stats()
    assert st.ct.cont_over_cap > 0 and s["cont_over_cap"] == st.ct.cont_over_cap
    assert np.array_equal(cnt, st.contact_arrays["cnb_count"])
    assert np.array_equal(lst, st.contact_arrays["cnb_list"])


@pytest.mark.parametrize("law,mu", [(_abi.WF_CONTACT_REFERENCE, 0.0),
                                    (_abi.WF_CONTACT_REFERENCE, 0.3),
                                    (_abi.WF_CONTACT_STIFFNESS, 0.2)])
def test_contact_forces_single_call(oracle, wflib, law, mu):
    # slip over one step must exceed mu x penetration for some particles to
    # slide: tangential speeds of O(1e3) m/s mix both friction branches
    sc = _plate_case(law=law, mu=mu, seed=3, v_scale=3000.0)
    sc.cfg.fric_heat = 0.8
    rng = np.random.default_rng(4)
    sc.fields["a"] = rng.uniform(-1e5, 1e5, (sc.n, 3))
    _case(f"forces law={law} mu={mu}", sc)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_surface(sc.cfg)
    st.contact_nb(sc.cfg)
    st.mesh_move(sc.cfg.dt)       # element velocities for the friction term
    st.contact_forces(sc.cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalculateSurface()
        d.MeshUpdate(sc.cfg.dt)
        d.CalcContactNb()
        d.CalcContactForces()
        a, cf = d.get("a"), d.get("contforce")
        qf = d.get("q_friction")
        fsum = d.contact_force_sum()
    ref_cf = st.contact_arrays["cf"]
    if mu > 0.0:  # sliding friction heat, an extension
        assert np.count_nonzero(st["q_friction"]) > 0
        assert rel_l2(qf, st["q_friction"]) < 1e-12
    else:
        assert not np.any(qf)
    assert np.count_nonzero(np.linalg.norm(ref_cf, axis=1)) > 10
    assert rel_l2(a, st["a"]) < 1e-12
    assert rel_l2(cf, ref_cf) < 1e-12
    exp_sum = np.linalg.norm(ref_cf, axis=1)[st["ID"] == 1].sum()
    assert fsum == pytest.approx(exp_sum, rel=1e-12)


@pytest.mark.parametrize("law", [_abi.WF_CONTACT_REFERENCE, _abi.WF_CONTACT_STIFFNESS])
def test_steps_plate_compression(oracle, wflib, law):
    """The plate starts one h above the top layer and closes in at 2 m/s; the
    fields after each block of steps, the contact force where the plate
    touches and the sliding-friction heat match the oracle (with the
    reference law's softer static threshold, 0.4 mu delta, surface particles
    slide in these steps; under the stiffness law they stick)."""
    sc = _plate_case(law=law, mu=0.2, sink=0.0, seed=5)
    sc.cfg.fric_heat = 0.9
    _case(f"steps law={law}", sc)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    touched = heated = 0
    with to_gpu(sc, wflib) as d:
        for _ in range(6):
            st.step(sc.cfg, sc.bcs, 5)
            d.step(5)
            cf = st.contact_arrays["cf"]
            touched += np.count_nonzero(np.linalg.norm(cf, axis=1))
            heated += np.count_nonzero(st["q_friction"])
            assert rel_l2(d.get("contforce"), cf) < STEP_TOL
            for k in ["u", "v", "sigma", "pl_strain", "T", "q_friction"]:
                assert rel_l2(d.get(k), st[k]) < STEP_TOL, k
        assert np.array_equal(d.get("ID"), st["ID"])
        nodes = d.get_mesh(0)["nodes"]
    assert np.max(np.abs(nodes - st.mesh_arrays["x"])) < 1e-14
    print(f"  contacts {touched}, sliding-heated {heated}")
    assert touched > 0
    assert heated > 0 or law == _abi.WF_CONTACT_STIFFNESS


def test_steps_plate_compression_mixed(oracle, wflib):
    sc = _plate_case(law=_abi.WF_CONTACT_STIFFNESS, mu=0.2, sink=0.0)
    _case("steps mixed", sc)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    nsteps = 30
    st.step(sc.cfg, sc.bcs, nsteps)
    sc.cfg.precision = _abi.WF_MIXED
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        u, sig, cf = d.get("u"), d.get("sigma"), d.get("contforce")
    for name, a, b in [("u", u, st["u"]), ("sigma", sig, st["sigma"]),
                       ("contforce", cf, st.contact_arrays["cf"])]:
        err = rel_l2(a, b)
        print(f"  {name}: rel L2 {err:.2e}")
        assert err < 1e-4, name
