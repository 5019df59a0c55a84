"""Particles of unequal smoothing length and mass.  The generated lattices
give every particle the same h and m, and the library then takes loops that
read h_i and m once per block (the uniform-h pair loops and the uniform-mass
surface pass); these tests perturb h and m per particle by a few percent so
the general loops run: each pair uses h_ij = (h_i + h_j) / 2 and the
neighbour's own mass, as InteractionAlt.cu:18-383 and Thermal.cu:101-142 do,
and the surface normal sums m_j x_ij (Geometry.cu:26-88).

The neighbour search radius stays 2 h_glob on both sides, so the pair sets
are the same as in the uniform case.

Tolerances: FP64 one kernel call relative L2 <= 1e-12 and tens of steps
<= 1e-9; mixed precision one call <= 2e-5 and coupled runs <= 1e-4; surface
IDs bit-exact and surface normals <= 1e-12 (FP64) or 1e-6 (mixed).
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import randomize, rel_l2, small_c1, to_gpu, to_oracle

pytestmark = pytest.mark.gpu


def _perturb(sc, seed, dh=0.03, dm=0.05):
    rng = np.random.default_rng(seed)
    n = sc.n
    sc.fields["h"] = sc.fields["h"] * rng.uniform(1.0 - dh, 1.0 + dh, n)
    sc.fields["m"] = sc.fields["m"] * rng.uniform(1.0 - dm, 1.0 + dm, n)
    return sc


@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-12), (_abi.WF_MIXED, 2e-5)])
def test_pair_kernels_nonuniform(oracle, wflib, precision, tol):
    sc = _perturb(randomize(small_c1(0.4), seed=31), seed=32)
    sc.cfg.precision = precision
    print(f"[case] nonuniform kernels prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    cfg = sc.cfg
    st.nsearch(cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        st.calc_dens_inc(cfg)
        d.CalcDensInc()
        assert rel_l2(d.get("drho"), st["drho"]) < tol
        st.calc_rate_tensors(cfg)
        d.CalcRateTensors()
        assert rel_l2(d.get("strrate"), st["strrate"]) < tol
        assert rel_l2(d.get("rotrate"), st["rotrate"]) < tol
        st.calc_accel(cfg)
        d.CalcAccel()
        assert rel_l2(d.get("a"), st["a"]) < tol


@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-12), (_abi.WF_MIXED, 2e-5)])
def test_conduction_nonuniform(oracle, wflib, precision, tol):
    sc = _perturb(scenarios.c2_thermal_cube(L=10e-3), seed=33)
    sc.cfg.precision = precision
    print(f"[case] nonuniform conduction prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_temp_inc(sc.cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalcTempInc()
        g = d.get("dTdt")
    assert np.abs(st["dTdt"]).max() > 0
    assert rel_l2(g, st["dTdt"]) < tol


@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-9), (_abi.WF_MIXED, 1e-4)])
def test_steps_nonuniform(oracle, wflib, precision, tol):
    sc = _perturb(randomize(small_c1(0.4), seed=34, v_scale=0.5, s_scale=250e6), seed=35)
    sc.cfg.precision = precision
    print(f"[case] nonuniform steps prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    nsteps = 25
    st.step(sc.cfg, sc.bcs, nsteps)
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        for k in ["u", "v", "rho", "sigma", "pl_strain"]:
            e = rel_l2(d.get(k), st[k])
            print(f"  {k}: rel L2 {e:.2e}")
            assert e < tol, k


@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-12), (_abi.WF_MIXED, 1e-6)])
def test_surface_ids_nonuniform_mass(oracle, wflib, precision, tol):
    sc = scenarios.c4_contact(scale=0.2, r=0.3e-3, dens=6, offset=0.137)
    sc.cfg.precision = precision
    sc.meshes[0]["nodes"][:, 2] -= 0.4 * sc.meta["h"]
    rng = np.random.default_rng(36)
    sc.fields["m"] = sc.fields["m"] * rng.uniform(0.9, 1.1, sc.n)
    print(f"[case] surface nonuniform mass n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_surface(sc.cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalculateSurface()
        ID, nrm = d.get("ID"), d.get("normal")
    assert np.array_equal(ID, st["ID"])
    assert (ID == 1).sum() > 0
    assert rel_l2(nrm, st.contact_arrays["normal"]) < tol
