"""GPU parity of the laws the benchmark workloads do not reach, against the
FP64 oracle on the same seeded inputs:
  - Hollomon hardening (Material.cuh:192-201) under the reference return map
    (Mechanical.cu:504-535) and the consistent return (extension);
  - the convection boundary condition (CalcConvHeatKernel, Thermal.cu:36-64);
  - Monaghan viscosity with beta != 0 (InteractionAlt.cu:93-104).

Tolerances, as in test_gpu_parity / test_gpu_mixed:
  - FP64, one kernel call: relative L2 <= 1e-12 per field;
  - FP64, tens of steps: relative L2 <= 1e-9 per field;
  - mixed precision, one pair-kernel call: relative L2 <= 2e-5;
  - mixed precision, tens of steps: relative L2 <= 1e-4.
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import randomize, rel_l2, small_c1, to_gpu, to_oracle

pytestmark = pytest.mark.gpu

KERNEL_TOL = 1e-12
STEP_TOL = 1e-9
MIXED_KERNEL_TOL = 2e-5
MIXED_RUN_TOL = 1e-4

# steel-like Hollomon law with a short plateau: eps0 = sy0/E = 1.5e-3,
# eps1 = (sy0/K)^(1/m) = 0.6^10 = 6.0e-3, so eps_p in [0, 0.05] reaches both
# the plateau and the power branch
HOLLOMON = dict(model=_abi.WF_MAT_HOLLOMON, sy0=300e6, K_h=500e6, m_h=0.1)


def _hollomon_case(seed, rm, precision=_abi.WF_FP64):
    sc = randomize(small_c1(0.4), seed=seed, v_scale=0.5, s_scale=300e6)
    for k, v in HOLLOMON.items():
        setattr(sc.cfg.mat, k, v)
    sc.cfg.return_map = rm
    sc.cfg.precision = precision
    rng = np.random.default_rng(seed + 100)
    sc.fields["pl_strain"] = rng.uniform(0.0, 0.05, sc.n)
    sc.fields["sigma_y"] = rng.uniform(250e6, 400e6, sc.n)
    return sc


@pytest.mark.parametrize("rm", [_abi.WF_RETURN_REFERENCE, _abi.WF_RETURN_NEWTON])
def test_hollomon_stress_kernel(oracle, wflib, rm):
    sc = _hollomon_case(31, rm)
    print(f"[case] hollomon stress rm={rm} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    cfg = sc.cfg
    st.nsearch(cfg)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        st.calc_dens_inc(cfg)
        d.CalcDensInc()
        st.update_density(cfg.dt)
        d.UpdateDensity(cfg.dt)
        st.calc_rate_tensors(cfg)
        d.CalcRateTensors()
        st.pressure()
        d.Pressure()
        st.stress_strain(cfg)
        d.StressStrain()
        for k in ["sigma", "shearstress", "sigma_y", "pl_strain", "sigma_eq"]:
            assert rel_l2(d.get(k), st[k]) < KERNEL_TOL, k
    # both branches of the law are reached
    ep0, ep1 = sc.fields["pl_strain"], st["pl_strain"]
    yielded = ep1 > ep0
    knee = 0.6 ** 10 - HOLLOMON["sy0"] / 200e9
    assert yielded.sum() > 50
    assert (ep1[yielded] < knee).any() and (ep1[yielded] > knee).any()


@pytest.mark.parametrize("rm", [_abi.WF_RETURN_REFERENCE, _abi.WF_RETURN_NEWTON])
def test_hollomon_steps(oracle, wflib, rm):
    sc = _hollomon_case(32, rm)
    print(f"[case] hollomon steps rm={rm} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    nsteps = 25
    st.step(sc.cfg, sc.bcs, nsteps)
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        for k in ["u", "v", "sigma", "sigma_y", "sigma_eq", "pl_strain"]:
            assert rel_l2(d.get(k), st[k]) < STEP_TOL, k


def test_hollomon_steps_mixed(oracle, wflib):
    sc = _hollomon_case(33, _abi.WF_RETURN_NEWTON, _abi.WF_MIXED)
    print(f"[case] hollomon steps mixed n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    nsteps = 25
    st.step(sc.cfg, sc.bcs, nsteps)
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        for k in ["u", "sigma", "sigma_eq", "pl_strain"]:
            e = rel_l2(d.get(k), st[k])
            print(f"  {k}: rel L2 {e:.2e}")
            assert e < MIXED_RUN_TOL, k


def _convection_cube(precision):
    """14 mm cube, dx 1 mm: the outer layer convects to T_inf = 250 K with
    h = 1e4 W/m^2K (a time constant of 0.24 s against 40 steps of 5.2 ms)."""
    sc = scenarios.c2_thermal_cube(L=14e-3)
    bct = sc.fields["BC_T"]
    bct[bct == _abi.WF_BCT_FIXED] = _abi.WF_BCT_CONVECTION
    sc.cfg.T_inf = 250.0
    sc.cfg.h_conv = 1e4
    sc.cfg.precision = precision
    return sc


def _no_convection(oracle, sc, cfg, nsteps=None):
    """The oracle's result with the convecting faces made adiabatic."""
    st = to_oracle(oracle, sc)
    bct = sc.fields["BC_T"].copy()
    st["BC_T"] = np.where(bct == _abi.WF_BCT_CONVECTION, _abi.WF_BCT_NONE, bct)
    st.nsearch(cfg)
    if nsteps is None:
        st.calc_temp_inc(cfg)
        return st["dTdt"].copy()
    st.step(cfg, [], nsteps)
    return st["T"].copy()


@pytest.mark.parametrize("precision", [_abi.WF_FP64, _abi.WF_MIXED])
def test_convection_single_call(oracle, wflib, precision):
    sc = _convection_cube(precision)
    print(f"[case] convection single call prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.calc_temp_inc(sc.cfg)
    ref = st["dTdt"].copy()
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        d.CalcTempInc()
        g = d.get("dTdt")
    conv = sc.fields["BC_T"] == _abi.WF_BCT_CONVECTION
    tol = KERNEL_TOL if precision == _abi.WF_FP64 else MIXED_KERNEL_TOL
    assert rel_l2(g, ref) < tol
    assert rel_l2(g[conv], ref[conv]) < tol
    # the convective term is a sizeable part of the faces' rate: h A (T_inf - T)/(m cp)
    adiabatic = _no_convection(oracle, sc, sc.cfg)
    assert np.all(ref[conv] < adiabatic[conv])
    assert rel_l2(ref[conv], adiabatic[conv]) > 0.01


@pytest.mark.parametrize("precision", [_abi.WF_FP64, _abi.WF_MIXED])
def test_convection_steps(oracle, wflib, precision):
    sc = _convection_cube(precision)
    print(f"[case] convection steps prec={precision} n={sc.n}", flush=True)
    cfg64 = _convection_cube(_abi.WF_FP64).cfg
    st = to_oracle(oracle, sc)
    st.nsearch(cfg64)
    nsteps = 40
    st.step(cfg64, [], nsteps)
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        T = d.get("T")
    conv = sc.fields["BC_T"] == _abi.WF_BCT_CONVECTION
    adiabatic = _no_convection(oracle, sc, cfg64, nsteps)
    assert np.all(st["T"][conv] < adiabatic[conv])  # the faces lose heat to T_inf
    tol = STEP_TOL if precision == _abi.WF_FP64 else MIXED_RUN_TOL
    e = rel_l2(T - 300.0, st["T"] - 300.0)
    print(f"  T-300: rel L2 {e:.2e}")
    assert e < tol


def _beta_case(precision, seed=41, v_scale=2.0, s_scale=100e6, beta=1.5):
    sc = randomize(small_c1(0.4), seed=seed, v_scale=v_scale, s_scale=s_scale)
    sc.cfg.alpha = 1.0
    sc.cfg.beta = beta
    sc.cfg.precision = precision
    return sc


@pytest.mark.parametrize("precision", [_abi.WF_FP64, _abi.WF_MIXED])
def test_viscosity_beta_accel(oracle, wflib, precision):
    """Fast approach (|v| up to 10 m/s) under small stresses, where the
    beta mu^2 term moves a by ~4e-3.  Besides a itself, the beta part alone
    (a with beta = 1.5 minus a with beta = 0, both runs on each side) is
    compared: 1e-9 in FP64, 1e-3 in mixed precision (a difference of two
    FP32-term sums that agree to ~99.6 %)."""
    out = {}
    for beta in (1.5, 0.0):
        sc = _beta_case(precision, v_scale=10.0, s_scale=1e6, beta=beta)
        print(f"[case] beta accel prec={precision} beta={beta} n={sc.n}", flush=True)
        st = to_oracle(oracle, sc)
        st.nsearch(sc.cfg)
        st.calc_accel(sc.cfg)
        with to_gpu(sc, wflib) as d:
            d.nsearch()
            d.CalcAccel()
            out[beta] = (d.get("a"), st["a"].copy())
    tol = KERNEL_TOL if precision == _abi.WF_FP64 else MIXED_KERNEL_TOL
    g, o = out[1.5]
    assert rel_l2(g, o) < tol
    o_delta = o - out[0.0][1]
    g_delta = g - out[0.0][0]
    assert rel_l2(o_delta, np.zeros_like(o)) > 1e-3 * rel_l2(o, np.zeros_like(o))
    assert rel_l2(g_delta, o_delta) < (1e-9 if precision == _abi.WF_FP64 else 1e-3)


@pytest.mark.parametrize("precision", [_abi.WF_FP64, _abi.WF_MIXED])
def test_viscosity_beta_steps(oracle, wflib, precision):
    sc = _beta_case(precision, seed=42)
    print(f"[case] beta steps prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    cfg64 = _beta_case(_abi.WF_FP64, seed=42).cfg
    st.nsearch(cfg64)
    nsteps = 25
    st.step(cfg64, sc.bcs, nsteps)
    tol = STEP_TOL if precision == _abi.WF_FP64 else MIXED_RUN_TOL
    with to_gpu(sc, wflib) as d:
        d.step(nsteps)
        for k in ["u", "v", "sigma", "pl_strain"]:
            e = rel_l2(d.get(k), st[k])
            print(f"  {k}: rel L2 {e:.2e}")
            assert e < tol, k
