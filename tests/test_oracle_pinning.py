"""Pin the FP64 oracle against facts that do not come from the oracle itself:
textbook closed forms (cubic spline, Johnson-Cook, radial return), lattice
sums, conservation laws and brute force.  The reference ships no golden
vectors, so these checks are what makes the oracle trustworthy; each is
chosen so that a dropped term, a wrong sign or a transposed index fails.
"""
import math

import numpy as np
import pytest
from scipy import integrate, optimize

from weldformgpu_amd import _abi, scenarios

from harness import randomize, small_c1, to_oracle


# ---------------------------------------------------------------------------
# cubic spline (Monaghan & Lattanzio 1985), 3D normalisation 1/(pi h^3)
def w_textbook(q, h):
    s = 1.0 / (math.pi * h ** 3)
    if q < 1:
        return s * (1 - 1.5 * q * q + 0.75 * q ** 3)
    if q < 2:
        return s * 0.25 * (2 - q) ** 3
    return 0.0


@pytest.mark.parametrize("q", [0.0, 0.3, 0.5, 0.999, 1.0, 1.5, 1.99, 2.0, 2.5])
def test_kernel_closed_form(oracle, q):
    h = 0.6e-3
    assert oracle.kernel(q, h) == pytest.approx(w_textbook(q, h), rel=1e-14, abs=0.0)


def test_kernel_normalised(oracle):
    h = 0.6e-3
    val, _ = integrate.quad(lambda r: 4 * math.pi * r * r * oracle.kernel(r / h, h), 0, 2 * h,
                            points=[h], epsabs=1e-14, epsrel=1e-12)
    assert val == pytest.approx(1.0, rel=1e-10)


@pytest.mark.parametrize("q", [0.2, 0.7, 1.2, 1.7, 1.95])
def test_grad_kernel_is_derivative_over_r(oracle, q):
    h = 0.6e-3
    r = q * h
    e = 1e-6 * h
    dwdr = (w_textbook((r + e) / h, h) - w_textbook((r - e) / h, h)) / (2 * e)
    assert oracle.grad_kernel(q, h) == pytest.approx(dwdr / r, rel=1e-7)


# ---------------------------------------------------------------------------
# geometry
def test_box_lattice_count_and_mass(oracle):
    xyz, mass = oracle.gen_box([0, 0, 0], [46e-3] * 3, 0.5e-3, 2700.0)
    assert len(xyz) == 46 ** 3
    assert mass * len(xyz) == pytest.approx(2700.0 * 46e-3 ** 3, rel=1e-12)
    assert np.allclose(np.unique(np.round(xyz[:, 0] / 1e-3, 9)), np.arange(46) + 0.5)


def test_cylinder_lattice_structure(oracle):
    """C1 deck: every particle inside R, 2r spacing, mirror symmetric, equal
    layers; the reference's accumulated zp += 2r loses the 30th layer to
    rounding (0.01475000000000001 > 0.01475), giving 29 x 648 = 18,792."""
    R, H, r = 7.5e-3, 15e-3, 0.25e-3
    xyz, mass = oracle.gen_cylinder([0, 0, 0], R, H, r, 7850.0)
    assert len(xyz) == 18792
    zs = np.unique(xyz[:, 2])
    assert len(zs) == 29
    layer = xyz[xyz[:, 2] == zs[0]]
    assert len(layer) == 648
    assert np.all(np.hypot(layer[:, 0], layer[:, 1]) <= R)
    key = {(round(a / r), round(b / r)) for a, b in layer[:, :2]}
    assert key == {(-a, b) for a, b in key} == {(a, -b) for a, b in key}
    assert all(a % 2 and b % 2 for a, b in key)  # odd multiples of r: 2r grid
    assert mass * len(xyz) == pytest.approx(7850 * math.pi * R * R * H, rel=1e-12)


def test_product_lattices_match_oracle(oracle):
    for R, H, r in [(7.5e-3, 15e-3, 0.25e-3), (3e-3, 6e-3, 0.15e-3)]:
        a, ma = oracle.gen_cylinder([0, 0, 0], R, H, r, 2700.0)
        b, mb = scenarios.gen_cylinder([0, 0, 0], R, H, r, 2700.0)
        assert np.array_equal(a, b) and ma == mb
    a, ma = oracle.gen_box([1e-3, -2e-3, 0], [10e-3, 6e-3, 4e-3], 0.5e-3, 7850.0)
    b, mb = scenarios.gen_box([1e-3, -2e-3, 0], [10e-3, 6e-3, 4e-3], 0.5e-3, 7850.0)
    assert np.array_equal(a, b) and ma == mb


# ---------------------------------------------------------------------------
# neighbour search
def test_neighbours_cell_list_equals_brute_force(oracle):
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 5e-3, (2500, 3))
    rad = 0.9e-3
    a = oracle.neighbors(pts, rad)
    b = oracle.neighbors(pts, rad, brute=True)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_neighbours_float_criterion_at_cutoff(oracle):
    """Pairs straddling the radius are decided by the float test
    0 < d2 < r*r (cuNSearchKernels.cu:163-164), not by the double one."""
    rf = np.float32(1e-3)
    r2 = rf * rf
    pts = []
    for k, d in enumerate(np.linspace(0.999999e-3, 1.000001e-3, 41)):
        pts += [[0.0, 0.0, 0.1 * k], [d, 0.0, 0.1 * k]]
    pts = np.array(pts)
    c, o, l, _ = oracle.neighbors(pts, 1e-3)
    p32 = pts.astype(np.float32)
    for k in range(41):
        d = p32[2 * k + 1] - p32[2 * k]
        d2 = np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        assert (c[2 * k] == 1) == bool(d2 < r2)


def test_neighbours_lattice_shells(oracle):
    """h = 1.2 dx: the bulk count is the number of lattice points with
    0 < |k|^2 < 2.4^2, i.e. shells 1..5: 6 + 12 + 8 + 6 + 24 = 56; rows are
    symmetric and never contain the particle itself."""
    xyz, _ = oracle.gen_box([0, 0, 0], [12e-3] * 3, 0.5e-3, 1.0)
    c, o, l, over = oracle.neighbors(xyz, 2.4e-3)
    assert c.max() == 56 and over == 0
    inner = np.all((xyz > 2.5e-3) & (xyz < 9.5e-3), axis=1)
    assert np.all(c[inner] == 56)
    rows = {i: set(l[o[i]:o[i] + c[i]]) for i in range(len(c))}
    for i in range(0, len(c), 97):
        assert i not in rows[i]
        for j in rows[i]:
            assert i in rows[j]


# ---------------------------------------------------------------------------
# pair loops on a lattice: lattice-sum closed forms
def _lattice_state(oracle, v_fn, sigma=None, n_side=11, dx=1e-3, rho=1000.0):
    L = n_side * dx
    xyz, mass = oracle.gen_box([0, 0, 0], [L] * 3, dx / 2, rho)
    n = len(xyz)
    st = oracle.State(n)
    st["x"] = xyz
    st["xs"] = xyz
    st["v"] = v_fn(xyz)
    st["rho"] = rho
    st["rho0"] = rho
    st["m"] = mass
    st["h"] = 1.2 * dx
    st["Cs"] = 5000.0
    st["G"] = 1e9
    if sigma is not None:
        st["sigma"] = sigma
    cfg = _abi.default_config()
    cfg.n_particles = n
    cfg.h_glob = 1.2 * dx
    cfg.dt = 1e-7
    st.nsearch(cfg)
    c = L / 2
    inner = np.all(np.abs(xyz - c) < L / 2 - 2.6 * dx, axis=1)
    return st, cfg, xyz, mass, inner


def _shell_vectors(dx, h):
    ks = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)
                   if 0 < a * a + b * b + c * c < 5.76])
    r = np.linalg.norm(ks, axis=1) * dx
    def gk(rr):
        q = rr / h
        C = 1.0 / (math.pi * h ** 4)
        if q < 1:
            return C / (q * h) * (-3 * q + 2.25 * q * q)
        return C / (q * h) * (-0.75 * (2 - q) ** 2)
    return ks * dx, np.array([gk(x) for x in r])


def test_dens_inc_uniform_expansion_lattice_sum(oracle):
    edot = 3.0
    st, cfg, xyz, m, inner = _lattice_state(oracle, lambda x: edot * x)
    st.calc_dens_inc(cfg)
    vecs, gk = _shell_vectors(1e-3, 1.2e-3)
    expected = edot * m * np.sum(gk * (vecs ** 2).sum(1))
    assert np.allclose(st["drho"][inner], expected, rtol=1e-12)
    assert expected < 0  # expansion lowers density


def test_dens_inc_uniform_and_rigid_rotation(oracle):
    st, cfg, xyz, m, inner = _lattice_state(oracle, lambda x: np.tile([0.3, -1.0, 2.0], (len(x), 1)))
    st.calc_dens_inc(cfg)
    assert np.all(st["drho"] == 0.0)
    w = np.array([1.0, -2.0, 0.5])
    st2, cfg2, *_ = _lattice_state(oracle, lambda x: np.cross(w, x))
    st2.calc_dens_inc(cfg2)
    # each term m (rho_i/rho_j) GK (v_ij . x_ij) vanishes up to rounding of v_ij
    scale = m * np.abs(_shell_vectors(1e-3, 1.2e-3)[1]).sum() * 2.0 * 11e-3 * 1e-3
    assert np.abs(st2["drho"]).max() < 1e-13 * scale


@pytest.mark.parametrize("quirks", [_abi.WF_Q_ALL, 0])
def test_rate_tensors_linear_field(oracle, quirks):
    Lg = np.array([[0.3, 1.1, -0.7], [0.2, -0.5, 0.9], [1.3, -0.4, 0.8]])
    st, cfg, xyz, m, inner = _lattice_state(oracle, lambda x: x @ Lg.T)
    cfg.quirks = quirks
    st.calc_rate_tensors(cfg)
    vecs, gk = _shell_vectors(1e-3, 1.2e-3)
    c = np.sum(gk * vecs[:, 0] ** 2)  # sum GK x_x^2 = sum GK x_y^2 = sum GK x_z^2
    mr = m / 1000.0
    sym = 0.5 * (Lg + Lg.T)
    # eps_ab = -(m/rho) c sym(L)_ab
    exp_d = -mr * c * np.array([sym[0, 0], sym[1, 1], sym[2, 2], sym[0, 1], sym[1, 2], sym[0, 2]])
    got = st["strrate"][inner]
    assert np.allclose(got, exp_d, rtol=1e-10, atol=1e-12 * np.abs(exp_d).max())
    # W_ab = -1/2 (m/rho) c (L_ab - L_ba); the reference negates xz
    skew = Lg - Lg.T
    s_xz = -1.0 if quirks & _abi.WF_Q_ROT_XZ_SIGN else 1.0
    exp_w = -0.5 * mr * c * np.array([skew[0, 1], skew[1, 2], s_xz * skew[0, 2]])
    gw = st["rotrate"][inner]
    assert np.all(gw[:, :3] == 0.0)
    assert np.allclose(gw[:, 3:], exp_w, rtol=1e-10, atol=1e-12 * np.abs(exp_w).max())


def test_accel_uniform_stress_and_momentum(oracle):
    sig = np.array([1e8, -2e8, 3e7, 4e7, -5e7, 6e7])
    st, cfg, xyz, m, inner = _lattice_state(oracle, lambda x: np.zeros_like(x), sigma=sig)
    st.calc_accel(cfg)
    a = st["a"]
    scale = m * 2e8 / 1000.0 ** 2 * 1e3 * 2 / (math.pi * (1.2e-3) ** 4) * 1e-3
    assert np.abs(a[inner]).max() < 1e-10 * scale
    # random sigma and v, all particles free: sum_i m_i a_i = 0 (pair antisymmetry)
    rng = np.random.default_rng(2)
    sc = randomize(small_c1(0.4), seed=4)
    sc.fields["rho"] = 7850.0 * (1 + rng.uniform(-0.01, 0.01, sc.n))
    st2 = to_oracle(oracle, sc)
    st2.nsearch(sc.cfg)
    st2.calc_accel(sc.cfg)
    mom = (st2["m"][:, None] * st2["a"]).sum(0)
    mag = (st2["m"][:, None] * np.abs(st2["a"])).sum(0)
    assert np.all(np.abs(mom) < 1e-11 * mag)


def test_accel_viscosity_only_on_approach(oracle):
    sig = np.zeros(6)
    st, cfg, *_ = _lattice_state(oracle, lambda x: 2.0 * x, sigma=sig)  # receding pairs
    st.calc_accel(cfg)
    assert np.all(st["a"] == 0.0)
    st2, cfg2, _, _, inner = _lattice_state(oracle, lambda x: -2.0 * x, sigma=sig)  # approaching
    st2.calc_accel(cfg2)
    assert np.abs(st2["a"]).max() > 0.0


# ---------------------------------------------------------------------------
# stress update / return map on single particles (no neighbours needed)
def _one(oracle, S0, D, W6, model, sy, G=26.5e9, dt=1e-7, quirks=_abi.WF_Q_ALL,
         rm=_abi.WF_RETURN_REFERENCE, ep=0.0, T=300.0, mat=None):
    st = oracle.State(1)
    st["shearstress"] = S0
    st["strrate"] = D
    st["rotrate"] = W6
    st["G"] = G
    st["sigma_y"] = sy
    st["pl_strain"] = ep
    st["rho"] = 2700.0
    st["rho0"] = 2700.0
    st["m"] = 1.0
    st["T"] = T
    cfg = _abi.default_config()
    cfg.n_particles = 1
    cfg.dt = dt
    cfg.quirks = quirks
    cfg.return_map = rm
    cfg.mat.model = model
    cfg.mat.E = 68.9e9
    cfg.mat.Et = 1e9
    if mat:
        for k, v in mat.items():
            setattr(cfg.mat, k, v)
    st.stress_strain(cfg)
    return st, cfg


def _vm(s6):
    s = np.array([[s6[0], s6[3], s6[5]], [s6[3], s6[1], s6[4]], [s6[5], s6[4], s6[2]]])
    return math.sqrt(1.5 * np.sum(s * s))


def test_stress_elastic_predictor(oracle):
    D = np.array([1.0, -0.4, -0.2, 0.3, 0.1, -0.2])
    st, cfg = _one(oracle, np.zeros(6), D, np.zeros(6), _abi.WF_MAT_BILINEAR, 1e12)
    tr = D[:3].sum() / 3
    exp = 2 * 26.5e9 * 1e-7 * (D - np.array([tr, tr, tr, 0, 0, 0]))
    assert np.allclose(st["shearstress"][0], exp, rtol=1e-14)
    assert st["pl_strain"][0] == 0.0


@pytest.mark.parametrize("quirks", [0, _abi.WF_Q_TENSOR_PRODUCT])
def test_jaumann_rotation_terms(oracle, quirks):
    """D = 0: the update is dt*(S W^T + W S).  The true product gives the
    textbook Jaumann terms; the reference's product (first column A*B,
    others A*B^T, tensor3.cu:241-255) makes the y and z columns S W + W S,
    whose diagonal entries vanish."""
    S6 = np.array([1e8, -3e7, -7e7, 2e7, 5e7, -4e7])
    W6 = np.array([0.0, 0.0, 0.0, 300.0, -200.0, 150.0])
    dt = 1e-7
    st, _ = _one(oracle, S6, np.zeros(6), W6, _abi.WF_MAT_BILINEAR, 1e12, dt=dt, quirks=quirks)
    S = np.array([[S6[0], S6[3], S6[5]], [S6[3], S6[1], S6[4]], [S6[5], S6[4], S6[2]]])
    W = np.array([[0, W6[3], W6[5]], [-W6[3], 0, W6[4]], [-W6[5], -W6[4], 0]])
    if quirks:
        A = S @ W.T + W @ S
        B = S @ W + W @ S
        J = np.column_stack([A[:, 0], B[:, 1], B[:, 2]])
        assert abs(J[1, 1]) < 1e-6 and abs(J[2, 2]) < 1e-6
    else:
        J = S @ W.T + W @ S
    exp = S6 + dt * np.array([J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[1, 2], J[0, 2]])
    assert np.allclose(st["shearstress"][0], exp, rtol=1e-14, atol=1e-7)


def test_reference_return_map_bilinear(oracle):
    S6 = np.array([3e8, -1e8, -2e8, 1e8, 0.5e8, -0.8e8])
    sy, G, E, Et = 250e6, 26.5e9, 68.9e9, 1e9
    st, _ = _one(oracle, S6, np.zeros(6), np.zeros(6), _abi.WF_MAT_BILINEAR, sy, G=G)
    st_trial = _vm(S6)
    assert st_trial > sy
    assert _vm(st["shearstress"][0]) == pytest.approx(sy, rel=1e-13)  # scaled with the old s_y
    dep = (st_trial - sy) / (3 * G)  # Mechanical.cu:527, no Ep in the denominator
    assert st["pl_strain"][0] == pytest.approx(dep, rel=1e-13)
    Ep = E * Et / (E - Et)
    assert st["sigma_y"][0] == pytest.approx(sy + dep * Ep, rel=1e-13)
    assert st["sigma_eq"][0] == pytest.approx(sy, rel=1e-13)


def test_newton_return_map_bilinear_closed_form(oracle):
    S6 = np.array([3e8, -1e8, -2e8, 1e8, 0.5e8, -0.8e8])
    sy, G, E, Et = 250e6, 26.5e9, 68.9e9, 1e9
    st, _ = _one(oracle, S6, np.zeros(6), np.zeros(6), _abi.WF_MAT_BILINEAR, sy, G=G,
                 rm=_abi.WF_RETURN_NEWTON)
    Ep = E * Et / (E - Et)
    dg = (_vm(S6) - sy) / (3 * G + Ep)
    assert st["pl_strain"][0] == pytest.approx(dg, rel=1e-13)
    assert _vm(st["shearstress"][0]) == pytest.approx(sy + Ep * dg, rel=1e-12)


JC = dict(A=324e6, B=114e6, n=0.42, C=0.002, eps0=1.0, m=1.34, Tm=925.0, Tt=300.0)


def jc_textbook(ep, rate, T):
    th = min(max((T - JC["Tt"]) / (JC["Tm"] - JC["Tt"]), 0.0), 1.0)
    return (JC["A"] + JC["B"] * ep ** JC["n"]) * (1 + JC["C"] * math.log(rate / JC["eps0"])) * \
        (1 - th ** JC["m"])


def test_jc_yield_closed_form(oracle):
    mp = _abi.wf_material_params(**JC)
    for ep, rate, T in [(0.0, 1.0, 300.0), (0.1, 10.0, 400.0), (0.5, 1e3, 700.0)]:
        assert oracle.lib().or_jc_yield(mp, ep, rate, T) == pytest.approx(
            jc_textbook(ep, rate, T), rel=1e-14)


@pytest.mark.parametrize("ep", [0.0, 0.05, 0.4])
def test_jc_newton_matches_root(oracle, ep):
    import ctypes as C
    mp = _abi.wf_material_params(**JC)
    G, rate, T = 26.5e9, 50.0, 450.0
    s_trial = jc_textbook(ep, rate, T) + 120e6
    it = C.c_int()
    g = oracle.lib().or_jc_newton(mp, s_trial, G, ep, rate, T, 0.0, 1e-12, 60, C.byref(it))
    f = lambda x: s_trial - 3 * G * x - jc_textbook(ep + x, rate, T)
    root = optimize.brentq(f, 0.0, (s_trial - jc_textbook(ep, rate, T)) / (3 * G), xtol=1e-20,
                           rtol=1e-15)
    assert g == pytest.approx(root, rel=1e-10)
    assert abs(f(g)) <= 1e-12 * s_trial * 1.0001
    assert 0 < it.value < 60


def test_jc_linear_hardening_closed_form(oracle):
    """n = 1, C = 0, T = Tt: s_y = A + B ep, so dg = (s_t - A - B ep)/(3G + B)."""
    import ctypes as C
    mp = _abi.wf_material_params(**dict(JC, n=1.0, C=0.0))
    G, ep = 26.5e9, 0.1
    s_trial = 600e6
    it = C.c_int()
    g = oracle.lib().or_jc_newton(mp, s_trial, G, ep, 1.0, 300.0, 0.0, 1e-13, 60, C.byref(it))
    assert g == pytest.approx((s_trial - JC["A"] - JC["B"] * ep) / (3 * G + JC["B"]), rel=1e-12)


# ---------------------------------------------------------------------------
# thermal
def test_conduction_uniform_and_conservative(oracle):
    sc = scenarios.c2_thermal_cube(L=10e-3)
    st = to_oracle(oracle, sc)
    st["BC_T"] = 0
    st["T"] = 345.0
    st.nsearch(sc.cfg)
    st.calc_temp_inc(sc.cfg)
    assert np.all(st["dTdt"] == 0.0)
    rng = np.random.default_rng(1)
    st["T"] = 300 + 100 * rng.random(sc.n)
    st["k_T"] = 150 + 100 * rng.random(sc.n)
    st.calc_temp_inc(sc.cfg)
    heat = st["m"] * st["cp_T"] * st["dTdt"]
    assert abs(heat.sum()) < 1e-11 * np.abs(heat).sum()


def test_conduction_gaussian_decay(oracle):
    """Interior decay of a Gaussian hot spot follows the heat equation's
    closed form to SPH discretisation accuracy (a few percent)."""
    sc = scenarios.c2_thermal_cube(L=30e-3, r=0.5e-3)
    st = to_oracle(oracle, sc)
    st["BC_T"] = 0
    xyz = sc.fields["x"]
    c = xyz.mean(0)
    s0 = 4e-3
    d2 = ((xyz - c) ** 2).sum(1)
    st["T"] = 300 + 100 * np.exp(-d2 / (2 * s0 * s0))
    st.nsearch(sc.cfg)
    nst = 60
    st.step(sc.cfg, [], nst)
    alpha = 200.0 / (2700.0 * 900.0)
    t = (nst - 1) * sc.cfg.dt  # the first leapfrog step does not advance T
    s2 = s0 * s0 + 2 * alpha * t
    peak = 100 * (s0 * s0 / s2) ** 1.5
    i = np.argmin(d2)
    assert st["T"][i] - 300 == pytest.approx(peak, rel=0.05)


def test_int_energy_single_particle(oracle):
    st = oracle.State(1)
    st["shearstress"] = [1e8, -2e8, 1e8, 3e7, -4e7, 5e7]
    st["strrate"] = [10.0, -3.0, -7.0, 2.0, 1.0, -4.0]
    st["m"] = 2e-6
    st["rho"] = 2700.0
    cfg = _abi.default_config()
    cfg.n_particles = 1
    cfg.dt = 1e-7
    st.int_energy(cfg)
    s, d = st["shearstress"][0], st["strrate"][0]
    exp = 1e-7 * 2e-6 / 2700 * (s[:3] @ d[:3] + 2 * s[3:] @ d[3:])
    assert st["int_energy"][0] == pytest.approx(exp, rel=1e-14)


def test_integration_and_bcs(oracle):
    st = oracle.State(3)
    st["x"] = [[0, 0, 0], [1, 1, 1], [2, 2, 2]]
    st["v"] = [[1, 2, 3], [4, 5, 6], [7, 8, 9]]
    st["a"] = [[10, 0, -10], [1, 1, 1], [2, 2, 2]]
    st["ID"] = [0, 3, 0]
    dt = 0.01
    st.apply_bc_vel(3, (0, 0, -1))
    st.update_pos_fraser(dt)
    st.update_vel(dt)
    st.apply_bc_vel(3, (0, 0, -1))
    assert np.allclose(st["x"][0], [0.01 + 0.0005, 0.02, 0.03 - 0.0005], rtol=1e-15)
    assert np.allclose(st["v"][0], [1.1, 2.0, 2.9], rtol=1e-15)
    assert np.allclose(st["x"][1], [1, 1, 1 - 0.01], rtol=1e-15)
    assert np.all(st["v"][1] == [0, 0, -1]) and np.all(st["a"][1] == 0)


def test_oracle_c1_steps_hold_bcs(oracle):
    sc = small_c1(0.4)
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.step(sc.cfg, sc.bcs, 30)
    v = st["v"]
    assert np.isfinite(v).all()
    assert np.all(v[sc.fields["ID"] == 3] == [0, 0, -1.0])
    assert np.all(v[sc.fields["ID"] == 2] == 0.0)
    assert st["u"][sc.fields["ID"] == 3][:, 2].max() < 0


# ---------------------------------------------------------------------------
# Hollomon law (Material.cuh:51-59, 192-201): eps0 = sy0/E, eps1 = (sy0/K)^(1/m);
# sigma_y = sy0 on the plateau (eps_p + eps0 <= eps1), K (eps_p + eps0)^m above.
# With m = 1/2 and m = 1/3 the power law is a square or cube root, so the
# expected values below come from sqrt/cbrt, not from pow.
HOL = dict(E=68.9e9, sy0=1e8, K_h=1e9, m_h=0.5)


def _hol_params(**kw):
    return _abi.wf_material_params(model=_abi.WF_MAT_HOLLOMON, **dict(HOL, **kw))


def test_hollomon_yield_closed_form(oracle):
    mp = _hol_params()
    eps0 = HOL["sy0"] / HOL["E"]
    eps1 = (HOL["sy0"] / HOL["K_h"]) ** 2  # m = 1/2: (sy0/K)^2 = 0.01
    assert eps1 == pytest.approx(0.01, rel=1e-15)
    y = lambda ep: oracle.lib().or_hollomon_yield(mp, ep)
    # plateau, up to just below the knee eps_p = eps1 - eps0
    for ep in (0.0, 1e-4, 0.5 * (eps1 - eps0), (eps1 - eps0) * (1 - 1e-9)):
        assert y(ep) == HOL["sy0"]
    # power branch: K sqrt(eps_p + eps0)
    for ep in (0.009, 0.015, 0.1, 0.7):
        assert y(ep) == pytest.approx(1e9 * math.sqrt(ep + eps0), rel=1e-14)
    # continuous at the knee, increasing after it
    knee = eps1 - eps0
    assert y(knee * (1 + 1e-9)) == pytest.approx(HOL["sy0"], rel=1e-8)
    assert y(0.02) < y(0.03) < y(0.3)
    # m = 1/3: cube roots
    mp3 = _hol_params(K_h=5e8, m_h=1.0 / 3.0)
    for ep in (0.05, 0.2, 0.9):
        assert oracle.lib().or_hollomon_yield(mp3, ep) == pytest.approx(
            5e8 * np.cbrt(ep + eps0), rel=1e-14)
    assert oracle.lib().or_hollomon_yield(mp3, 0.0) == HOL["sy0"]  # (0.2)^3 = 8e-3 > eps0


def test_hollomon_reference_return_map(oracle):
    """Reference return map (Mechanical.cu:504-535) with Hollomon: the trial
    deviator is scaled with the OLD sigma_y, sigma_y is re-evaluated at the
    OLD eps_p, dep = (s_trial - sigma_y)/(3G) and sigma_eq = the new sigma_y."""
    S6 = np.array([3e8, -1e8, -2e8, 1e8, 0.5e8, -0.8e8])
    G, ep, sy_old = 26.5e9, 0.02, 1.3e8
    st, _ = _one(oracle, S6, np.zeros(6), np.zeros(6), _abi.WF_MAT_HOLLOMON, sy_old, G=G,
                 ep=ep, mat=HOL)
    s_trial = _vm(S6)
    assert s_trial > sy_old
    sy_new = 1e9 * math.sqrt(ep + HOL["sy0"] / HOL["E"])
    assert _vm(st["shearstress"][0]) == pytest.approx(sy_old, rel=1e-13)
    assert st["sigma_y"][0] == pytest.approx(sy_new, rel=1e-14)
    dep = (s_trial - sy_new) / (3 * G)
    assert st["pl_strain"][0] == pytest.approx(ep + dep, rel=1e-13)
    assert st["sigma_eq"][0] == pytest.approx(sy_new, rel=1e-14)
    # elastic particle: nothing changes but sigma_y (re-evaluated at eps_p)
    st2, _ = _one(oracle, S6 * 0.2, np.zeros(6), np.zeros(6), _abi.WF_MAT_HOLLOMON, 2e8, G=G,
                  ep=ep, mat=HOL)
    assert st2["pl_strain"][0] == ep
    assert np.allclose(st2["shearstress"][0], S6 * 0.2, rtol=1e-15)
    assert st2["sigma_y"][0] == pytest.approx(sy_new, rel=1e-14)


@pytest.mark.parametrize("ep", [0.0, 0.02, 0.2])
def test_hollomon_consistent_return_closed_form(oracle, ep):
    """Consistent radial return: s_t - 3 G dg = K sqrt(ep + dg + eps0).  With
    u = sqrt(ep + dg + eps0) that is 3G u^2 + K u - (s_t + 3G (ep + eps0)) = 0,
    a quadratic with one positive root; a root short of the knee (u^2 below
    eps1) means the return ends on the plateau, dg = (s_t - sy0)/(3G).
    ep = 0 ends on the plateau, 0.02 and 0.2 on the power branch."""
    S6 = np.array([3e8, -1e8, -2e8, 1e8, 0.5e8, -0.8e8])
    G = 26.5e9
    st, _ = _one(oracle, S6, np.zeros(6), np.zeros(6), _abi.WF_MAT_HOLLOMON, 1e8, G=G, ep=ep,
                 rm=_abi.WF_RETURN_NEWTON, mat=HOL)
    K, eps0, s_t = HOL["K_h"], HOL["sy0"] / HOL["E"], _vm(S6)
    c = s_t + 3 * G * (ep + eps0)
    u = (-K + math.sqrt(K * K + 12 * G * c)) / (6 * G)
    if u * u > 0.01:  # the return ends past the knee, on the power branch
        dg, sy = u * u - ep - eps0, K * u
    else:  # it ends on the plateau: perfect plasticity at sy0
        dg, sy = (s_t - HOL["sy0"]) / (3 * G), HOL["sy0"]
        assert ep + dg + eps0 <= 0.01
    assert dg > 0
    assert st["pl_strain"][0] == pytest.approx(ep + dg, rel=1e-10)
    assert st["sigma_y"][0] == pytest.approx(sy, rel=1e-10)
    assert st["sigma_eq"][0] == pytest.approx(s_t - 3 * G * dg, rel=1e-10)
    assert _vm(st["shearstress"][0]) == pytest.approx(sy, rel=1e-10)


# ---------------------------------------------------------------------------
# equation of state, case 0 (Functions.cu:207-283): p = P0 + Cs^2 (rho - rho0)
# with P0 = 0 (never uploaded by the reference driver); DensitySolid inverts it.
def test_eos_round_trip_and_bulk_modulus(oracle):
    rng = np.random.default_rng(12)
    n = 64
    st = oracle.State(n)
    rho0 = rng.uniform(2000.0, 8000.0, n)
    Cs = rng.uniform(1000.0, 6000.0, n)
    rho = rho0 * (1 + rng.uniform(-0.02, 0.02, n))
    st["rho"], st["rho0"], st["Cs"] = rho, rho0, Cs
    st.pressure()
    p = st["p"].copy()
    # DensitySolid case 0 (Functions.cu:259-283): (p - P0)/Cs^2 + rho0 = rho
    assert np.allclose(p / (Cs * Cs) + rho0, rho, rtol=1e-14, atol=0.0)
    # sign: compression (rho > rho0) gives positive pressure
    assert np.all(np.sign(p) == np.sign(rho - rho0))
    # with Cs^2 = K/rho0 (the driver's Cs, WeldFormGPU.cu:253-267), the EOS is
    # the small-strain bulk law p = K (rho/rho0 - 1) = -K dV/V to first order
    E, nu, r0 = 68.9e9, 0.3, 2700.0
    K = E / (3 * (1 - 2 * nu))
    st1 = oracle.State(1)
    st1["rho0"] = r0
    st1["Cs"] = math.sqrt(K / r0)
    for vol in (1e-4, -3e-4):
        st1["rho"] = r0 / (1 + vol)
        st1.pressure()
        assert st1["p"][0] == pytest.approx(-K * vol, rel=2 * abs(vol))


# ---------------------------------------------------------------------------
# convection BC (CalcConvHeatKernel, Thermal.cu:36-64): one particle of
# volume V = m/rho with exposed face A = V^(2/3) (a dx^2 face for a dx^3
# particle) cools by Newton's law m cp dT/dt = h A (T_inf - T).
def _convecting_particle(oracle, T0, dx=1e-3, rho=2700.0, cp=900.0, h_conv=100.0,
                         T_inf=300.0, dt=0.1):
    st = oracle.State(1)
    st["x"] = [[0.0, 0.0, 0.0]]
    st["h"] = 1.2 * dx
    st["rho"], st["rho0"] = rho, rho
    st["m"] = rho * dx ** 3
    st["T"] = T0
    st["k_T"] = 200.0
    st["cp_T"] = cp
    st["BC_T"] = _abi.WF_BCT_CONVECTION
    cfg = _abi.default_config()
    cfg.n_particles = cfg.solid_count = 1
    cfg.h_glob = 1.2 * dx
    cfg.mechanics, cfg.thermal = 0, 1
    cfg.dt = dt
    cfg.h_conv, cfg.T_inf = h_conv, T_inf
    st.nsearch(cfg)
    assert st["nb_count"][0] == 0
    return st, cfg


def test_convection_newton_cooling_rate(oracle):
    dx, rho, cp, h, Tinf = 1e-3, 2700.0, 900.0, 100.0, 300.0
    for T0 in (500.0, 250.0):
        st, cfg = _convecting_particle(oracle, T0, dx, rho, cp, h, Tinf)
        st.calc_temp_inc(cfg)
        m = rho * dx ** 3
        # exponent 0.666666666 instead of 2/3 changes (1e-9)^(2/3) by 1.4e-8
        assert st["dTdt"][0] == pytest.approx(h * dx * dx * (Tinf - T0) / (m * cp), rel=1e-7)


def test_convection_exponential_decay(oracle):
    """Leapfrog over 200 steps follows T_inf + (T0 - T_inf) exp(-t/tau),
    tau = m cp/(h A) = 24.3 s, to the scheme's O(dt^2) accuracy."""
    dx, rho, cp, h, Tinf, T0, dt = 1e-3, 2700.0, 900.0, 100.0, 300.0, 500.0, 0.1
    st, cfg = _convecting_particle(oracle, T0, dx, rho, cp, h, Tinf, dt)
    nst = 200
    st.step(cfg, [], nst)
    tau = rho * dx ** 3 * cp / (h * dx * dx)
    t = (nst - 1) * dt  # the first leapfrog step does not advance T
    exp = Tinf + (T0 - Tinf) * math.exp(-t / tau)
    assert st["T"][0] - Tinf == pytest.approx(exp - Tinf, rel=1e-3)
    assert st["T"][0] < T0
