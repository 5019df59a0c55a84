"""Pinning of the oracle's contact restatement (Contact.cu, Geometry.cu,
Mesh.cu) by closed forms, lattice sums and brute force, on the CPU.

Single particle under a z-normal plate (normal -z, plane z = 1), mass m,
smoothing length h, time step dt, gap g = 1 - z below the plate:
  penetration delta = h - g, reference law F = 0.4 (2 m/dt^2) delta (-z),
  so a_z = -0.8 delta/dt^2; stiffness law F = kappa E dx delta (-z).
A particle on the shared edge of two triangles is inside both: the second
contact sees the position predicted with the first contact's acceleration,
delta_2 = 0.6 delta, so a_z = -0.8 (1 + 0.6) delta/dt^2 and the stored
contact force is the second one (Contact.cu:247-250).
Friction (Contact.cu:294-308), tangential speed v_t, plate at rest:
static if k v_t dt < mu_s |F_n| with k = 2 m/dt^2, then a_t = -2 v_t/dt
(the predicted slip is cancelled); otherwise a_t = -mu_d |F_n|/m.
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

M, H, DT = 2.0, 0.3, 0.01


def _cfg(**kw):
    c = _abi.default_config()
    c.dt = DT
    c.contact = 1
    c.h_glob = H
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _plate(po, dens=2):
    nodes, el = po.gen_plane_mesh([0.0, 0.0, 1.0], [2.0, 2.0, 1.0], dens)
    return {"nodes": nodes, "elnode": el}


def _one(po, pos, v=(0, 0, 0), mesh=None, **cfgkw):
    st = po.State(1)
    st["x"] = [pos]
    st["xs"] = [pos]
    st["v"] = [v]
    st["m"] = [M]
    st["h"] = [H]
    st["ID"] = [1]
    st.add_contact([mesh or _plate(po)])
    cfg = _cfg(**cfgkw)
    st.contact_nb(cfg)
    st.contact_forces(cfg)
    return st


def test_plate_mesh_matches_library_and_geometry(oracle):
    p1, p2, dens = [-0.02, -0.02, 0.031], [0.02, 0.02, 0.031], 16
    nodes, el = oracle.gen_plane_mesh(p1, p2, dens)
    n2, e2 = scenarios.gen_plane_mesh(p1, p2, dens)
    assert np.array_equal(nodes, n2) and np.array_equal(el, e2)
    assert nodes.shape == ((dens + 1) ** 2, 3) and el.shape == (2 * dens * dens, 3)
    st = oracle.State(1)
    st.add_contact([{"nodes": nodes, "elnode": el}])
    nrm, pp, c = (st.mesh_arrays[k] for k in ("nrm", "pp", "c"))
    assert np.array_equal(nrm, np.tile([0.0, 0.0, -1.0], (len(el), 1)))
    assert np.all(pp == -0.031)
    assert np.allclose(c, nodes[el].mean(axis=1), rtol=0, atol=1e-15)
    # the triangles tile the square: areas sum to L^2, every quad split once
    P = nodes[el]
    area = 0.5 * np.abs(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])[:, 2])
    assert abs(area.sum() - 0.04 ** 2) < 1e-15
    assert len({tuple(sorted(t)) for t in el.tolist()}) == len(el)


def test_single_particle_inside_one_triangle(oracle):
    st = _one(oracle, (0.2, 0.3, 0.9))
    delta = H - (1.0 - 0.9)
    az = -0.8 * delta / DT ** 2
    assert st["a"][0, 2] == pytest.approx(az, rel=1e-12)
    assert st["a"][0, 0] == 0.0 and st["a"][0, 1] == 0.0
    assert st.contact_arrays["cf"][0, 2] == pytest.approx(M * az, rel=1e-12)


def test_shared_edge_contacts_twice_with_updated_prediction(oracle):
    st = _one(oracle, (0.5, 0.5, 0.9))  # on the diagonal x + y = 1 of quad (0,0)
    delta = H - 0.1
    assert st.contact_arrays["cnb_count"][0] >= 2
    assert st["a"][0, 2] == pytest.approx(-0.8 * delta * 1.6 / DT ** 2, rel=1e-12)
    assert st.contact_arrays["cf"][0, 2] == pytest.approx(-0.8 * M * 0.6 * delta / DT ** 2,
                                                          rel=1e-12)


def test_no_force_beyond_h_or_outside_the_plate(oracle):
    assert np.all(_one(oracle, (0.2, 0.3, 0.6))["a"] == 0.0)       # gap 0.4 > h
    st = _one(oracle, (2.1, 0.5, 0.9))                              # beyond the edge x = 2
    assert st.contact_arrays["cnb_count"][0] > 0 and np.all(st["a"] == 0.0)


def test_stiffness_law(oracle):
    E, dx, kappa = 70e9, 0.25e-3, 0.1
    st = _one(oracle, (0.2, 0.3, 0.9), contact_law=_abi.WF_CONTACT_STIFFNESS,
              contact_kappa=kappa, contact_dx=dx, mat=_abi.wf_material_params(E=E))
    assert st["a"][0, 2] == pytest.approx(-kappa * E * dx * (H - 0.1) / M, rel=1e-12)


@pytest.mark.parametrize("vt,static", [(1.0, True), (10.0, False), (-3.5, True)])
def test_coulomb_friction_branches(oracle, vt, static):
    mu_s, mu_d = 0.5, 0.3
    st = _one(oracle, (0.2, 0.3, 0.9), v=(vt, 0.0, 0.0), fric_sta=mu_s, fric_dyn=mu_d)
    delta = H - 0.1
    fn = 0.8 * M * delta / DT ** 2
    assert (abs(vt) * DT < 0.4 * mu_s * delta) == static  # the branch the case means
    a, cf = st["a"][0], st.contact_arrays["cf"][0]
    if static:
        assert a[0] == pytest.approx(-2.0 * vt / DT, rel=1e-12)
        assert cf[0] == 0.0
    else:
        assert a[0] == pytest.approx(-np.sign(vt) * mu_d * fn / M, rel=1e-12)
        assert cf[0] == pytest.approx(-np.sign(vt) * mu_d * fn, rel=1e-12)
    assert a[2] == pytest.approx(-fn / M, rel=1e-12)


@pytest.mark.parametrize("vt,static", [(1.0, True), (10.0, False), (-14.0, False)])
def test_sliding_friction_heat(oracle, vt, static):
    """Friction heat (an extension): while sliding, the dissipated power
    fric_heat x mu_d |F_n| |v_t| enters the particle (the slip over dt is
    v_t dt, the tangential acceleration being zero before the contact);
    sticking dissipates nothing.  It then heats the particle at
    q / (m cp) in the conduction pass (a lone particle: no conduction)."""
    mu_s, mu_d, beta = 0.5, 0.3, 0.7
    st = _one(oracle, (0.2, 0.3, 0.9), v=(vt, 0.0, 0.0), fric_sta=mu_s, fric_dyn=mu_d,
              fric_heat=beta)
    fn = 0.8 * M * (H - 0.1) / DT ** 2
    q = st["q_friction"][0]
    if static:
        assert q == 0.0
    else:
        assert q == pytest.approx(beta * mu_d * fn * abs(vt), rel=1e-12)
    cp, rho = 900.0, 2700.0
    st["cp_T"] = [cp]
    st["k_T"] = [200.0]
    st["rho"] = [rho]
    st["T"] = [300.0]
    cfg = _cfg(fric_sta=mu_s, fric_dyn=mu_d, fric_heat=beta)
    cfg.thermal = 1
    st.nsearch(cfg)
    st.calc_temp_inc(cfg)
    assert st["dTdt"][0] == pytest.approx(q / (M * cp), rel=1e-12)


def test_friction_heat_off_by_default(oracle):
    st = _one(oracle, (0.2, 0.3, 0.9), v=(10.0, 0.0, 0.0), fric_sta=0.5, fric_dyn=0.3)
    assert st["q_friction"][0] == 0.0 and st.contact_arrays["cf"][0, 0] != 0.0


def test_friction_against_a_moving_plate_uses_node_velocity(oracle):
    """Plate sliding with the particle: no relative slip, no tangential force;
    the element velocity comes from the last mesh move (zero before it)."""
    st = oracle.State(1)
    st["x"] = [(0.2, 0.3, 0.9)]
    st["v"] = [(1.0, 0.0, 0.0)]
    st["m"] = [M]
    st["h"] = [H]
    st["ID"] = [1]
    st.add_contact([_plate(oracle)])
    st.set_mesh_motion(0, vel=(1.0, 0.0, 0.0))
    cfg = _cfg(fric_sta=0.5, fric_dyn=0.3)
    st.mesh_move(1e-9)  # node velocities set, plate barely moved
    assert np.all(st.mesh_arrays["v"][:, 0] == 1.0)
    st.contact_nb(cfg)
    st.contact_forces(cfg)
    assert abs(st["a"][0, 0]) < 1e-6 * abs(st["a"][0, 2])


def test_mesh_rotation_about_a_centre(oracle):
    nodes, el = oracle.gen_plane_mesh([-1.0, -1.0, 0.5], [1.0, 1.0, 0.5], 4)
    st = oracle.State(1)
    st.add_contact([{"nodes": nodes, "elnode": el}])
    w, c, dt = np.array([0.0, 0.0, 3.0]), np.array([0.25, -0.5, 0.0]), 1e-3
    st.set_mesh_motion(0, vel=(0.1, 0.0, -2.0), omega=w, center=c)
    st.mesh_move(dt)
    v_exp = np.array([0.1, 0.0, -2.0]) + np.cross(w, nodes - c)
    assert np.allclose(st.mesh_arrays["v"], v_exp, rtol=0, atol=1e-14)
    assert np.allclose(st.mesh_arrays["x"], nodes + dt * v_exp, rtol=0, atol=1e-15)
    assert np.allclose(st.mesh_arrays["nrm"], [0.0, 0.0, -1.0], rtol=0, atol=1e-15)
    assert np.allclose(st.mesh_arrays["pp"], -(0.5 - 2e-3), rtol=0, atol=1e-15)


def test_contact_candidates_brute_force(oracle):
    rng = np.random.default_rng(5)
    nodes, el = oracle.gen_plane_mesh([0.0, 0.0, 1.0], [2.0, 2.0, 1.0], 12)
    n = 400
    st = oracle.State(n)
    x = rng.uniform([-0.3, -0.3, 0.5], [2.3, 2.3, 1.5], (n, 3))
    st["x"] = x
    st["h"] = rng.uniform(0.05, 0.45, n)
    ID = rng.integers(0, 3, n).astype(np.int32)
    st["ID"] = ID
    st.add_contact([{"nodes": nodes, "elnode": el}])
    st.contact_nb(_cfg())
    cent = nodes[el].sum(axis=1) / 3.0
    cnt, lst = st.contact_arrays["cnb_count"], st.contact_arrays["cnb_list"]
    over = 0
    for i in range(n):
        d = np.sqrt(((x[i] - cent) ** 2).sum(1))
        cand = np.nonzero(d < 2 * st["h"][i])[0] if ID[i] == 1 else np.array([], int)
        over += len(cand) > 20
        assert cnt[i] == min(len(cand), 20)
        assert np.array_equal(lst[i, :cnt[i]], cand[:20])
    assert over > 0 and st.ct.cont_over_cap == over


def _lattice_surface(po, nside, ids=None):
    """Cube lattice of spacing 1, unit masses, h = 1.2: oracle surface normals
    and flags, with the neighbour lists of the oracle's search."""
    xyz, _ = po.gen_box([0.0, 0.0, 0.0], [nside * 1.0] * 3, 0.5, 1.0)
    n = len(xyz)
    st = po.State(n)
    st["x"] = xyz
    st["xs"] = xyz
    st["m"] = np.ones(n)
    st["h"] = np.full(n, 1.2)
    st["ID"] = np.zeros(n, np.int32) if ids is None else ids(xyz)
    cfg = _cfg(h_glob=1.2)
    st.nsearch(cfg)
    st.add_contact([])  # no rigid triangles
    st.calc_surface(cfg)
    return xyz, st


def test_surface_normals_lattice_sums(oracle):
    xyz, st = _lattice_surface(oracle, 8)
    idx = np.rint(xyz - xyz.min(0)).astype(int)   # integer lattice coordinates 0..7
    offs = np.array([(i, j, k) for i in range(-2, 3) for j in range(-2, 3)
                     for k in range(-2, 3) if 0 < i * i + j * j + k * k <= 5])
    normal = st.contact_arrays["normal"]
    for p in range(len(xyz)):
        q = idx[p] + offs
        inside = np.all((q >= 0) & (q <= 7), axis=1)
        nb = inside.sum()
        exp = -offs[inside].sum(0) / nb     # sum_j (x_i - x_j) / nb_i, unit spacing and mass
        assert np.allclose(normal[p], exp, rtol=0, atol=1e-12), p
        flag = np.linalg.norm(exp) >= 0.25 * 1.2 and nb <= 46
        assert (st["ID"][p] == 1) == flag, p
    # interior particles have zero normal and keep ID 0; face centres point out
    assert np.all(st["ID"][np.all((idx >= 2) & (idx <= 5), axis=1)] == 0)
    c = np.nonzero((idx[:, 0] == 7) & (idx[:, 1] == 3) & (idx[:, 2] == 4))[0][0]
    assert normal[c, 0] > 0.25 * 1.2 and abs(normal[c, 1]) < 1e-12


def test_surface_keeps_nonzero_ids(oracle):
    _, st = _lattice_surface(oracle, 6, ids=lambda x: np.where(x[:, 2] < 1.0, 2, 0)
                             .astype(np.int32))
    z = st["x"][:, 2]
    assert np.all(st["ID"][z < 1.0] == 2)
    assert np.any(st["ID"][z > 1.0] == 1)


def test_oracle_step_with_plate_contact(oracle):
    """A plate driven into a cylinder: contact forces appear, push the top
    down, and no surface particle passes through the plate."""
    sc = scenarios.c4_contact(scale=0.15, r=0.3e-3, dens=4, thermal=False,
                              law=_abi.WF_CONTACT_REFERENCE)
    st = oracle.State(sc.n)
    for k, v in sc.fields.items():
        st[{"x_snapshot": "xs"}.get(k, k)] = v
    st["xs"] = sc.fields["x"]
    st.add_contact(sc.meshes)
    st.nsearch(sc.cfg)
    st.step(sc.cfg, sc.bcs, 400)
    cf = st.contact_arrays["cf"]
    zp = st.mesh_arrays["x"][0, 2]
    assert zp < sc.meta["plate_z"] - 399 * 2.0 * sc.cfg.dt
    assert np.linalg.norm(cf, axis=1).sum() > 0.0
    assert np.all(cf[:, 2] <= 0.0)
    top = st["ID"] == 1
    assert st["x"][top, 2].max() < zp + 1e-12
    assert st["u"][:, 2].min() < 0.0
