"""Synthetic input decks for the step pipeline (the workloads of the tests
and the benchmark), built the way WeldFormGPU.cu's main builds a domain
(WeldFormGPU.cu:199-906):

  K = E/(3(1-2nu)), G = E/(2(1+nu)), dx = 2r, h = hFactor*dx, Cs = sqrt(K/rho)
  (WeldFormGPU.cu:253-267); dt = 0.4 h/(Cs + VMAX), VMAX = 10
  (WeldFormGPU.cu:62,904-906); lattice from AddCylinderLength/AddBoxLength;
  zones by AssignZone; velocity BCs per zone.

The lattices come from libwfhip's host builders (wf_gen_cylinder /
wf_gen_box), which need no GPU.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi


@dataclass
class Scenario:
    name: str
    cfg: _abi.wf_config
    fields: dict
    bcs: list = field(default_factory=list)
    meta: dict = field(default_factory=dict)
    # rigid meshes: dicts with nodes (nn,3), elnode (ne,3), zone, and the
    # rigid motion vel, omega, center
    meshes: list = field(default_factory=list)

    @property
    def n(self):
        return self.cfg.n_particles


def _lib():
    lib = C.CDLL(_abi.LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    lib.wf_gen_box.restype = I
    lib.wf_gen_box.argtypes = [P, D, D, D, D, D, P, I, P]
    lib.wf_gen_cylinder.restype = I
    lib.wf_gen_cylinder.argtypes = [P, D, D, D, D, P, I, P]
    lib.wf_gen_cylinder_polar.restype = I
    lib.wf_gen_cylinder_polar.argtypes = [I, D, D, D, D, D, I, D, P, P, I, P]
    lib.wf_assign_zone.restype = I
    lib.wf_assign_zone.argtypes = [P, I, P, P, I, P]
    lib.wf_gen_plane_mesh.restype = I
    lib.wf_gen_plane_mesh.argtypes = [P, P, I, P, I, P, I, C.POINTER(I)]
    lib.wf_read_nastran.restype = I
    lib.wf_read_nastran.argtypes = [C.c_char_p, I, P, I, P, I, C.POINTER(I)]
    return lib


def gen_cylinder(start, R, Lz, r, density):
    lib = _lib()
    V = np.ascontiguousarray(start, np.float64)
    n = lib.wf_gen_cylinder(V.ctypes.data, R, Lz, r, density, None, 0, None)
    xyz = np.zeros((n, 3))
    m = C.c_double()
    lib.wf_gen_cylinder(V.ctypes.data, R, Lz, r, density, xyz.ctypes.data, n, C.byref(m))
    return xyz, m.value


def gen_cylinder_polar(R, Lz, r, density, ang=2.0 * np.pi, rows=1, r_i=0.0, tag=0):
    """The cylindrical grid (Domain::AddCylUniformLength): positions, IDs, mass."""
    lib = _lib()
    n = lib.wf_gen_cylinder_polar(tag, R, Lz, r, density, ang, rows, r_i, None, None, 0, None)
    if n < 0:
        raise ValueError(f"wf_gen_cylinder_polar: error {n}")
    xyz = np.zeros((n, 3))
    ID = np.zeros(n, np.int32)
    m = C.c_double()
    lib.wf_gen_cylinder_polar(tag, R, Lz, r, density, ang, rows, r_i, xyz.ctypes.data,
                              ID.ctypes.data, n, C.byref(m))
    return xyz, ID, m.value


def gen_box(start, L, r, density):
    lib = _lib()
    V = np.ascontiguousarray(start, np.float64)
    n = lib.wf_gen_box(V.ctypes.data, L[0], L[1], L[2], r, density, None, 0, None)
    xyz = np.zeros((n, 3))
    m = C.c_double()
    lib.wf_gen_box(V.ctypes.data, L[0], L[1], L[2], r, density, xyz.ctypes.data, n, C.byref(m))
    return xyz, m.value


def gen_plane_mesh(p1, p2, dens):
    """z-normal plate of 2 dens^2 triangles (AxisPlaneMesh): nodes, elnode."""
    lib = _lib()
    a = np.ascontiguousarray(p1, np.float64)
    b = np.ascontiguousarray(p2, np.float64)
    nn = C.c_int()
    ne = lib.wf_gen_plane_mesh(a.ctypes.data, b.ctypes.data, int(dens), None, 0, None, 0,
                               C.byref(nn))
    if ne < 0:
        raise ValueError("bad plate mesh arguments")
    nodes = np.zeros((nn.value, 3))
    el = np.zeros((ne, 3), np.int32)
    lib.wf_gen_plane_mesh(a.ctypes.data, b.ctypes.data, int(dens), nodes.ctypes.data, nn.value,
                          el.ctypes.data, ne, C.byref(nn))
    return nodes, el


def read_nastran(path, flip=False):
    """GRID/CTRIA3 triangle mesh of a Nastran small-field file: nodes, elnode."""
    lib = _lib()
    nn = C.c_int()
    bp = str(path).encode()
    ne = lib.wf_read_nastran(bp, int(flip), None, 0, None, 0, C.byref(nn))
    if ne < 0:
        raise ValueError(f"cannot read Nastran mesh {path}")
    nodes = np.zeros((nn.value, 3))
    el = np.zeros((ne, 3), np.int32)
    lib.wf_read_nastran(bp, int(flip), nodes.ctypes.data, nn.value, el.ctypes.data, ne,
                        C.byref(nn))
    return nodes, el


def assign_zone(xyz, ID, start, end, zid):
    lib = _lib()
    s = np.ascontiguousarray(start, np.float64)
    e = np.ascontiguousarray(end, np.float64)
    return lib.wf_assign_zone(xyz.ctypes.data, len(xyz), s.ctypes.data, e.ctypes.data, zid,
                              ID.ctypes.data)


def _mech_fields(xyz, mass, rho, h, Cs, G, sy0):
    n = len(xyz)
    return {
        "x": xyz.copy(), "v": np.zeros((n, 3)), "rho": np.full(n, rho),
        "rho0": np.full(n, rho), "m": np.full(n, mass), "h": np.full(n, h),
        "Cs": np.full(n, Cs), "G": np.full(n, G), "sigma_y": np.full(n, sy0),
        "ID": np.zeros(n, np.int32),
    }


def cylinder_compression(R, H, r, *, rho, E, nu, material, v_top, hfac=1.2, thermal=None,
                         name="cylinder", layers=2, slice_angle=None):
    """Cylinder on the z axis, bottom `layers` layers held at v=0 (zone 2),
    top `layers` layers driven at (0,0,v_top) (zone 3).

    slice_angle None: the Cartesian lattice (AddCylinderLength).  Otherwise
    the cylindrical grid a deck with gridCoordSys "Cylindrical" and this
    sliceAngle gets (WeldFormGPU.cu:366-375): the whole circle for exactly
    2 pi, else a pi/16 slice of the outer half of the radius with the
    axisymmetric acceleration clamp; the grid's IDs (2, 3 on the slice edges,
    4 on the axis) stand where no zone overrides them."""
    ID0 = None
    axisymm = 0
    if slice_angle is None:
        xyz, mass = gen_cylinder([0.0, 0.0, 0.0], R, H, r, rho)
    elif slice_angle == 2.0 * np.pi:
        xyz, ID0, mass = gen_cylinder_polar(R, H, r, rho)
    else:
        xyz, ID0, mass = gen_cylinder_polar(R, H, r, rho, np.pi / 16.0, 1, R / 2.0)
        axisymm = 1
    n = len(xyz)
    dx = 2.0 * r
    h = dx * hfac
    K = E / (3. * (1. - 2 * nu))
    G = E / (2. * (1. + nu))
    Cs = np.sqrt(K / rho)
    cfg = _abi.default_config()
    cfg.n_particles = n
    cfg.solid_count = n
    cfg.axisymm = axisymm
    cfg.h_glob = h
    cfg.dt = 0.4 * h / (Cs + 10.0)
    mp = cfg.mat
    mp.E, mp.nu = E, nu
    for k, v in material.items():
        setattr(mp, k, v)
    sy0 = material.get("sy0", 0.0)
    f = _mech_fields(xyz, mass, rho, h, Cs, G, sy0)
    if ID0 is not None:
        f["ID"][:] = ID0
    zmin, zmax = xyz[:, 2].min(), xyz[:, 2].max()
    big = 10 * R
    lay = (layers - 0.5) * dx
    assign_zone(xyz, f["ID"], [-big, -big, -big], [big, big, zmin + lay], 2)
    assign_zone(xyz, f["ID"], [-big, -big, zmax - lay], [big, big, big], 3)
    bcs = [(2, (0.0, 0.0, 0.0)), (3, (0.0, 0.0, v_top))]
    if thermal:
        cfg.thermal = 1
        f["T"] = np.full(n, thermal["T0"])
        f["k_T"] = np.full(n, thermal["k"])
        f["cp_T"] = np.full(n, thermal["cp"])
        cfg.taylor_quinney = thermal.get("taylor_quinney", 0.0)
    return Scenario(name, cfg, f, bcs, {"R": R, "H": H, "r": r, "dx": dx, "h": h, "Cs": Cs,
                                        "mass": mass})


def c1_bilinear(scale=1.0):
    """Elastic-plastic steel cylinder, bilinear hardening: R 7.5 mm, H 15 mm,
    dx 0.5 mm (scale multiplies R and H), top at -1 m/s."""
    E = 200e9
    return cylinder_compression(7.5e-3 * scale, 15e-3 * scale, 0.25e-3, rho=7850.0, E=E, nu=0.3,
                                material={"model": _abi.WF_MAT_BILINEAR, "sy0": 300e6,
                                          "Et": 1e9},
                                v_top=-1.0, name="c1_bilinear_cylinder")


JC_AL = {"model": _abi.WF_MAT_JOHNSON_COOK, "A": 324e6, "B": 114e6, "n": 0.42, "C": 0.002,
         "eps0": 1.0, "m": 1.34, "Tm": 925.0, "Tt": 300.0}


def _adaptive_training(cfg, interval=200, interval_max=3200):
    """Surrogate rounds every `interval` steps while yielding, the interval
    doubling up to `interval_max` while the network still fits the new
    samples (a round's starting loss within 2x of the last final loss)."""
    cfg.mlp_train_interval = interval
    cfg.mlp_interval_max = interval_max
    cfg.mlp_refit_ratio = 2.0


def c3_johnson_cook(scale=1.0, r=0.15e-3):
    """Coupled thermo-mechanical Johnson-Cook compression: R 15 mm, H 30 mm,
    dx 0.3 mm, aluminium-like JC, Taylor-Quinney 0.9, top at -2 m/s,
    FP64 Newton return map."""
    T0 = 300.0
    sc = cylinder_compression(15e-3 * scale, 30e-3 * scale, r, rho=2700.0, E=68.9e9, nu=0.3,
                              material=dict(JC_AL, sy0=JC_AL["A"]), v_top=-2.0,
                              thermal={"T0": T0, "k": 200.0, "cp": 900.0,
                                       "taylor_quinney": 0.9},
                              name="c3_jc_thermomech_cylinder")
    sc.cfg.return_map = _abi.WF_RETURN_NEWTON
    _adaptive_training(sc.cfg)
    # initial flow stress at eps_p = 0, rate 0 (-> 1e-5), T0: Material.cuh:209-219
    mp = sc.cfg.mat
    th = min(max((T0 - mp.Tt) / (mp.Tm - mp.Tt), 0.0), 1.0)
    sy = mp.A * (1.0 + mp.C * np.log(1e-5 / mp.eps0)) * (1.0 - th ** mp.m)
    sc.fields["sigma_y"][:] = sy
    return sc


def c4_contact(scale=1.0, r=0.125e-3, dens=64, law=_abi.WF_CONTACT_STIFFNESS, mu=0.2,
               plate=40e-3, offset=0.0, thermal=True):
    """Johnson-Cook cylinder (R 15 mm, H 30 mm; scale multiplies R, H and the
    plate) compressed by a rigid plate of plate x plate with 2 dens^2
    triangles moving at -2 m/s, starting one h above the top layer; bottom
    two layers fixed.  Penalty law `law` (stiffness 0.1 E dx or the
    reference 0.4 * 2 m/dt^2), Coulomb friction mu (static = dynamic).
    `offset` shifts the plate in x and y by that fraction of its spacing."""
    sc = c3_johnson_cook(scale=scale, r=r)
    sc.name = "c4_jc_plate_contact"
    if not thermal:
        sc.cfg.thermal = 0
        sc.cfg.taylor_quinney = 0.0
    sc.bcs = [b for b in sc.bcs if b[0] != 3]  # the plate drives the top
    ID = sc.fields["ID"]
    ID[ID == 3] = 0
    cfg = sc.cfg
    cfg.contact = 1
    cfg.contact_law = law
    cfg.contact_kappa = 0.1
    cfg.contact_dx = sc.meta["dx"]
    cfg.fric_sta = mu
    cfg.fric_dyn = mu
    L = plate * scale
    zt = sc.fields["x"][:, 2].max() + sc.meta["h"]
    sh = offset * L / dens
    p1 = np.array([-L / 2 + sh, -L / 2 + sh, zt])
    nodes, el = gen_plane_mesh(p1, p1 + np.array([L, L, 0.0]), dens)
    sc.meshes = [{"nodes": nodes, "elnode": el, "zone": 10, "vel": (0.0, 0.0, -2.0),
                  "omega": (0.0, 0.0, 0.0), "center": (0.0, 0.0, 0.0)}]
    sc.meta.update({"plate": L, "dens": dens, "plate_z": zt})
    return sc


def gen_tool_mesh(center, zs, R_sh, R_pin, L_pin, n_theta=128, n_rs=20, n_zp=24, n_rc=6):
    """Friction-stir tool surface, axis z through `center` (x, y): the
    shoulder face (annulus R_pin..R_sh at z = zs), the pin side (radius R_pin,
    z from zs - L_pin to zs) and the pin tip (disk at zs - L_pin).  Triangle
    normals (n0n1 x n0n2) point out of the tool, into the workpiece: -z on the
    shoulder face and tip, radially outward on the pin side.  Every triangle
    has its own three nodes.  Returns nodes (3 ne, 3), elnode (ne, 3)."""
    cx, cy = center
    th = np.linspace(0.0, 2.0 * np.pi, n_theta + 1)[:-1]
    t0, t1 = th, np.roll(th, -1)

    def ring(r, t, z):
        return np.stack([cx + r * np.cos(t), cy + r * np.sin(t), np.full_like(t, z)], -1)

    tris = []
    # shoulder annulus and pin-tip rings: (p00, p01, p10), (p01, p11, p10),
    # i the inner radius, j the angle -> normal -z
    def annulus(radii, z):
        for ri, ro in zip(radii[:-1], radii[1:]):
            p00, p01 = ring(ri, t0, z), ring(ri, t1, z)
            p10, p11 = ring(ro, t0, z), ring(ro, t1, z)
            tris.append(np.stack([p00, p01, p10], 1))
            tris.append(np.stack([p01, p11, p10], 1))
    annulus(np.linspace(R_pin, R_sh, n_rs + 1), zs)
    zb = zs - L_pin
    # pin side: k the height -> normal e_r
    for z0, z1 in zip(np.linspace(zb, zs, n_zp + 1)[:-1], np.linspace(zb, zs, n_zp + 1)[1:]):
        p00, p01 = ring(R_pin, t0, z0), ring(R_pin, t1, z0)
        p10, p11 = ring(R_pin, t0, z1), ring(R_pin, t1, z1)
        tris.append(np.stack([p00, p01, p10], 1))
        tris.append(np.stack([p01, p11, p10], 1))
    # pin tip: a fan round the axis, then rings out to R_pin -> normal -z
    radii = np.linspace(0.0, R_pin, n_rc + 1)
    c = np.tile(np.array([cx, cy, zb]), (n_theta, 1))
    tris.append(np.stack([c, ring(radii[1], t1, zb), ring(radii[1], t0, zb)], 1))
    annulus(radii[1:], zb)
    T = np.concatenate(tris, 0)
    nodes = T.reshape(-1, 3).copy()
    el = np.arange(len(nodes), dtype=np.int32).reshape(-1, 3)
    return nodes, el


def c5_friction_stir(scale=1.0, r=0.125e-3, n_theta=128, n_rs=20, n_zp=24, n_rc=6, mu=0.2,
                     skin=0.4, ep_hot=0.05, thick=1.0, offset=(0.0, 0.0)):
    """Friction-stir-like plate: 120 x 60 x 6 mm of Johnson-Cook aluminium,
    dx = 2r, and a rigid tool (shoulder R 9 mm, pin R 3 mm, pin 5 mm long,
    ~12k triangles at the default resolution; `scale` multiplies the plate
    and tool dimensions, `thick` the plate thickness) with its shoulder on the top face and its pin
    plunged, centred at 1/4 of the length and 1/2 of the width, turning at
    1000 rpm about its axis and travelling at 2 mm/s along x.  Particles
    inside the pin are left out.  Penalty contact 0.1 E/dx with Coulomb
    friction mu, all of the sliding-friction power as heat into the plate,
    Taylor-Quinney 0.9, bottom layer held at 300 K and the bottom two layers
    at v = 0; lists with a skin of `skin` dx rebuilt when a particle with
    eps_p > ep_hot has moved half the skin (on the h = 1.2 dx lattice any skin
    from 0.05 to 0.43 dx lists the same 80 neighbours, since the next lattice
    shell lies at sqrt(8) dx, so the widest such skin rebuilds least often).
    `offset` shifts the tool axis by
    that many spacings in x and y; unequal shifts keep lattice points off the
    mesh's radial edges (a point on an edge makes the inside-triangle test a
    matter of rounding)."""
    rho, E, nu, T0 = 2700.0, 68.9e9, 0.3, 300.0
    Lx, Ly, Lz = 120e-3 * scale, 60e-3 * scale, 6e-3 * scale * thick
    R_sh, R_pin, L_pin = 9e-3 * scale, 3e-3 * scale, 5e-3 * scale
    xyz, mass = gen_box([0.0, 0.0, 0.0], [Lx, Ly, Lz], r, rho)
    dx = 2.0 * r
    ztop = xyz[:, 2].max()
    cx, cy = 0.25 * Lx + offset[0] * dx, 0.5 * Ly + offset[1] * dx
    rr = np.hypot(xyz[:, 0] - cx, xyz[:, 1] - cy)
    xyz = xyz[~((rr < R_pin) & (xyz[:, 2] > ztop - L_pin))]
    n = len(xyz)
    h = 1.2 * dx
    K = E / (3. * (1. - 2 * nu))
    G = E / (2. * (1. + nu))
    Cs = np.sqrt(K / rho)
    cfg = _abi.default_config()
    cfg.n_particles = n
    cfg.solid_count = n
    cfg.h_glob = h
    cfg.dt = 0.4 * h / (Cs + 10.0)
    mp = cfg.mat
    mp.E, mp.nu = E, nu
    for k, v in JC_AL.items():
        setattr(mp, k, v)
    cfg.return_map = _abi.WF_RETURN_NEWTON
    _adaptive_training(cfg)
    th = min(max((T0 - mp.Tt) / (mp.Tm - mp.Tt), 0.0), 1.0)
    sy = mp.A * (1.0 + mp.C * np.log(1e-5 / mp.eps0)) * (1.0 - th ** mp.m)
    f = _mech_fields(xyz, mass, rho, h, Cs, G, sy)
    big = 10 * Lx
    assign_zone(xyz, f["ID"], [-big, -big, -big], [big, big, xyz[:, 2].min() + 1.5 * dx], 2)
    bcs = [(2, (0.0, 0.0, 0.0))]
    cfg.thermal = 1
    cfg.taylor_quinney = 0.9
    f["T"] = np.full(n, T0)
    f["k_T"] = np.full(n, 200.0)
    f["cp_T"] = np.full(n, 900.0)
    bottom = np.isclose(xyz[:, 2], xyz[:, 2].min())
    f["BC_T"] = np.where(bottom, _abi.WF_BCT_FIXED, _abi.WF_BCT_NONE).astype(np.int32)
    cfg.contact = 1
    cfg.contact_law = _abi.WF_CONTACT_STIFFNESS
    cfg.contact_kappa = 0.1
    cfg.contact_dx = dx
    cfg.fric_sta = mu
    cfg.fric_dyn = mu
    cfg.fric_heat = 1.0
    cfg.rebuild_policy = _abi.WF_REBUILD_PLASTIC_ZONE
    cfg.ns_skin = skin * dx
    cfg.ep_hot = ep_hot
    zs = ztop + r  # the shoulder face half a spacing above the top layer
    nodes, el = gen_tool_mesh((cx, cy), zs, R_sh, R_pin, L_pin, n_theta, n_rs, n_zp, n_rc)
    omega = 1000.0 * 2.0 * np.pi / 60.0
    sc = Scenario("c5_friction_stir_plate", cfg, f, bcs,
                  {"r": r, "dx": dx, "h": h, "Cs": Cs, "mass": mass, "tool_center": (cx, cy),
                   "shoulder_z": zs, "R_sh": R_sh, "R_pin": R_pin, "L_pin": L_pin})
    sc.meshes = [{"nodes": nodes, "elnode": el, "zone": 10, "vel": (2e-3, 0.0, 0.0),
                  "omega": (0.0, 0.0, omega), "center": (cx, cy, zs)}]
    return sc


def c2_thermal_cube(L=46e-3, r=0.5e-3, seed=42):
    """Conduction only: cube of edge L, dx = 2r, rho 2700, cp 900, k 200,
    T = 300 + 500 exp(-|x-c|^2/(2 sigma^2)) + U(-1,1), sigma 5 mm; the first
    and last lattice layer on each axis held at 300 K; dt = 0.3 h^2 rho cp/k
    (Domain_d.cu:509)."""
    rho, cp, k = 2700.0, 900.0, 200.0
    xyz, mass = gen_box([0.0, 0.0, 0.0], [L, L, L], r, rho)
    n = len(xyz)
    dx = 2.0 * r
    h = 1.2 * dx
    cfg = _abi.default_config()
    cfg.n_particles = n
    cfg.solid_count = n
    cfg.h_glob = h
    cfg.mechanics = 0
    cfg.thermal = 1
    cfg.dt = 0.3 * h * h * rho * cp / k
    c = np.array([L / 2, L / 2, L / 2])
    sig = 5e-3
    rng = np.random.default_rng(seed)
    d2 = ((xyz - c) ** 2).sum(1)
    T = 300.0 + 500.0 * np.exp(-d2 / (2 * sig * sig)) + rng.uniform(-1.0, 1.0, n)
    lo, hi = xyz.min(0), xyz.max(0)
    face = np.zeros(n, bool)
    for a in range(3):
        face |= np.isclose(xyz[:, a], lo[a]) | np.isclose(xyz[:, a], hi[a])
    T[face] = 300.0
    bct = np.where(face, _abi.WF_BCT_FIXED, _abi.WF_BCT_NONE).astype(np.int32)
    f = {"x": xyz.copy(), "rho": np.full(n, rho), "rho0": np.full(n, rho),
         "m": np.full(n, mass), "h": np.full(n, h), "T": T, "k_T": np.full(n, k),
         "cp_T": np.full(n, cp), "BC_T": bct, "ID": np.zeros(n, np.int32)}
    return Scenario("c2_thermal_cube", cfg, f, [], {"L": L, "r": r, "dx": dx, "h": h})


def load_domain(dom, sc):
    """Upload a scenario's fields and BCs into a Domain."""
    for k, v in sc.fields.items():
        dom.set(k, v)
    dom.set_bcs(sc.bcs)
    for m in sc.meshes:
        mi = dom.add_mesh(m["nodes"], m["elnode"], m.get("zone", 0))
        dom.set_mesh_motion(mi, m.get("vel", (0, 0, 0)), m.get("omega", (0, 0, 0)),
                            m.get("center", (0, 0, 0)))
