"""ctypes wrapper of oracle/liboracle.so — the FP64 CPU restatement.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from weldformgpu_amd._abi import wf_config, wf_material_params  # noqa: F401  (types only)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_DBL3 = ["x", "v", "a", "u", "va", "xs"]
_DBL1 = ["rho", "rho0", "drho", "m", "h", "p", "Cs", "G", "sigma_y", "pl_strain", "sigma_eq",
         "eff_strain_rate", "int_energy", "T", "Ta", "Tb", "dTdt", "k_T", "cp_T", "q_plastic",
         "q_friction"]
_DBL6 = ["sigma", "shearstress", "strain", "strrate", "rotrate"]
_INT1 = ["ID", "BC_T"]


class or_state(C.Structure):
    _fields_ = ([("n", C.c_int), ("solid", C.c_int)]
                + [(k, C.c_void_p) for k in _DBL3 + _DBL1 + _DBL6 + _INT1]
                + [("nb_count", C.c_void_p), ("nb_offset", C.c_void_p), ("nb_list", C.c_void_p),
                   ("nb_cap", C.c_int64), ("newton_iters", C.c_int64),
                   ("newton_calls", C.c_int64), ("contact", C.c_void_p), ("xb", C.c_void_p)])


class or_mesh(C.Structure):
    _fields_ = [("nn", C.c_int), ("ne", C.c_int), ("n_meshes", C.c_int)] + [
        (k, C.c_void_p) for k in ["x", "v", "node_mesh", "el", "c", "nrm", "pp", "motion"]]


class or_contact(C.Structure):
    _fields_ = [("mesh", C.c_void_p)] + [
        (k, C.c_void_p) for k in ["normal", "cf", "idlock", "cnb_count", "cnb_list"]] + [
        ("surf_init", C.c_int), ("cont_over_cap", C.c_int), ("totmass", C.c_double)]


class or_run(C.Structure):
    _fields_ = [("time", C.c_double), ("step", C.c_int64), ("ts_i", C.c_int),
                ("is_yielding", C.c_int), ("isfirst_step", C.c_int), ("ns_builds", C.c_int),
                ("max_pl_strain", C.c_double), ("nb_over_cap", C.c_int),
                ("rebuild_pending", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        S, CF, R = C.POINTER(or_state), C.POINTER(wf_config), C.POINTER(or_run)
        sig = {
            "or_gen_box": (I, [P, D, D, D, D, P, I, D, P]),
            "or_gen_cylinder": (I, [P, D, D, D, P, I, D, P]),
            "or_gen_cylinder_polar": (I, [I, D, D, D, D, D, I, D, P, P, I, P]),
            "or_half_part_count": (I, [D, D, I]),
            "or_assign_zone": (I, [P, I, P, P, I, P]),
            "or_neighbors": (I64, [P, I, D, P, P, P, I64, P]),
            "or_neighbors_bruteforce": (I64, [P, I, D, P, P, P, I64]),
            "or_kernel": (D, [D, D]),
            "or_grad_kernel": (D, [D, D]),
            "or_jc_yield": (D, [C.POINTER(wf_material_params), D, D, D]),
            "or_jc_hardening": (D, [C.POINTER(wf_material_params), D, D, D]),
            "or_hollomon_yield": (D, [C.POINTER(wf_material_params), D]),
            "or_jc_newton": (D, [C.POINTER(wf_material_params), D, D, D, D, D, D, D, I,
                                 C.POINTER(I)]),
            "or_calc_dens_inc": (None, [S, CF]),
            "or_update_density": (None, [S, D]),
            "or_calc_rate_tensors": (None, [S, CF]),
            "or_pressure": (None, [S]),
            "or_stress_strain": (None, [S, CF]),
            "or_calc_accel": (None, [S, CF]),
            "or_apply_axisymm_bc": (None, [S]),
            "or_apply_bc_vel": (None, [S, I, D, D, D]),
            "or_update_pos_fraser": (None, [S, D]),
            "or_update_vel": (None, [S, D]),
            "or_calc_temp_inc": (None, [S, CF]),
            "or_thermal_calcs": (None, [S, CF, C.POINTER(I)]),
            "or_int_energy": (None, [S, CF]),
            "or_int_energy_sum": (D, [S]),
            "or_nsearch_build": (I, [S, CF, C.POINTER(I)]),
            "or_plastic_zone_moved": (I, [S, CF]),
            "or_step": (I, [S, CF, R, I, P, P]),
            "or_gen_plane_mesh": (I, [P, P, I, P, P]),
            "or_mesh_geom": (None, [P]),
            "or_mesh_move": (None, [P, D]),
            "or_calc_surface": (None, [S, CF]),
            "or_contact_nb": (None, [S, CF]),
            "or_contact_forces": (None, [S, CF]),
            "or_mlp_forward": (D, [P, P]),
            "or_mlp_loss_grad": (D, [P, P, P, I, P]),
            "or_set_threads": (None, [I]),
            "or_get_threads": (I, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- geometry -------------------------------------------------------------
def gen_box(V, L, r, density=1.0):
    V = np.ascontiguousarray(V, dtype=np.float64)
    n = lib().or_gen_box(_ptr(V), L[0], L[1], L[2], r, None, 0, density, None)
    xyz = np.zeros((n, 3))
    mass = C.c_double(0.0)
    lib().or_gen_box(_ptr(V), L[0], L[1], L[2], r, _ptr(xyz), n, density, C.byref(mass))
    return xyz, mass.value


def gen_cylinder(V, R, Lz, r, density=1.0):
    V = np.ascontiguousarray(V, dtype=np.float64)
    n = lib().or_gen_cylinder(_ptr(V), R, Lz, r, None, 0, density, None)
    xyz = np.zeros((n, 3))
    mass = C.c_double(0.0)
    lib().or_gen_cylinder(_ptr(V), R, Lz, r, _ptr(xyz), n, density, C.byref(mass))
    return xyz, mass.value


def gen_cylinder_polar(R, Lz, r, density=1.0, ang=2.0 * np.pi, rows=1, r_i=0.0, tag=0):
    """Positions (n, 3), IDs (n,) and particle mass of the cylindrical grid."""
    n = lib().or_gen_cylinder_polar(tag, R, Lz, r, density, ang, rows, r_i, None, None, 0, None)
    xyz = np.zeros((n, 3))
    ID = np.zeros(n, dtype=np.int32)
    mass = C.c_double(0.0)
    lib().or_gen_cylinder_polar(tag, R, Lz, r, density, ang, rows, r_i, _ptr(xyz), _ptr(ID), n,
                                C.byref(mass))
    return xyz, ID, mass.value


def assign_zone(xyz, ID, start, end, zid):
    s = np.ascontiguousarray(start, dtype=np.float64)
    e = np.ascontiguousarray(end, dtype=np.float64)
    return lib().or_assign_zone(_ptr(xyz), len(xyz), _ptr(s), _ptr(e), zid, _ptr(ID))


def gen_plane_mesh(p1, p2, dens):
    """AxisPlaneMesh(2, false, p1, p2, dens): (nodes (nn,3), elnode (ne,3))."""
    p1 = np.ascontiguousarray(p1, np.float64)
    p2 = np.ascontiguousarray(p2, np.float64)
    nodes = np.zeros(((dens + 1) ** 2, 3))
    el = np.zeros((2 * dens * dens, 3), np.int32)
    lib().or_gen_plane_mesh(_ptr(p1), _ptr(p2), dens, _ptr(nodes), _ptr(el))
    return nodes, el


# ---- neighbours -----------------------------------------------------------
def neighbors(xyz, radius, brute=False):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    n = len(xyz)
    counts = np.zeros(n, np.int32)
    offs = np.zeros(n + 1, np.int32)
    cap = max(64 * n, 16)
    while True:
        lst = np.zeros(cap, np.int32)
        oc = C.c_int(0)
        if brute:
            tot = lib().or_neighbors_bruteforce(_ptr(xyz), n, radius, _ptr(counts), _ptr(offs),
                                                _ptr(lst), cap)
        else:
            tot = lib().or_neighbors(_ptr(xyz), n, radius, _ptr(counts), _ptr(offs), _ptr(lst),
                                     cap, C.byref(oc))
        if tot <= cap:
            return counts, offs, lst[:tot].copy(), oc.value
        cap = int(tot)


# ---- state ----------------------------------------------------------------
class State:
    """Owns numpy arrays for one oracle domain (original particle order)."""

    def __init__(self, n, solid=None, nb_cap=None):
        self.n = n
        self.arrays = {}
        for k in _DBL3:
            self.arrays[k] = np.zeros((n, 3))
        for k in _DBL1:
            self.arrays[k] = np.zeros(n)
        for k in _DBL6:
            self.arrays[k] = np.zeros((n, 6))
        for k in _INT1:
            self.arrays[k] = np.zeros(n, np.int32)
        self.nb_cap = nb_cap or 100 * n + 16  # above the 96-entry skin-list cap
        self.arrays["nb_count"] = np.zeros(n, np.int32)
        self.arrays["nb_offset"] = np.zeros(n + 1, np.int32)
        self.arrays["nb_list"] = np.zeros(self.nb_cap, np.int32)
        self.st = or_state()
        self.st.n = n
        self.st.solid = n if solid is None else solid
        for k, a in self.arrays.items():
            setattr(self.st, k, a.ctypes.data)
        self.st.nb_cap = self.nb_cap
        self.arrays["xb"] = np.zeros((n, 3))  # positions at the last build
        self.st.xb = self.arrays["xb"].ctypes.data
        self.run = or_run()
        self.run.isfirst_step = 1

    def __getitem__(self, k):
        return self.arrays[k]

    def __setitem__(self, k, v):
        self.arrays[k][...] = v

    def ref(self):
        return C.byref(self.st)

    def nsearch(self, cfg):
        oc = C.c_int(0)
        rc = lib().or_nsearch_build(self.ref(), C.byref(cfg), C.byref(oc))
        if rc != 0:
            raise RuntimeError(f"oracle nsearch failed {rc}")
        self.run.nb_over_cap = oc.value
        return oc.value

    def plastic_zone_moved(self, cfg):
        return bool(lib().or_plastic_zone_moved(self.ref(), C.byref(cfg)))

    def neighbor_rows(self):
        c, o, l = self["nb_count"], self["nb_offset"], self["nb_list"]
        return c.copy(), o.copy(), l[: o[self.n]].copy()

    # -- rigid meshes and contact -------------------------------------------
    def add_contact(self, meshes):
        """meshes: list of dicts with nodes (nn,3), elnode (ne,3) and optional
        vel, omega, center (3 each).  Concatenates them like the library."""
        xs, els, nm, mot = [], [], [], []
        off = 0
        for k, m in enumerate(meshes):
            nodes = np.asarray(m["nodes"], np.float64).reshape(-1, 3)
            xs.append(nodes)
            els.append(np.asarray(m["elnode"], np.int32).reshape(-1, 3) + off)
            nm.append(np.full(len(nodes), k, np.int32))
            mot.append(np.concatenate([np.asarray(m.get(q, (0, 0, 0)), np.float64)
                                       for q in ("vel", "omega", "center")]))
            off += len(nodes)
        nn, ne = off, sum(len(e) for e in els)
        cat = (lambda a, shape, dt: np.ascontiguousarray(np.concatenate(a)) if a
               else np.zeros(shape, dt))
        M = self.mesh_arrays = {
            "x": cat(xs, (0, 3), np.float64), "v": np.zeros((nn, 3)),
            "node_mesh": cat(nm, 0, np.int32), "el": cat(els, (0, 3), np.int32),
            "c": np.zeros((ne, 3)), "nrm": np.zeros((ne, 3)), "pp": np.zeros(ne),
            "motion": cat(mot, 0, np.float64)}
        self.mesh = or_mesh()
        self.mesh.nn, self.mesh.ne, self.mesh.n_meshes = nn, ne, len(meshes)
        for k, a in M.items():
            setattr(self.mesh, k, a.ctypes.data)
        n = self.n
        self.contact_arrays = {"normal": np.zeros((n, 3)), "cf": np.zeros((n, 3)),
                               "idlock": np.zeros(n, np.int32),
                               "cnb_count": np.zeros(n, np.int32),
                               "cnb_list": np.full((n, 20), -1, np.int32)}
        self.ct = or_contact()
        self.ct.mesh = C.cast(C.pointer(self.mesh), C.c_void_p)
        for k, a in self.contact_arrays.items():
            setattr(self.ct, k, a.ctypes.data)
        self.st.contact = C.cast(C.pointer(self.ct), C.c_void_p)
        lib().or_mesh_geom(C.byref(self.mesh))

    def set_mesh_motion(self, mesh, vel=(0, 0, 0), omega=(0, 0, 0), center=(0, 0, 0)):
        self.mesh_arrays["motion"][9 * mesh:9 * mesh + 9] = np.concatenate(
            [np.asarray(q, np.float64) for q in (vel, omega, center)])

    def calc_surface(self, cfg):
        lib().or_calc_surface(self.ref(), C.byref(cfg))

    def contact_nb(self, cfg):
        lib().or_contact_nb(self.ref(), C.byref(cfg))

    def contact_forces(self, cfg):
        lib().or_contact_forces(self.ref(), C.byref(cfg))

    def mesh_move(self, dt):
        lib().or_mesh_move(C.byref(self.mesh), dt)

    def step(self, cfg, bcs=(), nsteps=1):
        zones = np.array([b[0] for b in bcs] or [0], np.int32)
        vals = np.array([list(b[1]) for b in bcs] or [[0, 0, 0]], np.float64)
        for _ in range(nsteps):
            rc = lib().or_step(self.ref(), C.byref(cfg), C.byref(self.run), len(bcs),
                               _ptr(zones), _ptr(vals))
            if rc != 0:
                raise RuntimeError(f"oracle step failed {rc}")

    # single kernels
    def calc_dens_inc(self, cfg):
        lib().or_calc_dens_inc(self.ref(), C.byref(cfg))

    def update_density(self, dt):
        lib().or_update_density(self.ref(), dt)

    def calc_rate_tensors(self, cfg):
        lib().or_calc_rate_tensors(self.ref(), C.byref(cfg))

    def pressure(self):
        lib().or_pressure(self.ref())

    def stress_strain(self, cfg):
        lib().or_stress_strain(self.ref(), C.byref(cfg))

    def calc_accel(self, cfg):
        lib().or_calc_accel(self.ref(), C.byref(cfg))

    def apply_axisymm_bc(self):
        lib().or_apply_axisymm_bc(self.ref())

    def apply_bc_vel(self, zone, v):
        lib().or_apply_bc_vel(self.ref(), zone, v[0], v[1], v[2])

    def update_pos_fraser(self, dt):
        lib().or_update_pos_fraser(self.ref(), dt)

    def update_vel(self, dt):
        lib().or_update_vel(self.ref(), dt)

    def calc_temp_inc(self, cfg):
        lib().or_calc_temp_inc(self.ref(), C.byref(cfg))

    def thermal_calcs(self, cfg):
        f = C.c_int(self.run.isfirst_step)
        lib().or_thermal_calcs(self.ref(), C.byref(cfg), C.byref(f))
        self.run.isfirst_step = f.value

    def int_energy(self, cfg):
        lib().or_int_energy(self.ref(), C.byref(cfg))

    def int_energy_sum(self):
        return lib().or_int_energy_sum(self.ref())


def set_threads(n):
    """OpenMP threads of the oracle's particle loops."""
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def kernel(q, h):
    return lib().or_kernel(q, h)


def grad_kernel(q, h):
    return lib().or_grad_kernel(q, h)


def mlp_forward(w, x):
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_mlp_forward(_ptr(w), _ptr(x))


def mlp_loss_grad(w, X, Y):
    w = np.ascontiguousarray(w, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    g = np.zeros(1249)
    loss = lib().or_mlp_loss_grad(_ptr(w), _ptr(X), _ptr(Y), len(Y), _ptr(g))
    return g, loss
