"""ctypes mirror of include/wf_abi.h (types, constants, library loader).

Only declarations live here; `load()` opens the built HIP library
`weldformgpu_amd/libwfhip.so` and raises if it is missing, so nothing in the
product can silently run without the native code.
"""
import ctypes as C
import os

WF_ABI_VERSION = 7

WF_OK = 0
WF_E_ARG = -1
WF_E_HIP = -2
WF_E_STATE = -3
WF_E_NBCAP = -4
WF_E_NONFINITE = -5
WF_E_NOMEM = -6
WF_E_UNSUPPORTED = -7

WF_MAX_NEIGHBORS = 70
WF_MAX_NEIGHBORS_SKIN = 96
WF_MAX_BCS = 16

WF_MAT_BILINEAR, WF_MAT_HOLLOMON, WF_MAT_JOHNSON_COOK = 0, 1, 2
WF_FP64, WF_MIXED = 0, 1
WF_RETURN_REFERENCE, WF_RETURN_NEWTON = 0, 1
WF_Q_ROT_XZ_SIGN = 0x1
WF_Q_TENSOR_PRODUCT = 0x2
WF_Q_STALE_NS_POS = 0x4
WF_Q_ALL = 0x7
WF_BCT_NONE, WF_BCT_CONVECTION, WF_BCT_FIXED = 0, 1, 2
WF_CONTACT_REFERENCE, WF_CONTACT_STIFFNESS = 0, 1
WF_REBUILD_EVERY_N, WF_REBUILD_PLASTIC_ZONE, WF_REBUILD_PLASTIC_ZONE_LOCAL = 0, 1, 2
WF_MAX_MESHES = 8
WF_MAX_CONT_NB = 20

# fields: name -> (id, components, numpy dtype)
FIELDS = {}
for _k, _n in enumerate(["x", "v", "a", "u", "va", "x_snapshot", "normal", "contforce"]):
    FIELDS[_n] = (_k, 3, "f8")
for _k, _n in enumerate(["rho", "rho0", "drho", "m", "h", "p", "Cs", "G", "sigma_y",
                         "pl_strain", "sigma_eq", "eff_strain_rate", "int_energy", "T",
                         "Ta", "Tb", "dTdt", "k_T", "cp_T", "q_plastic", "q_friction"]):
    FIELDS[_n] = (16 + _k, 1, "f8")
for _k, _n in enumerate(["sigma", "shearstress", "strain", "strrate", "rotrate"]):
    FIELDS[_n] = (48 + _k, 6, "f8")
for _k, _n in enumerate(["ID", "BC_T", "nb_count", "cont_nb_count"]):
    FIELDS[_n] = (64 + _k, 1, "i4")

KERNELS = ["dens", "rates", "stress", "accel", "integrate", "thermal", "nsearch", "mlp", "return",
           "contact", "cond_dens", "accel_integrate"]
WF_MLP_PARAMS = 1249


class wf_material_params(C.Structure):
    _fields_ = [("model", C.c_int), ("_pad0", C.c_int),
                ("E", C.c_double), ("nu", C.c_double), ("sy0", C.c_double),
                ("Et", C.c_double), ("K_h", C.c_double), ("m_h", C.c_double),
                ("A", C.c_double), ("B", C.c_double), ("n", C.c_double),
                ("C", C.c_double), ("eps0", C.c_double), ("m", C.c_double),
                ("Tm", C.c_double), ("Tt", C.c_double)]


class wf_config(C.Structure):
    _fields_ = [("abi_version", C.c_int), ("n_particles", C.c_int),
                ("solid_count", C.c_int), ("precision", C.c_int),
                ("dt", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("h_glob", C.c_double), ("mat", wf_material_params),
                ("return_map", C.c_int), ("quirks", C.c_int),
                ("mechanics", C.c_int), ("thermal", C.c_int),
                ("taylor_quinney", C.c_double), ("T_inf", C.c_double),
                ("h_conv", C.c_double), ("ns_interval", C.c_int),
                ("surrogate", C.c_int), ("ns_min_plastic", C.c_double),
                ("newton_max_iter", C.c_int), ("mlp_train_interval", C.c_int),
                ("newton_tol", C.c_double),
                ("contact", C.c_int), ("contact_law", C.c_int),
                ("contact_kappa", C.c_double), ("contact_dx", C.c_double),
                ("fric_sta", C.c_double), ("fric_dyn", C.c_double),
                ("id_free_surf", C.c_int), ("rebuild_policy", C.c_int),
                ("fric_heat", C.c_double), ("ns_skin", C.c_double), ("ep_hot", C.c_double),
                ("mlp_train_iters", C.c_int), ("mlp_interval_max", C.c_int),
                ("mlp_lr", C.c_double), ("mlp_refit_ratio", C.c_double),
                ("ns_global_every", C.c_int), ("axisymm", C.c_int),
                ("fuse_integrate", C.c_int), ("ns_compact_max", C.c_int)]


class wf_stats(C.Structure):
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("is_yielding", C.c_int),
                ("ns_builds", C.c_int), ("newton_iters", C.c_int64),
                ("newton_calls", C.c_int64), ("newton_iters_wave", C.c_int64),
                ("mlp_trainings", C.c_int), ("nb_over_cap", C.c_int), ("cont_over_cap", C.c_int),
                ("mlp_loss", C.c_double), ("mlp_samples", C.c_int64),
                ("mlp_interval", C.c_int), ("_pad0", C.c_int), ("mlp_start_loss", C.c_double),
                ("ns_local_builds", C.c_int), ("_pad1", C.c_int), ("ns_local_tiles", C.c_int64)]


def default_config():
    """Python twin of wf_config_default (the reference defaults)."""
    c = wf_config()
    c.abi_version = WF_ABI_VERSION
    c.precision = WF_FP64
    c.alpha, c.beta = 1.0, 0.0
    c.return_map = WF_RETURN_REFERENCE
    c.quirks = WF_Q_ALL
    c.mechanics, c.thermal = 1, 0
    c.taylor_quinney = 0.0
    c.T_inf, c.h_conv = 500.0, 100.0
    c.ns_interval = 5
    c.ns_min_plastic = 1e-6
    c.newton_max_iter = 50
    c.newton_tol = 1e-11
    c.surrogate = 0
    c.mlp_train_interval = 200
    c.mlp_train_iters = 250
    c.mlp_interval_max = 0
    c.mlp_lr = 2e-3
    c.mlp_refit_ratio = 2.0
    c.ns_global_every = 32
    c.fuse_integrate = 1
    c.contact = 0
    c.contact_law = WF_CONTACT_REFERENCE
    c.contact_kappa = 0.1
    c.id_free_surf = 1
    c.rebuild_policy = WF_REBUILD_EVERY_N
    c.fric_heat = 0.0
    c.ns_skin = 0.0
    c.ep_hot = 0.05
    return c


_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwfhip.so")


def _proto(lib):
    P, I, D, VP, SZ, I64 = C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_size_t, C.c_int64
    sig = {
        "wf_config_default": (None, [C.POINTER(wf_config)]),
        "wf_field_shape": (I, [I, C.POINTER(I), C.POINTER(I)]),
        "wf_create": (I, [C.POINTER(wf_config), C.POINTER(P)]),
        "wf_destroy": (I, [P]),
        "wf_last_error": (C.c_char_p, []),
        "wf_sync": (I, [P]),
        "wf_set_field": (I, [P, I, VP, SZ]),
        "wf_get_field": (I, [P, I, VP, SZ]),
        "wf_fill_field": (I, [P, I, D]),
        "wf_snapshot_positions": (I, [P]),
        "wf_set_bcs": (I, [P, I, VP, VP]),
        "wf_nsearch_build": (I, [P]),
        "wf_nsearch_stats": (I, [P, C.POINTER(I64), C.POINTER(I)]),
        "wf_get_neighbors": (I, [P, VP, VP, VP, I64]),
        "wf_calc_dens_inc": (I, [P]),
        "wf_update_density": (I, [P, D]),
        "wf_calc_rate_tensors": (I, [P]),
        "wf_pressure": (I, [P]),
        "wf_stress_strain": (I, [P]),
        "wf_calc_accel": (I, [P]),
        "wf_apply_axisymm_bc": (I, [P]),
        "wf_apply_bc_vel": (I, [P, I, D, D, D]),
        "wf_update_pos_fraser": (I, [P, D]),
        "wf_update_vel": (I, [P, D]),
        "wf_calc_temp_inc": (I, [P]),
        "wf_thermal_calcs": (I, [P, D]),
        "wf_int_energy": (I, [P, C.POINTER(D)]),
        "wf_step": (I, [P, I]),
        "wf_get_stats": (I, [P, C.POINTER(wf_stats)]),
        "wf_set_timing": (I, [P, I]),
        "wf_kernel_time": (I, [P, I, C.POINTER(D), C.POINTER(I64)]),
        "wf_time_pass": (I, [P, I, I, C.POINTER(D)]),
        "wf_mlp_get_params": (I, [P, VP, I]),
        "wf_mlp_set_params": (I, [P, VP, I]),
        "wf_mlp_reset_adam": (I, [P]),
        "wf_mlp_infer": (I, [P, VP, I, VP]),
        "wf_mlp_grad": (I, [P, VP, VP, I, VP, C.POINTER(C.c_float)]),
        "wf_mlp_fit": (I, [P, VP, VP, I, I, C.c_float, C.POINTER(C.c_float)]),
        "wf_mlp_train": (I, [P, I, C.c_float, C.POINTER(I64), C.POINTER(C.c_float)]),
        "wf_mlp_grad_noalloc": (I, [P, P, P, I, P, P, P]),
        "wf_gen_plane_mesh": (I, [VP, VP, I, VP, I, VP, I, C.POINTER(I)]),
        "wf_read_nastran": (I, [C.c_char_p, I, VP, I, VP, I, C.POINTER(I)]),
        "wf_add_mesh": (I, [P, VP, I, VP, I, I, C.POINTER(I)]),
        "wf_set_mesh_motion": (I, [P, I, VP, VP, VP]),
        "wf_mesh_info": (I, [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
        "wf_get_mesh": (I, [P, I, VP, VP, VP, VP, VP]),
        "wf_calc_surface": (I, [P]),
        "wf_calc_contact_nb": (I, [P]),
        "wf_calc_contact_forces": (I, [P]),
        "wf_mesh_update": (I, [P, D]),
        "wf_get_contact_neighbors": (I, [P, VP, VP]),
        "wf_contact_force_sum": (I, [P, C.POINTER(D)]),
        "wf_gen_box": (I, [VP, D, D, D, D, D, VP, I, VP]),
        "wf_gen_cylinder": (I, [VP, D, D, D, D, VP, I, VP]),
        "wf_gen_cylinder_polar": (I, [I, D, D, D, D, D, I, D, VP, VP, I, VP]),
        "wf_assign_zone": (I, [VP, I, VP, VP, I, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def declared_symbols():
    """Names of every function include/wf_abi.h declares."""
    return ["wf_config_default", "wf_field_shape", "wf_create", "wf_destroy", "wf_last_error",
            "wf_sync", "wf_set_field", "wf_get_field", "wf_fill_field", "wf_snapshot_positions",
            "wf_set_bcs", "wf_nsearch_build", "wf_nsearch_stats", "wf_get_neighbors",
            "wf_calc_dens_inc", "wf_update_density", "wf_calc_rate_tensors", "wf_pressure",
            "wf_stress_strain", "wf_calc_accel", "wf_apply_axisymm_bc", "wf_apply_bc_vel",
            "wf_update_pos_fraser", "wf_update_vel", "wf_calc_temp_inc", "wf_thermal_calcs",
            "wf_int_energy", "wf_step", "wf_get_stats", "wf_set_timing", "wf_kernel_time",
            "wf_time_pass", "wf_mlp_get_params", "wf_mlp_set_params", "wf_mlp_reset_adam",
            "wf_mlp_infer", "wf_mlp_grad", "wf_mlp_fit", "wf_mlp_train", "wf_mlp_grad_noalloc",
            "wf_gen_plane_mesh", "wf_read_nastran", "wf_add_mesh", "wf_set_mesh_motion",
            "wf_mesh_info", "wf_get_mesh", "wf_calc_surface", "wf_calc_contact_nb",
            "wf_calc_contact_forces", "wf_mesh_update", "wf_get_contact_neighbors",
            "wf_contact_force_sum", "wf_gen_box", "wf_gen_cylinder", "wf_gen_cylinder_polar",
            "wf_assign_zone"]


def load(path=None):
    """Open libwfhip.so (raises OSError if it has not been built)."""
    global _LIB
    if _LIB is None or path is not None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise OSError(f"libwfhip.so not built at {p}: run __graft_entry__.build()")
        _LIB = _proto(C.CDLL(p))
    return _LIB
