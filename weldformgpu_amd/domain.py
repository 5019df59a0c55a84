"""Python mirror of the device domain (the reference's Domain_d surface).

`Domain` wraps one `wf_domain` handle of libwfhip.so.  Method names follow
the reference's Domain_d entry points (src/cuda/Domain_d.cuh:101-426):
CalcDensInc, UpdateDensity, CalcRateTensors, Pressure, StressStrain,
CalcAccel, ApplyBCVel, UpdatePosFraser, UpdateVel, CalcTempInc,
ThermalCalcs, CalcIntEnergy and MechFraserSolve-style stepping.  Arrays are
numpy, in the original particle order.  Every failing call raises WfError
with the library's message; nothing here computes physics on the host.
"""
import ctypes as C

import numpy as np

from . import _abi


class WfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"wf error {code}: {msg}")
        self.code = code


def _check(lib, rc):
    if rc != 0:
        raise WfError(rc, lib.wf_last_error().decode(errors="replace"))
    return rc


class Domain:
    def __init__(self, cfg, lib=None):
        self.lib = lib or _abi.load()
        self.cfg = cfg
        self.n = cfg.n_particles
        h = C.c_void_p()
        _check(self.lib, self.lib.wf_create(C.byref(cfg), C.byref(h)))
        self.h = h

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if self.h:
            self.lib.wf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(self.lib, self.lib.wf_sync(self.h))

    # -- fields ------------------------------------------------------------
    def set(self, name, arr):
        fid, comps, dt = _abi.FIELDS[name]
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if a.size != self.n * comps:
            raise ValueError(f"{name}: expected {self.n * comps} values, got {a.size}")
        _check(self.lib, self.lib.wf_set_field(self.h, fid, a.ctypes.data, a.nbytes))

    def get(self, name):
        fid, comps, dt = _abi.FIELDS[name]
        out = np.empty(self.n * comps, dtype=dt)
        _check(self.lib, self.lib.wf_get_field(self.h, fid, out.ctypes.data, out.nbytes))
        return out.reshape(self.n, comps) if comps > 1 else out

    def fill(self, name, value):
        fid, _, _ = _abi.FIELDS[name]
        _check(self.lib, self.lib.wf_fill_field(self.h, fid, float(value)))

    def snapshot_positions(self):
        _check(self.lib, self.lib.wf_snapshot_positions(self.h))

    def set_bcs(self, bcs):
        zones = np.array([b[0] for b in bcs] or [0], np.int32)
        vals = np.array([list(b[1]) for b in bcs] or [[0.0, 0.0, 0.0]], np.float64)
        _check(self.lib, self.lib.wf_set_bcs(self.h, len(bcs), zones.ctypes.data,
                                             vals.ctypes.data))

    # -- neighbour search --------------------------------------------------
    def nsearch(self):
        _check(self.lib, self.lib.wf_nsearch_build(self.h))

    def nsearch_stats(self):
        tot, oc = C.c_int64(), C.c_int()
        _check(self.lib, self.lib.wf_nsearch_stats(self.h, C.byref(tot), C.byref(oc)))
        return tot.value, oc.value

    def neighbors(self):
        tot, _ = self.nsearch_stats()
        counts = np.zeros(self.n, np.int32)
        offs = np.zeros(self.n + 1, np.int32)
        lst = np.zeros(max(tot, 1), np.int32)
        _check(self.lib, self.lib.wf_get_neighbors(self.h, counts.ctypes.data, offs.ctypes.data,
                                                   lst.ctypes.data, lst.size))
        return counts, offs, lst[:tot]

    # -- reference entry points ---------------------------------------------
    def CalcDensInc(self):
        _check(self.lib, self.lib.wf_calc_dens_inc(self.h))

    def UpdateDensity(self, dt):
        _check(self.lib, self.lib.wf_update_density(self.h, dt))

    def CalcRateTensors(self):
        _check(self.lib, self.lib.wf_calc_rate_tensors(self.h))

    def Pressure(self):
        _check(self.lib, self.lib.wf_pressure(self.h))

    def StressStrain(self):
        _check(self.lib, self.lib.wf_stress_strain(self.h))

    def CalcAccel(self):
        _check(self.lib, self.lib.wf_calc_accel(self.h))

    def ApplyAxiSymmBC(self):
        _check(self.lib, self.lib.wf_apply_axisymm_bc(self.h))

    def ApplyBCVel(self, zone, v):
        _check(self.lib, self.lib.wf_apply_bc_vel(self.h, zone, v[0], v[1], v[2]))

    def UpdatePosFraser(self, dt):
        _check(self.lib, self.lib.wf_update_pos_fraser(self.h, dt))

    def UpdateVel(self, dt):
        _check(self.lib, self.lib.wf_update_vel(self.h, dt))

    def CalcTempInc(self):
        _check(self.lib, self.lib.wf_calc_temp_inc(self.h))

    def ThermalCalcs(self, dt):
        _check(self.lib, self.lib.wf_thermal_calcs(self.h, dt))

    def CalcIntEnergy(self):
        s = C.c_double()
        _check(self.lib, self.lib.wf_int_energy(self.h, C.byref(s)))
        return s.value

    def step(self, n=1):
        _check(self.lib, self.lib.wf_step(self.h, int(n)))

    # -- stats and timing ------------------------------------------------------
    def stats(self):
        s = _abi.wf_stats()
        _check(self.lib, self.lib.wf_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_ if not k.startswith("_")}

    def set_timing(self, on):
        _check(self.lib, self.lib.wf_set_timing(self.h, 1 if on else 0))

    def time_pass(self, name, reps=20):
        """Milliseconds per launch of the rates or accel pass, launched reps
        times back to back (no event or wait between launches)."""
        k = _abi.KERNELS.index(name)
        ms = C.c_double()
        _check(self.lib, self.lib.wf_time_pass(self.h, k, int(reps), C.byref(ms)))
        return ms.value

    def kernel_time(self, name):
        k = _abi.KERNELS.index(name)
        ms, n = C.c_double(), C.c_int64()
        _check(self.lib, self.lib.wf_kernel_time(self.h, k, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # -- return-map surrogate ------------------------------------------------
    def mlp_params(self):
        w = np.zeros(_abi.WF_MLP_PARAMS, np.float32)
        _check(self.lib, self.lib.wf_mlp_get_params(self.h, w.ctypes.data, w.size))
        return w

    def set_mlp_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(self.lib, self.lib.wf_mlp_set_params(self.h, w.ctypes.data, w.size))

    def mlp_reset_adam(self):
        _check(self.lib, self.lib.wf_mlp_reset_adam(self.h))

    def mlp_infer(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(len(X), np.float32)
        _check(self.lib, self.lib.wf_mlp_infer(self.h, X.ctypes.data, len(X), y.ctypes.data))
        return y

    def mlp_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(_abi.WF_MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(self.lib, self.lib.wf_mlp_grad(self.h, X.ctypes.data, Y.ctypes.data, len(X),
                                              g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def mlp_fit(self, X, Y, iters, lr):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        loss = C.c_float()
        _check(self.lib, self.lib.wf_mlp_fit(self.h, X.ctypes.data, Y.ctypes.data, len(X),
                                             int(iters), float(lr), C.byref(loss)))
        return loss.value

    def mlp_train(self, iters=250, lr=2e-3):
        ns, loss = C.c_int64(), C.c_float()
        _check(self.lib, self.lib.wf_mlp_train(self.h, int(iters), float(lr), C.byref(ns),
                                               C.byref(loss)))
        return ns.value, loss.value

    # -- rigid meshes and contact (Contact.cu, Geometry.cu, Mesh.cu) -----------
    def add_mesh(self, nodes, elnode, zone_id):
        """Upload a rigid triangle mesh; returns its index."""
        nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, 3)
        elnode = np.ascontiguousarray(elnode, np.int32).reshape(-1, 3)
        mi = C.c_int()
        _check(self.lib, self.lib.wf_add_mesh(self.h, nodes.ctypes.data, len(nodes),
                                              elnode.ctypes.data, len(elnode), int(zone_id),
                                              C.byref(mi)))
        return mi.value

    def set_mesh_motion(self, mesh, v=(0, 0, 0), w=(0, 0, 0), center=(0, 0, 0)):
        v, w, c = (np.ascontiguousarray(a, np.float64) for a in (v, w, center))
        _check(self.lib, self.lib.wf_set_mesh_motion(self.h, int(mesh), v.ctypes.data,
                                                     w.ctypes.data, c.ctypes.data))

    def mesh_info(self, mesh=-1):
        nn, ne, z = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib, self.lib.wf_mesh_info(self.h, int(mesh), C.byref(nn), C.byref(ne),
                                               C.byref(z)))
        return nn.value, ne.value, z.value

    def get_mesh(self, mesh):
        nn, ne, _ = self.mesh_info(mesh)
        out = {"nodes": np.zeros((nn, 3)), "node_v": np.zeros((nn, 3)),
               "centroids": np.zeros((ne, 3)), "normals": np.zeros((ne, 3)),
               "pplane": np.zeros(ne)}
        _check(self.lib, self.lib.wf_get_mesh(
            self.h, int(mesh), out["nodes"].ctypes.data, out["node_v"].ctypes.data,
            out["centroids"].ctypes.data, out["normals"].ctypes.data, out["pplane"].ctypes.data))
        return out

    def CalculateSurface(self):
        _check(self.lib, self.lib.wf_calc_surface(self.h))

    def CalcContactNb(self):
        _check(self.lib, self.lib.wf_calc_contact_nb(self.h))

    def CalcContactForces(self):
        _check(self.lib, self.lib.wf_calc_contact_forces(self.h))

    def MeshUpdate(self, dt):
        _check(self.lib, self.lib.wf_mesh_update(self.h, float(dt)))

    def contact_neighbors(self):
        """(counts[n], list[n, 20]) in original particle order, -1 padded."""
        cnt = np.zeros(self.n, np.int32)
        lst = np.zeros((self.n, _abi.WF_MAX_CONT_NB), np.int32)
        _check(self.lib, self.lib.wf_get_contact_neighbors(self.h, cnt.ctypes.data,
                                                           lst.ctypes.data))
        return cnt, lst

    def contact_force_sum(self):
        s = C.c_double()
        _check(self.lib, self.lib.wf_contact_force_sum(self.h, C.byref(s)))
        return s.value
