"""Generates the small golden fixtures under tests/golden/ from the FP64 oracle.

  lattice.npz   a cubic block (bilinear steel, seeded random v and stress):
                inputs, the outputs of each step kernel applied in step order,
                and the fields after 5 full steps with two BC zones
  stress.npz    single particles through the reference radial return
                (bilinear) and the Newton return map (Johnson-Cook)
  contact.npz   three particles under a 2-triangle plate: inside one
                triangle, on the shared edge, outside (force and friction)
  mlp.npz       16 samples through the surrogate network: outputs, the
                mean-squared-error loss and its gradient

Run from the repository root: python tests/golden/make_fixtures.py
The oracle is deterministic, so tests/test_golden.py re-derives every
output bit for bit."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from weldformgpu_amd import _abi  # noqa: E402


def lattice_case():
    """Inputs of the lattice fixture (shared with the test)."""
    rho, E, nu, r = 7850.0, 200e9, 0.3, 0.5e-3
    xyz, mass = po.gen_box([0.0, 0.0, 0.0], [8.5e-3] * 3, r, rho)
    n = len(xyz)
    dx = 2 * r
    h = 1.2 * dx
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    Cs = np.sqrt(K / rho)
    cfg = _abi.default_config()
    cfg.n_particles = cfg.solid_count = n
    cfg.h_glob = h
    cfg.dt = 0.4 * h / (Cs + 10.0)
    cfg.mat.model, cfg.mat.E, cfg.mat.nu = _abi.WF_MAT_BILINEAR, E, nu
    cfg.mat.sy0, cfg.mat.Et = 300e6, 1e9
    rng = np.random.default_rng(2024)
    ss = rng.uniform(-250e6, 250e6, (n, 6))
    ss[:, 2] = -(ss[:, 0] + ss[:, 1])
    p = rng.uniform(-100e6, 100e6, n)
    sig = ss.copy()
    sig[:, :3] -= p[:, None]
    ID = np.zeros(n, np.int32)
    po.assign_zone(xyz, ID, [-1, -1, -1], [1, 1, 0.6e-3], 2)
    po.assign_zone(xyz, ID, [-1, -1, xyz[:, 2].max() - 1.1e-3], [1, 1, 1], 3)
    fields = {"x": xyz, "v": rng.uniform(-1.0, 1.0, (n, 3)), "rho": np.full(n, rho) *
              (1 + rng.uniform(-1e-3, 1e-3, n)), "rho0": np.full(n, rho), "m": np.full(n, mass),
              "h": np.full(n, h), "Cs": np.full(n, Cs), "G": np.full(n, G),
              "sigma_y": np.full(n, 300e6), "shearstress": ss, "sigma": sig, "ID": ID}
    bcs = [(2, (0.0, 0.0, 0.0)), (3, (0.0, 0.0, -1.0))]
    return cfg, fields, bcs


def _state(cfg, fields):
    st = po.State(cfg.n_particles)
    for k, v in fields.items():
        st[k] = v
    st["xs"] = fields["x"]
    return st


def lattice_outputs(cfg, fields, bcs):
    st = _state(cfg, fields)
    st.nsearch(cfg)
    out = {}
    st.calc_dens_inc(cfg)
    out["drho"] = st["drho"].copy()
    st.update_density(cfg.dt)
    out["rho"] = st["rho"].copy()
    st.calc_rate_tensors(cfg)
    out["strrate"], out["rotrate"] = st["strrate"].copy(), st["rotrate"].copy()
    st.pressure()
    out["p"] = st["p"].copy()
    st.stress_strain(cfg)
    for k in ("sigma", "shearstress", "pl_strain", "sigma_eq", "sigma_y", "int_energy"):
        out[k] = st[k].copy()
    st.calc_accel(cfg)
    out["a"] = st["a"].copy()
    st5 = _state(cfg, fields)
    st5.nsearch(cfg)
    st5.step(cfg, bcs, 5)
    for k in ("u", "v", "sigma", "rho", "pl_strain"):
        out["step5_" + k] = st5[k].copy()
    return out


def stress_case():
    """Single particles: bilinear (reference return) and JC (Newton)."""
    rows = []
    rng = np.random.default_rng(7)
    for model in (_abi.WF_MAT_BILINEAR, _abi.WF_MAT_JOHNSON_COOK):
        for k in range(4):
            s_scale = 150e6 if model == _abi.WF_MAT_BILINEAR else 400e6
            rows.append((model, rng.uniform(-2e3, 2e3, 6), rng.uniform(-500, 500, 3),
                         rng.uniform(-s_scale, s_scale, 6), 0.02 * k))
    return rows


def stress_outputs(rows):
    outs = {k: [] for k in ("sigma", "shearstress", "pl_strain", "sigma_eq", "sigma_y")}
    for model, D, W3, S0, ep0 in rows:
        cfg = _abi.default_config()
        cfg.n_particles = cfg.solid_count = 1
        cfg.dt = 1e-7
        m = cfg.mat
        m.model, m.E, m.nu = model, 70e9, 0.3
        if model == _abi.WF_MAT_BILINEAR:
            m.sy0, m.Et = 200e6, 2e9
            sy = 200e6
        else:
            cfg.return_map = _abi.WF_RETURN_NEWTON
            m.A, m.B, m.n, m.C, m.eps0, m.m, m.Tm, m.Tt = 324e6, 114e6, 0.42, 0.002, 1.0, 1.34, \
                925.0, 300.0
            sy = 324e6
        st = po.State(1)
        S0 = S0.copy()
        S0[2] = -(S0[0] + S0[1])
        st["shearstress"] = [S0]
        st["strrate"] = [D]
        st["rotrate"] = [[0, 0, 0, W3[0], W3[1], W3[2]]]
        st["G"] = [26.5e9]
        st["rho"] = [2700.0]
        st["m"] = [1e-6]
        st["sigma_y"] = [sy]
        st["pl_strain"] = [ep0]
        st["T"] = [350.0]
        st.stress_strain(cfg)
        for k in outs:
            outs[k].append(st[k][0].copy())
    return {k: np.array(v) for k, v in outs.items()}


def contact_case():
    nodes, el = po.gen_plane_mesh([0.0, 0.0, 1.0], [2.0, 2.0, 1.0], 2)
    x = np.array([[0.2, 0.3, 0.9], [0.5, 0.5, 0.85], [2.3, 0.5, 0.9]])
    v = np.array([[1.0, 0.0, 0.0], [12.0, -3.0, 0.0], [0.0, 0.0, 1.0]])
    return nodes, el, x, v


def contact_outputs(nodes, el, x, v):
    cfg = _abi.default_config()
    cfg.n_particles = cfg.solid_count = len(x)
    cfg.dt, cfg.contact, cfg.h_glob = 0.01, 1, 0.3
    cfg.fric_sta, cfg.fric_dyn = 0.5, 0.3
    st = po.State(len(x))
    st["x"] = x
    st["v"] = v
    st["m"] = np.full(len(x), 2.0)
    st["h"] = np.full(len(x), 0.3)
    st["ID"] = np.ones(len(x), np.int32)
    st.add_contact([{"nodes": nodes, "elnode": el}])
    st.contact_nb(cfg)
    st.contact_forces(cfg)
    return {"a": st["a"].copy(), "cf": st.contact_arrays["cf"].copy(),
            "cnb_count": st.contact_arrays["cnb_count"].copy()}


def mlp_case():
    rng = np.random.default_rng(16)
    return (rng.normal(0, 0.3, 1249), rng.uniform(0, 1.5, (16, 4)), rng.uniform(0, 1, 16))


def mlp_outputs(w, X, Y):
    y = np.array([po.mlp_forward(w, x) for x in X])
    g, loss = po.mlp_loss_grad(w, X, Y)
    return {"y": y, "grad": g, "loss": np.array(loss)}


def main():
    cfg, fields, bcs = lattice_case()
    np.savez(os.path.join(HERE, "lattice.npz"),
             **{"in_" + k: v for k, v in fields.items()},
             **{"out_" + k: v for k, v in lattice_outputs(cfg, fields, bcs).items()})
    rows = stress_case()
    np.savez(os.path.join(HERE, "stress.npz"),
             in_model=np.array([r[0] for r in rows]), in_D=np.array([r[1] for r in rows]),
             in_W=np.array([r[2] for r in rows]), in_S0=np.array([r[3] for r in rows]),
             in_ep0=np.array([r[4] for r in rows]),
             **{"out_" + k: v for k, v in stress_outputs(rows).items()})
    nodes, el, x, v = contact_case()
    np.savez(os.path.join(HERE, "contact.npz"), in_nodes=nodes, in_el=el, in_x=x, in_v=v,
             **{"out_" + k: val for k, val in contact_outputs(nodes, el, x, v).items()})
    w, X, Y = mlp_case()
    np.savez(os.path.join(HERE, "mlp.npz"), in_w=w, in_X=X, in_Y=Y,
             **{"out_" + k: val for k, val in mlp_outputs(w, X, Y).items()})


if __name__ == "__main__":
    main()
