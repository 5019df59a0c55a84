"""The HIP path against the committed golden fixtures (no oracle run):
single kernels at 1e-12 relative L2, 5 full steps at 1e-9, the contact
case at 1e-12, the surrogate's FP32 gradient at 1e-4."""
import os
import sys

import numpy as np
import pytest

from weldformgpu_amd import _abi
from weldformgpu_amd.domain import Domain

from harness import rel_l2

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fixtures as mf  # noqa: E402


def _load(name):
    return np.load(os.path.join(HERE, "golden", name), allow_pickle=False)


def _domain(cfg, g, lib):
    d = Domain(cfg, lib)
    for k in ("x", "v", "rho", "rho0", "m", "h", "Cs", "G", "sigma_y", "shearstress", "sigma",
              "ID"):
        d.set(k, g["in_" + k])
    return d


def test_lattice_kernels(wflib):
    g = _load("lattice.npz")
    cfg, _, bcs = mf.lattice_case()
    print(f"[case] golden lattice n={cfg.n_particles}", flush=True)
    with _domain(cfg, g, wflib) as d:
        d.nsearch()
        d.CalcDensInc()
        assert rel_l2(d.get("drho"), g["out_drho"]) < 1e-12
        d.UpdateDensity(cfg.dt)
        d.CalcRateTensors()
        assert rel_l2(d.get("strrate"), g["out_strrate"]) < 1e-12
        assert rel_l2(d.get("rotrate"), g["out_rotrate"]) < 1e-12
        d.Pressure()
        d.StressStrain()
        for k in ("sigma", "pl_strain", "sigma_eq"):
            assert rel_l2(d.get(k), g["out_" + k]) < 1e-12, k
        d.CalcAccel()
        assert rel_l2(d.get("a"), g["out_a"]) < 1e-12
    with _domain(cfg, g, wflib) as d:
        d.set_bcs(bcs)
        d.step(5)
        for k in ("u", "v", "sigma", "rho", "pl_strain"):
            assert rel_l2(d.get(k), g["out_step5_" + k]) < 1e-9, k


def test_contact_case(wflib):
    g = _load("contact.npz")
    cfg = _abi.default_config()
    n = len(g["in_x"])
    cfg.n_particles = cfg.solid_count = n
    cfg.dt, cfg.contact, cfg.h_glob = 0.01, 1, 0.3
    cfg.fric_sta, cfg.fric_dyn = 0.5, 0.3
    print("[case] golden contact", flush=True)
    with Domain(cfg, wflib) as d:
        d.set("x", g["in_x"])
        d.set("v", g["in_v"])
        d.fill("m", 2.0)
        d.fill("h", 0.3)
        d.set("ID", np.ones(n, np.int32))
        d.add_mesh(g["in_nodes"], g["in_el"], 10)
        d.nsearch()
        d.CalcContactNb()
        d.CalcContactForces()
        cnt, _ = d.contact_neighbors()
        assert cnt.tolist() == g["out_cnb_count"].tolist()
        assert rel_l2(d.get("a"), g["out_a"]) < 1e-12
        assert rel_l2(d.get("contforce"), g["out_cf"]) < 1e-12


def test_mlp_gradient(wflib):
    g = _load("mlp.npz")
    cfg = _abi.default_config()
    cfg.n_particles = cfg.solid_count = 64
    print("[case] golden mlp", flush=True)
    with Domain(cfg, wflib) as d:
        d.set_mlp_params(g["in_w"].astype(np.float32))
        grad, loss = d.mlp_grad(g["in_X"].astype(np.float32), g["in_Y"].astype(np.float32))
    assert rel_l2(grad, g["out_grad"]) < 1e-4
    assert abs(loss - float(g["out_loss"])) < 1e-5 * max(1.0, float(g["out_loss"]))
