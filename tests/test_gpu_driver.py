"""The host driver end to end on the GPU: deck → domain → Fraser loop → CSV
snapshots, against the FP64 oracle run of the same setup built here from the
deck's numbers (a rigid body of type File is read here by an independent
parse of the Nastran fixed fields).  CSV values carry 7 significant digits (%.6e, rounding <=
5e-7 relative), so each field is compared at relative L2 <= 1e-6 and
pointwise to 1e-5 of the field's largest magnitude."""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios
from weldformgpu_amd import build as wbuild

pytestmark = pytest.mark.gpu

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")
HEADER = ("ID, X, Y, Z, posX,posY,posZ,Ux, Uy, Uz, Vx, Vy, Vz, Ax, Ay, Az, rho, p, SigmaEq, "
          "Pl_Strain, Nb, ContNb, CFx, CFy, CFz, Nx, Ny, Nz, h")


def _read_csv(path):
    with open(path) as f:
        assert f.readline().rstrip("\n") == HEADER
    return np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)


def _nas_real(t):
    """An 8-character Nastran real: "1.5-3", "2.5E-01", "-1.00D0", "0.0"."""
    t = t.strip()
    m = re.fullmatch(r"([+-]?[0-9.]+)(?:[EeDd]?([+-][0-9]+))?", t)
    assert m, t
    return float(m.group(1)) * 10.0 ** int(m.group(2) or 0)


def _read_nas(path):
    """GRID (fields 2, 4-6) and CTRIA3 (fields 4-6) of a small-field file."""
    grid, tri = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("GRID"):
                grid.append((int(line[8:16]),
                             [_nas_real(line[24 + 8 * c:32 + 8 * c]) for c in range(3)]))
            elif line.startswith("CTRIA3"):
                tri.append([int(line[24 + 8 * c:32 + 8 * c]) for c in range(3)])
    pos = {gid: k for k, (gid, _) in enumerate(grid)}
    return (np.array([g[1] for g in grid], float),
            np.array([[pos[t] for t in row] for row in tri], np.int32))


def _oracle_from_deck(po, dk):
    """The setup WeldFormGPU.cu builds from the deck, restated in Python."""
    c, m, blk = dk["Configuration"], dk["Materials"][0], dk["DomainBlocks"][0]
    r, rho, E, nu = c["particleRadius"], m["density0"], m["youngsModulus"], m["poissonsRatio"]
    h = 2 * r * c["hFactor"]
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    Cs = math.sqrt(K / rho)
    axisymm = 0
    ID = None
    if blk["type"] == "Box":
        xyz, mass = scenarios.gen_box(blk["start"], blk["dim"], r, rho)
    elif blk.get("gridCoordSys", "Cartesian") == "Cartesian":
        xyz, mass = scenarios.gen_cylinder(blk["start"], blk["dim"][0] / 2, blk["dim"][2], r,
                                           rho)
    else:  # WeldFormGPU.cu:366-375, sliceAngle default 2.0000001 pi (:311)
        L = blk["dim"]
        if blk.get("sliceAngle", 2.0000001 * math.pi) == 2.0 * math.pi:
            xyz, ID, mass = po.gen_cylinder_polar(L[0] / 2, L[2], r, rho)
        else:
            xyz, ID, mass = po.gen_cylinder_polar(L[0] / 2, L[2], r, rho, math.pi / 16, 1,
                                                  L[0] / 4)
            axisymm = 1
    n = len(xyz)
    if ID is None:
        ID = np.zeros(n, np.int32)
    for z in dk["DomainZones"]:
        scenarios.assign_zone(xyz, ID, z["start"], z["end"], z["id"])
    cfg = _abi.default_config()
    cfg.n_particles = cfg.solid_count = n
    cfg.axisymm = axisymm
    cfg.h_glob = h
    cfg.dt = 0.4 * h / (Cs + 10.0)
    cfg.mat.E, cfg.mat.nu, cfg.mat.sy0 = E, nu, m["yieldStress0"]
    ext = dk.get("Extensions", {})
    T0 = 0.0
    if m["type"] == "Bilinear":
        cfg.mat.model, cfg.mat.Et = _abi.WF_MAT_BILINEAR, m["const"][0]
    else:
        cfg.mat.model = _abi.WF_MAT_JOHNSON_COOK
        (cfg.mat.A, cfg.mat.B, cfg.mat.n, cfg.mat.C, cfg.mat.eps0, cfg.mat.m, cfg.mat.Tm,
         cfg.mat.Tt) = m["const"]
        T0 = ext.get("thermal", {}).get("T0", cfg.mat.Tt)
        th = min(max((T0 - cfg.mat.Tt) / (cfg.mat.Tm - cfg.mat.Tt), 0.0), 1.0)
        cfg.mat.sy0 = cfg.mat.A * (1 + cfg.mat.C * math.log(1e-5 / cfg.mat.eps0)) * (
            1 - th ** cfg.mat.m)
    if ext.get("returnMap") == "newton":
        cfg.return_map = _abi.WF_RETURN_NEWTON
    st = po.State(n)
    st["x"] = xyz
    st["xs"] = xyz
    for k, v in {"rho": rho, "rho0": rho, "h": h, "m": mass, "Cs": Cs, "G": G,
                 "sigma_y": cfg.mat.sy0, "T": T0}.items():
        st[k] = np.full(n, v)
    if "thermal" in ext:
        t = ext["thermal"]
        cfg.thermal, cfg.taylor_quinney = 1, t["taylorQuinney"]
        st["k_T"] = np.full(n, t["k"])
        st["cp_T"] = np.full(n, t["cp"])
    st["ID"] = ID
    bcs = [(b["zoneId"], tuple(b["value"])) for b in dk["BoundaryConditions"]]
    meshes = []
    for rb in dk.get("RigidBodies", []):
        if rb["type"] == "File":
            nodes, el = _read_nas(os.path.join(DECKS, rb["fileName"]))
            nodes = nodes + np.array(rb["moveDir"], float)
        else:
            p1 = np.array(rb["start"], float)
            p2 = p1 + np.array([rb["dim"][0], rb["dim"][1], 0.0])
            nodes, el = po.gen_plane_mesh(p1, p2, rb["partSide"])
        vel = [b[1] for b in bcs if b[0] == rb["zoneId"]][0]
        meshes.append({"nodes": nodes, "elnode": el, "vel": vel})
    if meshes:
        cfg.contact = 1
        ct = dk["Contact"][0]
        cfg.fric_sta, cfg.fric_dyn = ct["fricCoeffStatic"], ct["fricCoeffDynamic"]
        st.add_contact(meshes)
    return st, cfg, bcs


@pytest.mark.parametrize("deck", ["compression_mini.json", "contact_mini.json",
                                  "contact_file_press.json", "cylindrical_slice_mini.json",
                                  "cylindrical_full_mini.json"])
def test_driver_run_matches_oracle(oracle, tmp_path, deck):
    path = os.path.join(DECKS, deck)
    with open(path) as f:
        dk = json.load(f)
    print(f"[case] driver {deck}", flush=True)
    r = subprocess.run([wbuild.DRIVER, path, "--out", str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    st, cfg, bcs = _oracle_from_deck(oracle, dk)
    st.nsearch(cfg)
    # the loop: steps until Time >= simTime with Time accumulated step by step,
    # a snapshot file whenever Time reaches the next output time, named by
    # Time in seconds to six decimals (SolverFraser.cu:144,330), so outputs
    # less than a microsecond apart share one name
    t, steps, t_out = 0.0, 0, dk["Configuration"]["outTime"]
    names, n_out = {"out_0.000000.csv"}, 0
    while t < dk["Configuration"]["simTime"]:
        t += cfg.dt
        steps += 1
        if t >= t_out:
            names.add(f"out_{t:.6f}.csv")
            n_out += 1
            t_out += dk["Configuration"]["outTime"]
    st.step(cfg, bcs, steps)
    outs = sorted(p for p in os.listdir(tmp_path) if p.startswith("out_"))
    assert n_out == int(round(dk["Configuration"]["simTime"] / dk["Configuration"]["outTime"]))
    assert outs == sorted(names)
    got = _read_csv(tmp_path / "test.csv")
    n = st.n
    ne = sum(len(m) for m in ([st.mesh_arrays["el"]] if st.st.contact else []))
    assert got.shape[0] == n + ne
    P = got[:n]

    def close(col, ref, name):
        got_f = P[:, col].ravel()
        ref = np.asarray(ref, float).ravel()
        scale = max(np.abs(ref).max(), 1e-300)
        l2 = np.linalg.norm(got_f - ref) / max(np.linalg.norm(ref), 1e-300)
        assert l2 < 1e-6, (name, l2)                      # %.6e rounding: <= 5e-7
        assert np.abs(got_f - ref).max() < 1e-5 * scale, name

    assert np.array_equal(P[:, 0].astype(int), st["ID"])
    close(slice(1, 4), st["x"], "x")
    close(slice(7, 10), st["u"], "u")
    close(slice(10, 13), st["v"], "v")
    close(18, st["sigma_eq"], "sigma_eq")
    close(19, st["pl_strain"], "pl_strain")
    assert np.array_equal(P[:, 20].astype(int), st["nb_count"])
    if st.st.contact:
        close(slice(22, 25), st.contact_arrays["cf"], "contforce")
        sums = [float(line.rsplit(" ", 1)[1]) for line in r.stdout.splitlines()
                if "contact force sum" in line]
        assert len(sums) == n_out and max(sums) > 0.0  # the plate touched
        R = got[n:]
        assert np.all(R[:, 0] == dk["RigidBodies"][0]["zoneId"])
        cen = st.mesh_arrays["c"]
        assert np.allclose(R[:, 1:4], cen, rtol=2e-6, atol=1e-12)
