"""The command-line host driver (weldformgpu_amd/wfdriver): JSON deck parsing
and setup on the host (--dry-run, no GPU), checked against the deck's own
numbers and the library's lattice/zone builders, and its error handling.
The decks under tests/golden/decks use the reference's schema
(WeldFormGPU.cu:142-906, App. C of the deck format), including a
BoundaryConditions entry spelled "valuetype" (read as constant)."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from weldformgpu_amd import build as wbuild
from weldformgpu_amd import scenarios

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _dry(deck, expect_rc=0):
    r = subprocess.run([wbuild.DRIVER, "--dry-run", deck], capture_output=True, text=True)
    assert r.returncode == expect_rc, r.stderr
    return json.loads(r.stdout) if expect_rc == 0 else r.stderr


def _deck(name):
    with open(os.path.join(DECKS, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module", autouse=True)
def _driver_built():
    if not os.path.exists(wbuild.DRIVER):
        wbuild.build_driver()


def test_compression_deck_setup():
    dk = _deck("compression_mini.json")
    s = _dry(os.path.join(DECKS, "compression_mini.json"))
    c, m, blk = dk["Configuration"], dk["Materials"][0], dk["DomainBlocks"][0]
    r, rho, E, nu = c["particleRadius"], m["density0"], m["youngsModulus"], m["poissonsRatio"]
    h = 2 * r * c["hFactor"]
    K = E / (3 * (1 - 2 * nu))
    Cs = math.sqrt(K / rho)
    assert s["h"] == h and s["K"] == K and s["G"] == E / (2 * (1 + nu)) and s["Cs"] == Cs
    assert s["dt"] == 0.4 * h / (Cs + 10.0)          # VMAX = 10
    xyz, mass = scenarios.gen_cylinder(blk["start"], blk["dim"][0] / 2, blk["dim"][2], r, rho)
    assert s["n"] == len(xyz) and s["mass"] == mass
    ID = np.zeros(len(xyz), np.int32)
    counts = [scenarios.assign_zone(xyz, ID, z["start"], z["end"], z["id"])
              for z in dk["DomainZones"]]
    assert [zc for _, zc in s["zones"]] == counts and min(counts) > 0
    assert s["bcs"] == [[2, 0, 0, 0], [3, 0, 0, -10]]
    assert s["model"] == 0 and s["sy0"] == m["yieldStress0"] and s["contact"] == 0
    assert s["sim_time"] == c["simTime"] and s["out_time"] == c["outTime"]


def test_contact_deck_setup():
    dk = _deck("contact_mini.json")
    s = _dry(os.path.join(DECKS, "contact_mini.json"))
    rb = dk["RigidBodies"][0]
    dens = rb["partSide"]
    assert s["contact"] == 1 and s["contact_law"] == 0
    assert s["fric_sta"] == 0.2 and s["fric_dyn"] == 0.2
    # plate: (dens+1)^2 nodes, 2 dens^2 triangles, moved by the zone-10 BC
    assert s["meshes"] == [[10, (dens + 1) ** 2, 2 * dens * dens, 0, 0, -2]]
    # Johnson-Cook: flow stress at eps_p = 0, rate 1e-5, T0 = T_t
    A, B, n, C, eps0, mm, Tm, Tt = dk["Materials"][0]["const"]
    assert s["model"] == 2 and s["return_map"] == 1 and s["thermal"] == 1
    assert s["sy0"] == pytest.approx(A * (1 + C * math.log(1e-5 / eps0)), rel=1e-15)
    assert s["T0"] == 300.0


def test_extension_options(tmp_path):
    """Extensions of this library in the deck (the reference ignores the
    object): sliding-friction heat, the plastic-zone rebuild policies and the
    surrogate's training settings."""
    d = _deck("contact_mini.json")
    s = _dry(os.path.join(DECKS, "contact_mini.json"))
    assert s["fric_heat"] == 0.0 and s["rebuild_policy"] == 0 and s["ep_hot"] == 0.05
    assert s["ns_global_every"] == 32 and s["mlp"] == [200, 250, 2e-3, 0, 2.0]
    d["Extensions"]["frictionHeat"] = 0.9
    d["Extensions"]["rebuild"] = {"policy": "plasticZone", "skin": 2e-5, "epHot": 0.1}
    p = tmp_path / "ext.json"
    p.write_text(json.dumps(d))
    s = _dry(str(p))
    assert s["fric_heat"] == 0.9 and s["rebuild_policy"] == 1
    assert s["ns_skin"] == 2e-5 and s["ep_hot"] == 0.1
    d["Extensions"]["rebuild"] = {"policy": "plasticZoneLocal", "skin": 2e-5, "epHot": 0.1,
                                  "globalEvery": 12}
    d["Extensions"]["surrogateTraining"] = {"interval": 100, "iterations": 80,
                                            "learningRate": 1e-3, "intervalMax": 1600,
                                            "refitRatio": 1.5}
    p.write_text(json.dumps(d))
    s = _dry(str(p))
    assert s["rebuild_policy"] == 2 and s["ns_global_every"] == 12
    assert s["mlp"] == [100, 80, 1e-3, 1600, 1.5]
    d["Extensions"]["rebuild"]["policy"] = "sometimes"
    p.write_text(json.dumps(d))
    assert "plasticZoneLocal" in _dry(str(p), expect_rc=2)


def test_rotating_rigid_body(tmp_path):
    """A rigid body turning about its own axis: the angular velocity comes
    from the valueAng of its zone's boundary condition (SetMeshVel) and the
    centre from RigidBodies[].rotationCenter (an extension; the reference
    turns bodies about the origin)."""
    d = _deck("contact_mini.json")
    s = _dry(os.path.join(DECKS, "contact_mini.json"))
    assert s["mesh_rotation"] == [[0, 0, 0, 0, 0, 0]]
    zone = d["RigidBodies"][0]["zoneId"]
    for b in d["BoundaryConditions"]:
        if b["zoneId"] == zone:
            b["valueAng"] = [0.0, 0.0, 104.72]
    d["RigidBodies"][0]["rotationCenter"] = [0.001, -0.002, 0.0]
    p = tmp_path / "rot.json"
    p.write_text(json.dumps(d))
    s = _dry(str(p))
    assert s["mesh_rotation"] == [[0, 0, 104.72, 0.001, -0.002, 0.0]]


def test_bad_decks_fail_with_a_message(tmp_path):
    p = tmp_path / "broken.json"
    p.write_text('{"Configuration": {"particleRadius": 0.001,, }}')
    assert "JSON" in _dry(str(p), expect_rc=2)
    d = _deck("compression_mini.json")
    d["DomainBlocks"][0]["type"] = "File"
    p.write_text(json.dumps(d))
    assert "not supported" in _dry(str(p), expect_rc=2)
    d = _deck("contact_mini.json")
    d["RigidBodies"][0]["type"] = "Sphere"
    p.write_text(json.dumps(d))
    assert "not supported" in _dry(str(p), expect_rc=2)
