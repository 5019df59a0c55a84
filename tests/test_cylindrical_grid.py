"""The cylindrical grid (Domain::AddCylUniformLength, Domain.cpp:1071-1193)
and how a deck selects it (gridCoordSys "Cylindrical", WeldFormGPU.cu:
366-375), on the CPU.

The oracle's generator is pinned by facts of the construction that do not
restate its loop: the per-layer count of the whole circle is
1 + sum_k ceil(2 pi k) over rings k 2r < R, every particle lies on its ring
at equal angular spacing, the mass times the count is the whole cylinder's,
and the IDs mark the axis (4) and, on a slice, each ring's first and last
angle (2, 3).  The product's generator (wf_gen_cylinder_polar) must then
equal the oracle's bit for bit, and the driver must build the grid and the
axisymmetric clamp flag exactly as the reference's deck reader does: the
whole circle only for a sliceAngle of exactly 2 pi, otherwise (the default
2.0000001 pi included) a pi/16 slice of the outer half of the radius.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from weldformgpu_amd import build as wbuild
from weldformgpu_amd import scenarios

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")


def _rings(R, r, r_i=0.0):
    """Ring radii of the construction, in the loop's own accumulation."""
    out, ri = [], r_i
    while ri < R:
        out.append(ri)
        ri += 2 * r
    return out


def test_full_circle_counts_and_positions(oracle):
    R, H, r, rho = 2.9e-3, 4.1e-3, 0.25e-3, 7850.0
    xyz, ID, mass = oracle.gen_cylinder_polar(R, H, r, rho)
    # layers at z = r, 3r, ... <= H: 0.25 .. 3.75 mm
    zs = np.unique(xyz[:, 2])
    assert np.allclose(zs, 0.25e-3 + 0.5e-3 * np.arange(8), rtol=0, atol=1e-15)
    K = 5  # rings 0.5 .. 2.5 mm (3.0 mm is past R)
    per_layer = 1 + sum(math.ceil(2 * math.pi * k) for k in range(1, K + 1))
    assert per_layer == 1 + 7 + 13 + 19 + 26 + 32
    assert len(xyz) == 8 * per_layer
    assert mass * len(xyz) == pytest.approx(math.pi * R * R * H * rho, rel=1e-13)
    layer = xyz[xyz[:, 2] == zs[0]]
    rad = np.hypot(layer[:, 0], layer[:, 1])
    for k in range(K + 1):
        ring = layer[np.abs(rad - 2 * r * k) < 1e-9]
        n_k = 1 if k == 0 else math.ceil(2 * math.pi * k)
        assert len(ring) == n_k
        if k:
            ang = np.sort(np.mod(np.arctan2(ring[:, 1], ring[:, 0]), 2 * np.pi))
            assert ang[0] == pytest.approx(0.0, abs=1e-12)
            assert np.allclose(np.diff(ang), 2 * np.pi / n_k, rtol=0, atol=1e-12)
    # IDs: the axis particle of each layer is 4, the rest the tag
    assert np.array_equal(np.flatnonzero(ID == 4), np.flatnonzero(np.hypot(xyz[:, 0], xyz[:, 1]) == 0))
    assert (ID == 4).sum() == 8 and set(np.unique(ID)) == {0, 4}


def test_slice_ids_and_ring_counts(oracle):
    R, H, r, rho = 15e-3, 3e-3, 0.15e-3, 2700.0
    ang = math.pi / 16
    xyz, ID, mass = oracle.gen_cylinder_polar(R, H, r, rho, ang, 1, R / 2)
    rings = _rings(R, r, R / 2)
    counts = [math.ceil(ang * ri / (2 * r)) for ri in rings]
    nz = len(np.unique(xyz[:, 2]))
    assert nz == 10 and len(xyz) == nz * sum(counts)
    # the mass is the whole cylinder's volume over the slice's count
    assert mass * len(xyz) == pytest.approx(math.pi * R * R * H * rho, rel=1e-13)
    theta = np.arctan2(xyz[:, 1], xyz[:, 0])
    assert theta.min() == 0.0 and theta.max() < ang
    k = 0
    layer_ids = ID[:sum(counts)]
    for c in counts:
        ids = layer_ids[k:k + c]
        assert ids[0] == 2 and ids[-1] == 3 and (ids[1:-1] == 0).all()
        k += c
    assert (ID == 2).sum() == (ID == 3).sum() == nz * len(rings)


def test_rows_and_tag(oracle):
    """rows = 2 marks two angles at each end and the tag fills the rest; on
    a ring too short for both ends, the first-row mark 2 wins."""
    xyz, ID, _ = oracle.gen_cylinder_polar(4e-3, 1e-3, 0.25e-3, 1.0, math.pi / 2, 2, 0.5e-3, tag=7)
    rings = _rings(4e-3, 0.25e-3, 0.5e-3)
    k = 0
    for ri in rings:
        c = math.ceil((math.pi / 2) * ri / 0.5e-3)
        ids = ID[k:k + c]
        want = [2 if a < 2 else (3 if a > c - 3 else 7) for a in range(c)]
        assert ids.tolist() == want
        k += c


@pytest.mark.parametrize("args", [
    (2.9e-3, 4.1e-3, 0.25e-3, 7850.0, 2 * math.pi, 1, 0.0),
    (15e-3, 3e-3, 0.15e-3, 2700.0, math.pi / 16, 1, 7.5e-3),
    (7.5e-3, 15e-3, 0.25e-3, 7850.0, 2 * math.pi, 1, 0.0),
    (4e-3, 1e-3, 0.25e-3, 1.0, math.pi / 2, 2, 0.5e-3),
])
def test_product_grid_matches_oracle(oracle, args):
    R, H, r, rho, ang, rows, ri = args
    a, ia, ma = oracle.gen_cylinder_polar(R, H, r, rho, ang, rows, ri)
    b, ib, mb = scenarios.gen_cylinder_polar(R, H, r, rho, ang, rows, ri)
    assert np.array_equal(a, b) and np.array_equal(ia, ib) and ma == mb


def test_product_grid_rejects_bad_arguments():
    with pytest.raises(ValueError):
        scenarios.gen_cylinder_polar(1e-3, 1e-3, 0.0, 1.0)
    with pytest.raises(ValueError):
        scenarios.gen_cylinder_polar(1e-3, 1e-3, 1e-4, 1.0, ang=0.0)


def _dry(deck):
    r = subprocess.run([wbuild.DRIVER, "--dry-run", deck], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@pytest.fixture(scope="module", autouse=True)
def _driver_built():
    if not os.path.exists(wbuild.DRIVER):
        wbuild.build_driver()


def test_deck_default_slice_angle_builds_the_slice(oracle):
    path = os.path.join(DECKS, "cylindrical_slice_mini.json")
    dk = json.load(open(path))
    assert "sliceAngle" not in dk["DomainBlocks"][0]
    s = _dry(path)
    L = dk["DomainBlocks"][0]["dim"]
    r, rho = dk["Configuration"]["particleRadius"], dk["Materials"][0]["density0"]
    xyz, ID, mass = oracle.gen_cylinder_polar(L[0] / 2, L[2], r, rho, math.pi / 16, 1, L[0] / 4)
    assert s["n"] == len(xyz) and s["mass"] == mass and s["axisymm"] == 1
    counts = [oracle.assign_zone(xyz, ID, z["start"], z["end"], z["id"])
              for z in dk["DomainZones"]]
    assert [zc for _, zc in s["zones"]] == counts


def test_deck_exact_two_pi_builds_the_whole_circle(oracle, tmp_path):
    path = os.path.join(DECKS, "cylindrical_full_mini.json")
    dk = json.load(open(path))
    assert dk["DomainBlocks"][0]["sliceAngle"] == 2 * math.pi
    s = _dry(path)
    L = dk["DomainBlocks"][0]["dim"]
    r, rho = dk["Configuration"]["particleRadius"], dk["Materials"][0]["density0"]
    xyz, _, mass = oracle.gen_cylinder_polar(L[0] / 2, L[2], r, rho)
    assert s["n"] == len(xyz) and s["mass"] == mass and s["axisymm"] == 0
    # a sliceAngle one ulp away from 2 pi is a slice again
    dk["DomainBlocks"][0]["sliceAngle"] = math.nextafter(2 * math.pi, 7.0)
    p2 = tmp_path / "near.json"
    p2.write_text(json.dumps(dk))
    s2 = _dry(str(p2))
    assert s2["axisymm"] == 1 and s2["n"] != s["n"]
    # the Cartesian lattice of the same block is the other generator
    dk["DomainBlocks"][0]["gridCoordSys"] = "Cartesian"
    p3 = tmp_path / "cart.json"
    p3.write_text(json.dumps(dk))
    xyz_c, mass_c = oracle.gen_cylinder([0, 0, 0], L[0] / 2, L[2], r, rho)
    s3 = _dry(str(p3))
    assert s3["n"] == len(xyz_c) and s3["mass"] == mass_c and s3["axisymm"] == 0


def test_axisymm_clamp_oracle_condition(oracle):
    """The clamp zeroes a_x and a_y exactly where |a_xy| > 1e-3, x > h/2 and
    a_x > 1e-2, and leaves a_z and every other particle alone."""
    from harness import small_c1, to_oracle
    sc = small_c1()
    st = to_oracle(oracle, sc)
    n = sc.n
    rng = np.random.default_rng(3)
    a = rng.normal(0.0, 1.0, (n, 3)) * rng.choice([1e-4, 5e-3, 1.0, 1.0], (n, 1))
    x = st["x"].copy()
    x[:, 0] = rng.uniform(-1.0, 3.0, n) * sc.cfg.h_glob
    st["x"] = x
    st["a"] = a
    st.apply_axisymm_bc()
    got = st["a"]
    h = st["h"]
    hit = (np.hypot(a[:, 0], a[:, 1]) > 1e-3) & (x[:, 0] > h / 2) & (a[:, 0] > 1e-2)
    assert 0.1 * n < hit.sum() < 0.5 * n
    assert np.array_equal(got[hit, :2], np.zeros((hit.sum(), 2)))
    assert np.array_equal(got[~hit], a[~hit]) and np.array_equal(got[:, 2], a[:, 2])
