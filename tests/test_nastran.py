"""Nastran small-field triangle meshes (wf_read_nastran, the reference's
NastranReader + ReadFromNastran): known answers on a committed fixture,
orientation flip, use from a deck, and — when the reference's example mesh
is on disk — agreement with an independent reading of its fixed fields."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from weldformgpu_amd import build as wbuild
from weldformgpu_amd import scenarios

DECKS = os.path.join(os.path.dirname(__file__), "golden", "decks")
REF_NAS = "/root/reference/examples/input/tool_metal_cut_mm.nas"


def test_fixture_known_answers():
    nodes, el = scenarios.read_nastran(os.path.join(DECKS, "tool_mini.nas"))
    exp = [[0, 0, 5e-4], [1, 0, 5e-4], [1, 1, 5e-4], [0, 1, 5e-4], [-0.25, 1, -1],
           [0.25, 1, -1], [0, 2, 0]]
    assert np.array_equal(nodes, np.array(exp, float))
    assert el.tolist() == [[0, 1, 2], [0, 2, 3], [4, 5, 6]]
    _, fl = scenarios.read_nastran(os.path.join(DECKS, "tool_mini.nas"), flip=True)
    assert fl.tolist() == [[0, 2, 1], [0, 3, 2], [4, 6, 5]]
    # orientation: n0n1 x n0n2 of the square points +z, flipped -z
    P = nodes[el[0]]
    assert np.cross(P[1] - P[0], P[2] - P[0])[2] > 0


def test_missing_file_and_unknown_node(tmp_path):
    with pytest.raises(ValueError):
        scenarios.read_nastran(tmp_path / "none.nas")
    bad = tmp_path / "bad.nas"
    bad.write_text("GRID           1             0.0     0.0     0.0\n"
                   "CTRIA3         1       1       1       2       3\n")
    with pytest.raises(ValueError):
        scenarios.read_nastran(bad)


def test_deck_with_file_rigid_body():
    if not os.path.exists(wbuild.DRIVER):
        wbuild.build_driver()
    r = subprocess.run([wbuild.DRIVER, "--dry-run", os.path.join(DECKS, "contact_file_mini.json")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    s = json.loads(r.stdout)
    assert s["contact"] == 1 and s["meshes"] == [[10, 7, 3, 0, 0, -1]]


def _ref_real(t):
    """A second reading of an 8-character Nastran real, by regular expression."""
    t = t.strip()
    m = re.fullmatch(r"([+-]?[0-9.]+)(?:[EeDd]?([+-][0-9]+))?", t)
    assert m, t
    return float(m.group(1)) * 10.0 ** int(m.group(2) or 0)


@pytest.mark.skipif(not os.path.exists(REF_NAS), reason="reference example mesh not present")
def test_reference_example_mesh():
    nodes, el = scenarios.read_nastran(REF_NAS)
    grid, tri = [], []
    with open(REF_NAS) as f:
        for line in f:
            if line.startswith("GRID") and not line.startswith("GRID*"):
                grid.append((int(line[8:16]), [_ref_real(line[24 + 8 * c:32 + 8 * c])
                                               for c in range(3)]))
            elif line.startswith("CTRIA3"):
                tri.append([int(line[24 + 8 * c:32 + 8 * c]) for c in range(3)])
    pos = {gid: k for k, (gid, _) in enumerate(grid)}
    assert len(nodes) == len(grid) and len(el) == len(tri) > 0
    assert np.allclose(nodes, [g[1] for g in grid], rtol=1e-15, atol=0)
    assert el.tolist() == [[pos[t] for t in row] for row in tri]
