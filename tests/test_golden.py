"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from the FP64 oracle) are re-derived bit for
bit by the oracle: any change to the oracle's arithmetic or to the inputs'
construction shows up here."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_fixtures as mf  # noqa: E402


def _load(name):
    return np.load(os.path.join(HERE, "golden", name), allow_pickle=False)


def test_lattice_fixture():
    g = _load("lattice.npz")
    cfg, fields, bcs = mf.lattice_case()
    for k, v in fields.items():
        assert np.array_equal(g["in_" + k], v), k
    out = mf.lattice_outputs(cfg, fields, bcs)
    for k, v in out.items():
        assert np.array_equal(g["out_" + k], v), k
    assert g["out_step5_pl_strain"].max() > 0  # the case yields


def test_stress_fixture():
    g = _load("stress.npz")
    rows = mf.stress_case()
    out = mf.stress_outputs(rows)
    for k, v in out.items():
        assert np.array_equal(g["out_" + k], v), k
    ep0 = np.array([r[4] for r in rows])
    assert np.all(g["out_pl_strain"] >= ep0) and np.any(g["out_pl_strain"] > ep0)


def test_contact_fixture():
    g = _load("contact.npz")
    out = mf.contact_outputs(*mf.contact_case())
    for k, v in out.items():
        assert np.array_equal(g["out_" + k], v), k
    assert g["out_cnb_count"].tolist() == [1, 2, 0]


@pytest.mark.parametrize("name", ["mlp.npz"])
def test_mlp_fixture(name):
    g = _load(name)
    out = mf.mlp_outputs(*mf.mlp_case())
    for k, v in out.items():
        assert np.array_equal(g["out_" + k], v), k
