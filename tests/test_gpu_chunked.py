"""Pair loops over staged sets larger than their LDS capacity.

A tile's staged set (the particles some list of the tile references) is
written to LDS in one pass when it fits the kernel's 1,024 slots; otherwise
the same kernel streams it in chunks and each thread walks its list chunk by
chunk.  Lattices with h = 1.2 dx never need the chunked path (at most ~910
staged particles per tile), so this test widens the lists: the plastic-zone
rebuild policy with a skin of 0.59 dx gives a search radius of 2.99 dx, 92
neighbours per bulk particle (within the 96-entry skin cap) and, on the
93,800-particle Johnson-Cook cylinder (R 7.5 mm, dx 0.3 mm), staged sets of
~800 particles on average and up to ~1,090 (18 of 733 tiles over 1,024).
The test counts on the CPU, from the same cell ordering, how many tiles
exceed the stage, and checks the density rate, strain/rotation rates and acceleration
of both precisions against the FP64 oracle on those lists.

Tolerances: lists bit-exact; FP64 one kernel call relative L2 <= 1e-12;
mixed precision <= 2e-5.
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi, scenarios

from harness import randomize, rel_l2, to_gpu, to_oracle

STAGE = 1024
TILE = 128


def _tile_union_sizes(x, radius):
    """Staged-set size of every tile: particles ordered by search cell (cell
    = 1.01 radius on FP32 positions, 8^3 meta-blocks of Morton-ordered
    cells, ascending index within a cell), 128 per tile, union of the
    tile's neighbour rows."""
    from scipy.spatial import cKDTree
    xf = x.astype(np.float32).astype(np.float64)
    cs = float(np.float32(radius)) * 1.01
    org = xf.min(0) - cs
    c = np.floor((xf - org) / cs).astype(np.int64)
    mdim = (c.max(0) + 2 + 7) // 8
    m = np.zeros(len(xf), np.int64)
    for b in range(3):
        for a in range(3):
            m |= (((c[:, a] & 7) >> b) & 1) << (3 * b + a)
    key = (((c[:, 2] >> 3) * mdim[1] + (c[:, 1] >> 3)) * mdim[0] + (c[:, 0] >> 3)) * 512 + m
    xs = xf[np.lexsort((np.arange(len(xf)), key))]
    tree = cKDTree(xs)
    sizes = []
    for t in range(0, len(xs), TILE):
        rows = tree.query_ball_point(xs[t:t + TILE], radius * (1 - 1e-7))
        sizes.append(len(set().union(*rows)))
    return np.array(sizes)


def _case():
    sc = randomize(scenarios.c3_johnson_cook(scale=0.5), seed=41)
    dx = 2 * sc.meta["r"]
    sc.cfg.rebuild_policy = _abi.WF_REBUILD_PLASTIC_ZONE
    sc.cfg.ns_skin = 0.59 * dx
    sc.cfg.ep_hot = 0.05
    return sc


def test_chunked_staging_needed():
    sc = _case()
    radius = 2 * sc.cfg.h_glob + sc.cfg.ns_skin
    sizes = _tile_union_sizes(sc.fields["x"], radius)
    print(f"[case] tiles {len(sizes)}, staged max {sizes.max()}, over {STAGE}: "
          f"{(sizes > STAGE).sum()}", flush=True)
    assert (sizes > STAGE).sum() >= 10


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-12), (_abi.WF_MIXED, 2e-5)])
def test_chunked_pair_kernels(oracle, wflib, precision, tol):
    sc = _case()
    sc.cfg.precision = precision
    print(f"[case] chunked staging prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    cfg = sc.cfg
    assert st.nsearch(cfg) == 0  # no row over the skin-list cap
    oc, oo, ol = st.neighbor_rows()
    assert 80 < oc.max() <= _abi.WF_MAX_NEIGHBORS_SKIN
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        gc, go, gl = d.neighbors()
        _, over = d.nsearch_stats()
        assert over == 0
        assert np.array_equal(gc, oc) and np.array_equal(gl, ol)
        st.calc_dens_inc(cfg)
        d.CalcDensInc()
        assert rel_l2(d.get("drho"), st["drho"]) < tol
        st.calc_rate_tensors(cfg)
        d.CalcRateTensors()
        assert rel_l2(d.get("strrate"), st["strrate"]) < tol
        assert rel_l2(d.get("rotrate"), st["rotrate"]) < tol
        st.calc_accel(cfg)
        d.CalcAccel()
        assert rel_l2(d.get("a"), st["a"]) < tol


# ---------------------------------------------------------------------------
# Staged sets left uncompacted.  The list builder compacts a tile's
# adjacent-cell candidates to the particles its lists reference when there
# are at most cfg.ns_compact_max of them (default 4096, the LDS flag array),
# and stages the whole candidate set otherwise.  Lattices within the list caps
# have 700-2,200 candidates per tile, so the default never takes the second
# branch; ns_compact_max = 256 sends every tile of this case there (staged
# sets of the full candidate count, many above the 1,024-slot stage, so the
# chunked pair loops run too).

COMPACT = 256


def _candidate_counts(x, radius):
    """Candidates of every tile: the particles in the union of the 27-cell
    neighbourhoods of the tile's particles' cells (cells of 1.01 radius on
    FP32 positions, tiles of 128 particles in the search's cell order)."""
    xf = x.astype(np.float32).astype(np.float64)
    cs = float(np.float32(radius)) * 1.01
    org = xf.min(0) - cs
    c = np.floor((xf - org) / cs).astype(np.int64)
    mdim = (c.max(0) + 2 + 7) // 8
    m = np.zeros(len(xf), np.int64)
    for b in range(3):
        for a in range(3):
            m |= (((c[:, a] & 7) >> b) & 1) << (3 * b + a)
    key = (((c[:, 2] >> 3) * mdim[1] + (c[:, 1] >> 3)) * mdim[0] + (c[:, 0] >> 3)) * 512 + m
    order = np.lexsort((np.arange(len(xf)), key))
    cells = [tuple(v) for v in c[order]]
    count = {}
    for v in cells:
        count[v] = count.get(v, 0) + 1
    out = []
    for t in range(0, len(cells), TILE):
        adj = set()
        for (i, j, k) in set(cells[t:t + TILE]):
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    for dk in (-1, 0, 1):
                        adj.add((i + di, j + dj, k + dk))
        out.append(sum(count.get(v, 0) for v in adj))
    return np.array(out)


def _uncompacted_case():
    sc = randomize(scenarios.c3_johnson_cook(scale=0.3), seed=43)
    sc.cfg.ns_compact_max = COMPACT
    return sc


def test_uncompacted_staging_needed():
    sc = _uncompacted_case()
    u = _candidate_counts(sc.fields["x"], 2 * sc.cfg.h_glob)
    print(f"[case] tiles {len(u)}, candidates min {u.min()} max {u.max()}, over "
          f"{STAGE}: {(u > STAGE).sum()}", flush=True)
    assert (u > COMPACT).all() and (u > STAGE).sum() >= 10


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol,tol_steps", [(_abi.WF_FP64, 1e-12, 1e-9),
                                                     (_abi.WF_MIXED, 2e-5, 1e-4)])
def test_uncompacted_pair_kernels(oracle, wflib, precision, tol, tol_steps):
    sc = _uncompacted_case()
    sc.cfg.precision = precision
    print(f"[case] uncompacted staging prec={precision} n={sc.n}", flush=True)
    st = to_oracle(oracle, sc)
    cfg = sc.cfg
    st.nsearch(cfg)
    oc, oo, ol = st.neighbor_rows()
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        gc, go, gl = d.neighbors()
        assert np.array_equal(gc, oc) and np.array_equal(gl, ol)
        st.calc_dens_inc(cfg)
        d.CalcDensInc()
        assert rel_l2(d.get("drho"), st["drho"]) < tol
        st.calc_rate_tensors(cfg)
        d.CalcRateTensors()
        assert rel_l2(d.get("strrate"), st["strrate"]) < tol
        assert rel_l2(d.get("rotrate"), st["rotrate"]) < tol
        st.calc_accel(cfg)
        d.CalcAccel()
        assert rel_l2(d.get("a"), st["a"]) < tol
    # whole steps from a fresh state, through the step driver
    sc = _uncompacted_case()
    sc.cfg.precision = precision
    st = to_oracle(oracle, sc)
    st.nsearch(sc.cfg)
    st.step(sc.cfg, sc.bcs, 6)
    with to_gpu(sc, wflib) as d:
        d.step(6)
        for k in ("u", "v", "sigma", "rho"):
            err = rel_l2(d.get(k), st[k])
            print(f"  {k}: {err:.2e}")
            assert err < tol_steps, k
