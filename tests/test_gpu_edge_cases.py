"""Edge cases of the tiled pair loops on the GPU against the FP64 oracle:

- one particle (no neighbours, an empty staged set);
- 129 particles (a last tile holding a single particle);
- a sparse cloud where most particles have no neighbour and the rest a few;
- two particles at the same position (excluded from each other's lists by
  the reference's 0 < d^2 criterion, but neighbours of the same others).

Each case runs the neighbour search (sets bit-exact after sorting), one call
of each pair kernel and a few whole steps, in FP64 (relative L2 <= 1e-12 per
kernel, 1e-9 over steps) and mixed precision (2e-5 per kernel, 1e-4 over
steps).  A field that is zero in the oracle must be zero on the GPU.
"""
import numpy as np
import pytest

from weldformgpu_amd import _abi
from weldformgpu_amd.domain import WfError

from harness import randomize, rel_l2, small_c1, to_gpu, to_oracle

pytestmark = pytest.mark.gpu

PRECISIONS = [(_abi.WF_FP64, 1e-12, 1e-9), (_abi.WF_MIXED, 2e-5, 1e-4)]


def _subset(sc, idx, x=None):
    """The scenario restricted to particles idx (optionally at positions x)."""
    for k, v in list(sc.fields.items()):
        sc.fields[k] = np.ascontiguousarray(v[idx])
    if x is not None:
        sc.fields["x"] = np.ascontiguousarray(x)
    sc.cfg.n_particles = len(idx)
    sc.cfg.solid_count = len(idx)
    return sc


def _one_particle():
    sc = randomize(small_c1(0.4), seed=51)
    return _subset(sc, np.array([sc.n // 2]))


def _ragged():
    sc = randomize(small_c1(0.4), seed=52)
    return _subset(sc, np.arange(129))


def _sparse():
    sc = randomize(small_c1(0.4), seed=53)
    rng = np.random.default_rng(54)
    n = 300
    dx = 2 * sc.meta["r"]
    # spread over a box ~ 40 dx wide: most particles are alone
    x = rng.uniform(0.0, 40 * dx, (n, 3))
    return _subset(sc, np.arange(n), x)


def _coincident():
    sc = randomize(small_c1(0.4), seed=55)
    n = sc.n
    idx = np.concatenate([np.arange(n), [n // 3]])
    x = sc.fields["x"][idx].copy()
    sc = _subset(sc, idx, x)
    sc.meta["twins"] = (n // 3, n)  # particle n is a copy of particle n // 3
    return sc


CASES = {"one": _one_particle, "ragged": _ragged, "sparse": _sparse, "coincident": _coincident}


def _check(got, ref, tol, name):
    if np.abs(ref).max() == 0.0:
        assert np.abs(got).max() == 0.0, name
    else:
        assert rel_l2(got, ref) < tol, name


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision,ktol,stol", PRECISIONS)
def test_edge_case_kernels_and_steps(oracle, wflib, case, precision, ktol, stol):
    sc = CASES[case]()
    sc.cfg.precision = precision
    print(f"[case] edge {case} prec={precision} n={sc.n}", flush=True)
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
        _check(d.get("drho"), st["drho"], ktol, "drho")
        st.calc_rate_tensors(cfg)
        d.CalcRateTensors()
        _check(d.get("strrate"), st["strrate"], ktol, "strrate")
        st.calc_accel(cfg)
        d.CalcAccel()
        _check(d.get("a"), st["a"], ktol, "a")
    if case == "coincident":
        i, j = sc.meta["twins"]
        assert j not in ol[oo[i]:oo[i] + oc[i]] and i not in ol[oo[j]:oo[j] + oc[j]]
        assert oc[i] == oc[j] > 0
    # whole steps from the initial state
    st = to_oracle(oracle, sc)
    st.nsearch(cfg)
    st.step(cfg, sc.bcs, 5)
    with to_gpu(sc, wflib) as d:
        d.step(5)
        for k in ["x", "v", "rho", "sigma"]:
            _check(d.get(k), st[k], stol, k)


@pytest.mark.parametrize("nz, fits", [(2400, True), (2600, False)])
def test_mixed_grid_limit_of_the_packed_cell(oracle, wflib, nz, fits):
    """Mixed precision keeps each particle's search cell packed in 11/11/10
    bits, so its grid holds at most 1024 cells of 1.01 x 2h along z.  A line
    of 2,600 particles dx apart spans 2600 / 2.424 + 3 = 1075 cells: the
    search refuses it with WF_E_UNSUPPORTED and a message, where FP64 builds
    it; 2,400 particles (993 cells) build in both precisions, with the
    oracle's lists."""
    from weldformgpu_amd.domain import WfError
    sc = randomize(small_c1(0.4), seed=57)
    dx = 2 * sc.meta["r"]
    x = np.zeros((nz, 3))
    x[:, 2] = dx * np.arange(nz)
    sc = _subset(sc, np.arange(nz) % sc.n, x)
    print(f"[case] packed-cell limit nz={nz}", flush=True)
    cnt, off, lst, _ = oracle.neighbors(x, 2 * sc.cfg.h_glob)
    assert cnt.max() == 4  # +-1 and +-2 dx along the line
    for prec in (_abi.WF_FP64, _abi.WF_MIXED):
        sc.cfg.precision = prec
        with to_gpu(sc, wflib) as d:
            if prec == _abi.WF_MIXED and not fits:
                with pytest.raises(WfError) as ei:
                    d.nsearch()
                assert ei.value.code == _abi.WF_E_UNSUPPORTED
                assert "11/11/10 bits" in str(ei.value)
                continue
            d.nsearch()
            gc, go, gl = d.neighbors()
            assert np.array_equal(gc, cnt) and np.array_equal(gl, lst)


def _uncompacted_sizes(x, radius, tile=128):
    """Particles in the occupied cells adjacent to each tile's cells (cell =
    1.01 radius on FP32 positions, tiles of 128 in the search's cell order):
    the set the list builder stages before it compacts."""
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
    cells = {}
    for cc in map(tuple, c):
        cells[cc] = cells.get(cc, 0) + 1
    sizes = []
    for t in range(0, len(order), tile):
        own = {tuple(v) for v in c[order[t:t + tile]]}
        adj = {(a + i, b + j, d + k) for a, b, d in own
               for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)}
        sizes.append(sum(cells.get(q, 0) for q in adj))
    return np.array(sizes)


@pytest.mark.parametrize("precision,tol", [(_abi.WF_FP64, 1e-12), (_abi.WF_MIXED, 2e-5)])
def test_uncompacted_staged_sets_over_4096(oracle, wflib, precision, tol):
    """6,000 particles at random in a 2.4 mm cube with 2h = 1.2 mm: every
    tile's adjacent cells hold more than the 4,096 particles the list builder
    compacts, so the lists keep raw positions into the whole set and the pair
    loops take the chunked path over it.  Every particle has far more than
    96 candidates: each list holds 96 true neighbours (the storage limit),
    and every particle counts as over the reference's cap of 70.  The pair
    kernels are compared with the oracle run on the GPU's own lists."""
    rng = np.random.default_rng(58)
    sc = randomize(small_c1(0.4), seed=58, v_scale=0.5, s_scale=100e6)
    n = 6000
    x = rng.uniform(0.0, 2.4e-3, (n, 3))
    sc = _subset(sc, np.arange(n) % sc.n, x)
    sc.cfg.precision = precision
    radius = 2 * sc.cfg.h_glob
    sizes = _uncompacted_sizes(x, radius)
    print(f"[case] uncompacted sets: {len(sizes)} tiles, {sizes.min()}..{sizes.max()} "
          f"particles, precision {precision}", flush=True)
    assert sizes.min() > 4096
    st = to_oracle(oracle, sc)
    with to_gpu(sc, wflib) as d:
        d.nsearch()
        gc, go, gl = d.neighbors()
        _, over = d.nsearch_stats()
        assert over == n and np.all(gc == _abi.WF_MAX_NEIGHBORS_SKIN)
        cnt, off, lst, _ = oracle.neighbors(x, radius)
        for i in range(0, n, 97):
            assert set(gl[go[i]:go[i] + gc[i]]) <= set(lst[off[i]:off[i] + cnt[i]])
        st["nb_count"] = gc
        st["nb_offset"] = go
        st["nb_list"][:len(gl)] = gl
        for call, name, fields in (("CalcDensInc", "calc_dens_inc", ["drho"]),
                                   ("CalcRateTensors", "calc_rate_tensors",
                                    ["strrate", "rotrate"]),
                                   ("CalcAccel", "calc_accel", ["a"])):
            getattr(d, call)()
            getattr(st, name)(sc.cfg)
            for k in fields:
                e = rel_l2(d.get(k), st[k])
                print(f"  {k}: {e:.2e}")
                assert e < tol, k


@pytest.mark.parametrize("precision", [_abi.WF_FP64, _abi.WF_MIXED])
def test_packed_cell_limit(wflib, precision):
    """Mixed precision keeps each particle's search cell packed in 11/11/10
    bits, so its grid holds at most 2048 x 2048 x 1024 cells of 1.01 (2h):
    a line of particles longer than 1024 cells along z is refused with
    WF_E_UNSUPPORTED and a message naming the limit; FP64 builds its lists."""
    sc = randomize(small_c1(0.4), seed=55)
    dx = 2 * sc.meta["r"]
    cells = 2 * sc.cfg.h_glob * 1.01 / dx   # a search cell in units of dx
    nz = int(1100 * cells)                  # > 1024 cells along z
    idx = np.arange(nz) % sc.n
    x = np.zeros((nz, 3))
    x[:, 2] = np.arange(nz) * dx
    sc = _subset(sc, idx, x)
    sc.cfg.precision = precision
    print(f"[case] packed cell limit prec={precision} n={sc.n}", flush=True)
    with to_gpu(sc, wflib) as d:
        if precision == _abi.WF_MIXED:
            with pytest.raises(WfError, match="2048 x 2048 x 1024"):
                d.nsearch()
        else:
            d.nsearch()
            total, over = d.nsearch_stats()
            assert over == 0 and total == 2 * (nz - 1) + 2 * (nz - 2)
