"""GPU cell list vs the CPU oracle: pair lists are bit-exact as sorted (i, j)
sets, together with every reference-visible field (grid, offset, cellpnum,
pcell, pairsn, capacity, mppcell, мaxpairs)."""
import math
import os
import warnings

import numpy as np
import pytest
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def as_set(pairs):
    return np.unique(np.asarray(pairs), axis=0)


def assert_same_celllist(g, o, check_capacity=True):
    inf = o.info()
    assert g.grid == inf["grid"]
    assert g.offset == inf["offset"]
    assert g.cs == inf["cs"]
    assert g.pairsn == inf["pairsn"]
    if check_capacity:
        assert g.capacity == inf["capacity"]
        assert g.mppcell == inf["mppcell"]
    pc, cp = o.cells()
    assert np.array_equal(g.cellpnum, cp)
    assert np.array_equal(g.pcell, pc)
    assert np.array_equal(as_set(__import__("gclsph").neighborlistview(g)), as_set(o.neighborlistview()))


def test_config0(G, O, S):
    sc = S.config0()
    g = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    o = O.CellList(sc.points, sc.cellsize, sc.dist)
    assert_same_celllist(g, o)
    assert g.maxpairs() == o.maxpairs()


def test_reference_fixture(G, O, S):
    """test/runtests.jl: GPUCellList(points, (2h, 2h), 2h) on the Dp0.02 CSVs, then update!."""
    pts, rho, ptype = S.load_dualsphysics_csv(os.path.join(GOLDEN, "FluidPoints_Dp0.02.csv"),
                                              os.path.join(GOLDEN, "BoundaryPoints_Dp0.02.csv"))
    h = 1.2 * math.sqrt(2) * 0.02
    g = G.GPUCellList(pts, (2 * h, 2 * h), 2 * h)
    o = O.CellList(pts, (2 * h, 2 * h), 2 * h)
    G.update_(g)
    o.update()
    assert_same_celllist(g, o)


def test_config1_update_and_partialupdate(G, O, S):
    """200,000 uniform points, cellsize = dist = 0.016, jittered by N(0, 0.001^2) per step."""
    pts = S.config1_points()
    g = G.GPUCellList(pts, (0.016, 0.016), 0.016)
    o = O.CellList(pts, (0.016, 0.016), 0.016)
    assert_same_celllist(g, o)
    cur = pts
    for step in range(3):
        cur = S.jitter(cur, 12345, step)
        g.points = cur
        o.set_points(cur)
        if step % 2 == 0:
            G.partialupdate_(g)
            o.partialupdate()
        else:
            G.update_(g)
            o.update()
        assert_same_celllist(g, o)


def test_3d_random(G, O, S):
    pts = S.uniform_points(77, 40000, 3)
    g = G.GPUCellList(pts, (0.05, 0.06, 0.05), 0.05)
    o = O.CellList(pts, (0.05, 0.06, 0.05), 0.05)
    assert_same_celllist(g, o)
    moved = S.jitter(pts, 77, 0, sigma=0.004)
    g.points = moved
    o.set_points(moved)
    G.update_(g)
    o.update()
    assert_same_celllist(g, o)


def test_float32(G, O, S):
    pts = S.uniform_points(21, 30000, 2)
    g = G.GPUCellList(pts, (0.011, 0.011), 0.011, precision="f32")
    o = O.CellList(pts, (0.011, 0.011), 0.011, prec=32)
    assert_same_celllist(g, o)
    moved = S.jitter(pts, 21, 0, sigma=0.002)
    g.points = moved
    o.set_points(moved)
    G.update_(g)
    o.update()
    assert_same_celllist(g, o)


def test_small_mpairs_overflow_and_regrow(G, O, S):
    pts = S.uniform_points(5, 5000, 2)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        g = G.GPUCellList(pts, 0.03, 0.03, mpairs=1000)
        assert any("not full" in str(x.message) for x in w)
    o = O.CellList(pts, 0.03, 0.03, mpairs=1000)
    assert g.pairsn == o.pairsn > g.capacity == 1000
    with pytest.raises(IndexError):
        G.neighborlistview(g)
    G.update_(g)
    o.update()
    assert_same_celllist(g, o)


def test_partialupdate_overflow_keeps_true_count(G, O, S):
    pts = S.uniform_points(6, 4000, 2)
    g = G.GPUCellList(pts, 0.04, 0.04)
    o = O.CellList(pts, 0.04, 0.04)
    squeezed = tuple(0.5 + 0.6 * (p - 0.5) for p in pts)  # denser: more pairs, stale cells
    g.points = squeezed
    o.set_points(squeezed)
    G.partialupdate_(g)
    o.partialupdate()
    assert g.pairsn == o.pairsn > g.capacity == o.info()["capacity"]
    with pytest.raises(IndexError):
        G.neighborlistview(g)
    assert len(G.neighborlist(g)) == g.capacity


def test_fillzero_tail(G, O, S):
    """update!/partialupdate! without reallocation (src/celllist.jl:95-134,
    141-145): with fillzero = false the entries past pairsn keep the previous
    search's contents; with fillzero = true they are (0, 0)."""
    u = S.uniform_points(12, 30000, 2)
    pts = tuple(0.1 + 0.8 * p for p in u)
    kw = dict(MIN=(0.0, 0.0), MAX=(1.0, 1.0))
    g = G.GPUCellList(pts, 0.02, 0.02, **kw)
    o = O.CellList(pts, 0.02, 0.02, **kw)
    cap, n0 = g.capacity, g.pairsn
    spread = tuple(0.5 + 1.03 * (p - 0.5) for p in pts)  # fewer pairs, same grid, no reallocation
    for upd, oupd in ((G.update_, o.update), (G.partialupdate_, o.partialupdate)):
        g.points = pts
        o.set_points(pts)
        G.update_(g)
        o.update()
        before = G.neighborlist(g)
        g.points = spread
        o.set_points(spread)
        upd(g, False)
        oupd()
        n1 = g.pairsn
        assert n1 == o.pairsn and 0.6 * cap <= n1 < n0 and g.capacity == cap
        after = G.neighborlist(g)
        assert np.array_equal(after[n1:], before[n1:]) and np.any(after[n1:n0] != 0)
        assert np.array_equal(as_set(after[:n1]), as_set(o.neighborlistview()))
        upd(g, True)
        after = G.neighborlist(g)
        assert g.pairsn == n1 and not np.any(after[n1:])
        assert np.array_equal(as_set(after[:n1]), as_set(o.neighborlistview()))


@pytest.mark.parametrize("dim,dist", [(2, 0.05), (2, 0.015), (3, 0.124), (3, 0.06)])
def test_pair_emit_variants_agree(G, O, S, monkeypatch, dim, dist):
    """The one-lane-per-point and one-wave-per-point emits write the same
    buffer (dense sets of ~80 pairs per point and sparse ones of ~7)."""
    pts = S.uniform_points(21 + dim, 20000, dim)
    o = O.CellList(pts, dist, dist)
    bufs = []
    for mode in ("lane", "wave"):
        monkeypatch.setenv("GCLSPH_PAIR_EMIT", mode)
        g = G.GPUCellList(pts, dist, dist)
        G.update_(g)
        assert g.pairsn == o.info()["pairsn"]
        bufs.append(G.neighborlist(g))
    assert np.array_equal(bufs[0], bufs[1])
    assert np.array_equal(as_set(bufs[1][: o.info()["pairsn"]]), as_set(o.neighborlistview()))


@pytest.mark.parametrize("cs", [0.1, 0.3])
def test_stored_windows_with_large_cells(G, O, S, monkeypatch, cs):
    """The 2D count pass keeps each point's neighbour windows for the emit pass
    (at most 6; a point with more is tested again by the emit) and stages the
    candidates of a wave's stencil rows in LDS when they fit.  Points sorted by
    x in cells 5-15 times dist: a wave's rows hold far more candidates than the
    LDS slots (the global-memory path of both passes), while a point's
    neighbours lie in a run of x-neighbours per cell, so most points' windows
    fit the store and some do not.  Lane and wave emits write the same buffer,
    equal as a set to the oracle's, after update! and partialupdate!."""
    pts = S.uniform_points(31, 30000, 2)
    order = np.argsort(pts[0], kind="stable")
    pts = tuple(p[order] for p in pts)
    dist = 0.02
    o = O.CellList(pts, cs, dist)
    bufs = []
    for mode in ("lane", "wave"):
        monkeypatch.setenv("GCLSPH_PAIR_EMIT", mode)
        g = G.GPUCellList(pts, cs, dist)
        G.update_(g)
        assert g.pairsn == o.info()["pairsn"]
        bufs.append(G.neighborlist(g))
        g.points = S.jitter(pts, 5, 1, sigma=0.002)
        G.partialupdate_(g)
        bufs.append(G.neighborlist(g)[: g.pairsn])
    assert np.array_equal(bufs[0], bufs[2]) and np.array_equal(bufs[1], bufs[3])
    assert np.array_equal(as_set(bufs[0][: o.info()["pairsn"]]), as_set(o.neighborlistview()))
    o.set_points(S.jitter(pts, 5, 1, sigma=0.002))
    o.partialupdate()
    assert np.array_equal(as_set(bufs[1]), as_set(o.neighborlistview()))


def test_points_outside_min_max_clamp(G, O, S):
    pts = S.uniform_points(8, 20000, 2)
    kw = dict(MIN=(0.25, 0.2), MAX=(0.7, 0.9))
    g = G.GPUCellList(pts, 0.02, 0.02, **kw)
    o = O.CellList(pts, 0.02, 0.02, **kw)
    assert_same_celllist(g, o)


def test_cellsize_below_dist_warns(G, S):
    pts = S.uniform_points(3, 1000, 2)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        g = G.GPUCellList(pts, (0.01, 0.05), 0.03)
        assert any("Cell size" in str(x.message) for x in w)
    assert g.cs == (0.03, 0.05)


def test_degenerate_inputs(G, S):
    g = G.GPUCellList((np.array([0.5]), np.array([0.5])), 0.1, 0.1, MIN=(0.0, 0.0), MAX=(1.0, 1.0))
    assert g.pairsn == 0 and len(G.neighborlistview(g)) == 0
    # identical points: the lowered MIN (src/celllist.jl:42) still gives one cell; all pairs at r = 0
    g = G.GPUCellList((np.full(10, 0.3), np.full(10, 0.3)), 0.1, 0.1)
    assert g.grid == (1, 1) and g.pairsn == 45
    with pytest.raises(ValueError, match="no cells"):
        G.GPUCellList((np.full(10, 0.3), np.full(10, 0.3)), 0.1, 0.1, MIN=(0.5, 0.5), MAX=(0.2, 0.2))
    # every point in one cell: all pairs within dist
    pts = (0.5 + 1e-3 * S.uniform01(1, 50, 0), 0.5 + 1e-3 * S.uniform01(1, 50, 1))
    g = G.GPUCellList(pts, 0.1, 0.1)
    assert g.pairsn == 50 * 49 // 2


def test_deterministic_order(G, S):
    pts = S.uniform_points(31, 50000, 2)
    a = G.neighborlistview(G.GPUCellList(pts, 0.01, 0.01))
    b = G.neighborlistview(G.GPUCellList(pts, 0.01, 0.01))
    assert np.array_equal(a, b)
