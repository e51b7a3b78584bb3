"""HIP distance joins vs the oracle (CPU restatement) through libgsj's C ABI: ST_DWithin(geography) of polygon
layers and of points against polygons, with reported distances.

Bar: pair sets bit-exact; distances bit-exact as well (the engine and the oracle share the portable elementary
functions, so every FP64 operation rounds the same), with the metric's 1e-6 m kept as the documented bound."""
import numpy as np
import pytest

import gsj
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

DIST_ATOL = 1e-6   # metres: documented bound; expected to match to the bit
CELL = 0.6 / 1095  # the dwithin layer's lattice: 1095 x 1095 cells over 0.6 x 0.6 degrees (~61 m x 47 m at 38.9 N)


def _dw_layer(n=50, seed=2, **kw):
    return gsj.generate("parcels_mesh", nx=n, ny=n, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=seed, **kw)


def _check(got, exp, report):
    g = sorted_rows(got.pairs, got.dist if report else None)
    e = sorted_rows(exp["pairs"], exp["dist"] if report else None)
    assert g[0].shape == e[0].shape, f"row count {len(g[0])} != oracle {len(e[0])}"
    assert np.array_equal(g[0], e[0]), "pair sets differ"
    if report:
        np.testing.assert_allclose(g[1], e[1], rtol=0, atol=DIST_ATOL)
        assert np.array_equal(g[1], e[1]), f"distances not bit-identical (max {np.max(np.abs(g[1] - e[1]))} m)"


@pytest.mark.parametrize("d", [0.0, 5.0, 40.0])
def test_polygon_self_dwithin(ctx, d):
    M = _dw_layer()
    L = ctx.polys(M)
    got = ctx.dwithin(L, L, d, same=True, report_distance=True, upper=True)
    OL = O.Layer.polys(M)
    exp = O.dwithin(OL, OL, d, same=True, upper=True, report=True)
    assert got.stats["candidates"] == exp["candidates"]
    _check(got, exp, True)
    # the same pairs without distances
    _check(ctx.dwithin(L, L, d, same=True, upper=True), exp, False)
    assert (got.dist >= 0).all() and (got.dist <= d + 1e-9).all()


def test_subdivided_and_two_layer_dwithin(ctx):
    M = _dw_layer(40, seed=3)
    S = gsj.subdivide(M, 8)
    Q = _dw_layer(30, seed=4)
    for P, R, same in ((S, S, True), (Q, S, False), (S, Q, False)):
        LP, LR = ctx.polys(P), ctx.polys(R)
        got = ctx.dwithin(LP, LR, 5.0, same=same, report_distance=True, upper=same)
        OP, OR = O.Layer.polys(P), O.Layer.polys(R)
        _check(got, O.dwithin(OP, OR, 5.0, same=same, upper=same, report=True), True)


def test_dwithin_overflow_lists_and_batching(ctx):
    # regions with 500-vertex sides: a shared border puts ~500 edges of each side in one pair's window (> 128)
    R = gsj.generate("regions_grid", nx=40, ny=40, kx=3, ky=3, edge_segs=500, seed=0)
    L = ctx.polys(R)
    OL = O.Layer.polys(R)
    got = ctx.dwithin(L, L, 3.0, same=True, report_distance=True, upper=True)
    _check(got, O.dwithin(OL, OL, 3.0, same=True, upper=True, report=True), True)
    assert got.stats["n_over"] > 0
    # batching and sub-batching do not change the rows
    M = gsj.subdivide(_dw_layer(40, seed=5), 8)
    LM = ctx.polys(M)
    a = ctx.dwithin(LM, LM, 5.0, same=True, report_distance=True, upper=True)
    b = ctx.dwithin(LM, LM, 5.0, same=True, report_distance=True, upper=True, batch_q=97, max_pairs=500)
    ga, gb = sorted_rows(a.pairs, a.dist), sorted_rows(b.pairs, b.dist)
    assert b.stats["batches"] > 3 and np.array_equal(ga[0], gb[0]) and np.array_equal(ga[1], gb[1])


@pytest.mark.parametrize("subdiv", [0, 8])
def test_points_dwithin(ctx, subdiv):
    M = _dw_layer(40, seed=6)
    if subdiv:
        M = gsj.subdivide(M, subdiv)
    xy = gsj.generate_points(n=20000, x0=-95.2, y0=38.6, w=40 * CELL, h=40 * CELL, clusters=10, sigma_m=50.0, seed=2)
    # plus points exactly on vertices and edge midpoints of the layer (on the boundary: distance 0)
    xy = np.concatenate([xy, M.xy[::37], 0.5 * (M.xy[:-1:53] + M.xy[1::53])])
    R = ctx.polys(M)
    pts = ctx.points(xy)
    OP, OR = O.Layer.points_layer(xy), O.Layer.polys(M)
    for d in (0.0, 5.0):
        got = ctx.points_dwithin(pts, R, d, report_distance=True)
        exp = O.points_dwithin(OP, OR, d, report=True)
        assert got.stats["candidates"] == exp["candidates"]
        _check(got, exp, True)
    # with d = 0 every point in a polygon (interior or boundary) is a row; points within the frame's error band of a
    # boundary (~1e-9 m plus the edge-curvature term) are rows too: the frame test d2 <= (d - band)^2 of
    # engine/buffer.cuh:1039 accepts them at d = 0
    pip = {tuple(p) for p in ctx.points_in_polygons(pts, R).pairs.tolist()}
    dw0 = {tuple(p) for p in ctx.points_dwithin(pts, R, 0.0).pairs.tolist()}
    assert pip <= dw0


def test_dwithin_empty_and_errors(ctx):
    M = _dw_layer(5)
    L = ctx.polys(M)
    E = ctx.polys(gsj.HostSet(np.zeros(1, np.uint64), np.zeros(1, np.uint64), np.zeros((0, 2))))
    assert len(ctx.dwithin(E, L, 5.0).pairs) == 0 and len(ctx.dwithin(L, E, 5.0).pairs) == 0
    with pytest.raises(gsj.GsjError):
        ctx.dwithin(L, L, -1.0)
    pts = ctx.points(np.array([[-95.19, 38.61]]))
    with pytest.raises(gsj.GsjError):
        ctx.dwithin(pts, L, 5.0)
