"""HIP eroded intersects vs the oracle (CPU restatement) through libgsj's C ABI: polygon x polygon
(ST_Intersects(ST_Buffer(A::geography, -ra), ST_Buffer(B, -rb))) and points in eroded polygons. Bar: row sets
bit-exact (threshold decisions share the oracle's portable math)."""
import numpy as np
import pytest

import gsj
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

CELL = 0.6 / 1095


def _layer(n=40, seed=7):
    return gsj.generate("parcels_mesh", nx=n, ny=n, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=seed)


def _rows(r):
    return {tuple(p) for p in np.asarray(r).reshape(-1, 2).tolist()}


@pytest.mark.parametrize("ra,rb", [(1.0, 1.0), (0.5, 2.5), (0.0, 1.5), (3.0, 0.0), (0.0, 0.0)])
def test_self_erosion_matches_oracle(ctx, ra, rb):
    M = _layer()
    L = ctx.polys(M)
    got = ctx.intersects_buffered(L, L, ra, rb, same=True, upper=True)
    OL = O.Layer.polys(M)
    exp = O.erode(OL, OL, ra, rb, same=True, upper=True)
    assert got.stats["candidates"] == exp["candidates"]
    assert _rows(got.pairs) == _rows(exp["pairs"])
    assert len(got.pairs) == len(_rows(got.pairs))   # each parent pair once


def test_subdivided_need_self_and_two_layers(ctx):
    M = _layer(36, seed=8)
    S = gsj.subdivide(M, 6)
    need = (np.arange(M.n) % 3 == 0).astype(np.uint8)
    LS, OS = ctx.polys(S), O.Layer.polys(S)
    got = ctx.intersects_buffered(LS, LS, 1.0, 1.0, same=True, need_self=need, upper=True)
    exp = O.erode(OS, OS, 1.0, 1.0, same=True, upper=True, need_self=need)
    assert _rows(got.pairs) == _rows(exp["pairs"])
    # subdivision does not change the eroded polygons: same rows as the whole layer
    LM = ctx.polys(M)
    whole = ctx.intersects_buffered(LM, LM, 1.0, 1.0, same=True, need_self=need, upper=True)
    assert _rows(whole.pairs) == _rows(got.pairs)
    Q = _layer(30, seed=9)
    LQ, OQ = ctx.polys(Q), O.Layer.polys(Q)
    g2 = ctx.intersects_buffered(LQ, LS, 0.5, 2.0)
    assert _rows(g2.pairs) == _rows(O.erode(OQ, OS, 0.5, 2.0)["pairs"])


def test_large_polygons_use_global_lists(ctx):
    # config-1 parcels inside 500-vertex-per-side regions: the regions' windows hold hundreds of edges (> 64 in LDS)
    P = gsj.generate("parcels_grid", nx=30, ny=30, seed=0)
    R = gsj.generate("regions_grid", nx=30, ny=30, kx=2, ky=2, edge_segs=500, seed=0)
    LP, LR = ctx.polys(P, build_bvh=False), ctx.polys(R)
    OP, OR = O.Layer.polys(P), O.Layer.polys(R)
    got = ctx.intersects_buffered(LP, LR, 2.0, 3.0)
    exp = O.erode(OP, OR, 2.0, 3.0)
    assert _rows(got.pairs) == _rows(exp["pairs"]) and len(exp["pairs"]) > 0
    RR = ctx.intersects_buffered(LR, LR, 5.0, 5.0, same=True, upper=True)
    assert RR.stats["n_over"] > 0
    assert _rows(RR.pairs) == _rows(O.erode(OR, OR, 5.0, 5.0, same=True, upper=True)["pairs"])


@pytest.mark.parametrize("subdiv", [0, 6])
def test_points_in_eroded_polygons(ctx, subdiv):
    M = _layer(30, seed=10)
    if subdiv:
        M = gsj.subdivide(M, subdiv)
    xy = gsj.generate_points(n=30000, x0=-95.2, y0=38.6, w=30 * CELL, h=30 * CELL, clusters=10, sigma_m=40.0, seed=3)
    xy = np.concatenate([xy, M.xy[::41]])   # vertices: always excluded for r > 0
    pts, R = ctx.points(xy), ctx.polys(M)
    OP, OR = O.Layer.points_layer(xy), O.Layer.polys(M)
    for r in (0.0, 0.5, 2.0, 5.0):
        got = ctx.points_intersects_buffered(pts, R, r)
        exp = O.points_erode(OP, OR, r)
        assert got.stats["candidates"] == exp["candidates"]
        assert _rows(got.pairs) == _rows(exp["pairs"]), r
    # shrinking polygons only removes points
    assert _rows(ctx.points_intersects_buffered(pts, R, 5.0).pairs) <= _rows(ctx.points_intersects_buffered(pts, R, 0.5).pairs)


def test_erosion_errors(ctx):
    """erosion radii are >= 0: a negative one (a grown side, i.e. buffers of opposite sign) is refused, as the
    reference refuses opposite-sign buffers (daemon/gpujoind.cu:406-411); two grown sides are a distance join"""
    L = ctx.polys(_layer(5))
    for ra, rb in ((-1.0, 1.0), (1.0, -0.5), (1e5, 1.0)):
        with pytest.raises(gsj.GsjError, match=r"radii must be in \[0, 1e5\) metres"):
            ctx.intersects_buffered(L, L, ra, rb)
    pts = ctx.points(np.array([[-95.19, 38.61]]))
    with pytest.raises(gsj.GsjError):
        ctx.intersects_buffered(pts, L, 1.0, 1.0)


@pytest.mark.parametrize("lat,lon,expect", [(75.0, 20.0, True), (34.0, -117.0, False), (-78.0, 160.0, True)])
def test_polar_erosion_takes_scale_one(ctx, lat, lon, expect):
    """the engine's eroded intersects at polar latitudes (scale 1, as the reference) and on a UTM central meridian
    (scale 0.9996): rectangles overlapping by 1.0002 (ra + rb) metres meet only where the scale is 1; rows equal the
    oracle's (tests/test_erosion.py pins the oracle by the closed form)"""
    from test_erosion import _rects_at

    ra, rb = 2.0, 1.0
    hs = _rects_at(lat, lon, (ra + rb) * 1.0002)
    L = ctx.polys(hs)
    got = ctx.intersects_buffered(L, L, ra, rb, same=True, upper=True)
    OL = O.Layer.polys(hs)
    exp = O.erode(OL, OL, ra, rb, same=True, upper=True)
    assert _rows(got.pairs) == _rows(exp["pairs"])
    assert ((0, 1) in _rows(got.pairs)) == expect
