"""HIP engine vs the oracle (CPU restatement) through libgsj's C ABI: layer preparation, candidates, intersects,
intersection areas, DE-9IM, piece -> parent aggregation, batching, points in polygons, clustering.

Bar: pairs, flags and DE-9IM codes bit-exact; areas bit-exact as well (the engine adds every sum in the same order as
the oracle), with the 1e-9 relative tolerance of the metric kept as the documented fallback bound."""
import numpy as np
import pytest

import gsj
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

AREA_RTOL = 1e-9   # documented bound; the engine is expected to match to the bit


def _cfg1(n=100, segs=500, k=4):
    P = gsj.generate("parcels_grid", nx=n, ny=n, seed=0)
    R = gsj.generate("regions_grid", nx=n, ny=n, kx=k, ky=k, edge_segs=segs, seed=0)
    return P, R


def _mesh(n=120, seed=1, **kw):
    return gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=seed, **kw)


def _check_rows(got, exp, area=False, code=False):
    g = sorted_rows(got.pairs, got.area if area else None, got.code if code else None)
    e = sorted_rows(exp["pairs"], exp.get("area") if area else None, exp.get("code") if code else None)
    assert g[0].shape == e[0].shape, f"row count {len(g[0])} != oracle {len(e[0])}"
    assert np.array_equal(g[0], e[0]), "pair sets differ"
    if area:
        np.testing.assert_allclose(g[1], e[1], rtol=AREA_RTOL, atol=0)
        assert np.array_equal(g[1], e[1]), f"areas not bit-identical (max rel {np.max(np.abs(g[1] - e[1]) / np.maximum(np.abs(e[1]), 1e-300))})"
    if code:
        assert np.array_equal(g[2], e[2]), "DE-9IM codes differ"


def test_layer_prep_matches_oracle(ctx):
    P = gsj.subdivide(_mesh(60), 16)
    d = ctx.polys(P, build_bvh=False).export()
    o = O.Layer.polys(P).export()
    for k in ("origin", "ext", "bb", "q"):
        assert np.array_equal(d[k], o[k]), k
    np.testing.assert_allclose(d["blen"], o["blen"], rtol=1e-15)


def test_points_prep_matches_oracle(ctx):
    xy = gsj.generate_points(n=20000, x0=-95.5, y0=38.5, w=0.05, h=0.05, clusters=20, seed=2)
    assert np.array_equal(ctx.points(xy).export()["bb"], O.Layer.points_layer(xy).export()["bb"])


def test_cfg1_intersects_area(ctx):
    P, R = _cfg1()
    got = ctx.join(ctx.polys(P, build_bvh=False), ctx.polys(R), want_area=True)
    exp = O.join(O.Layer.polys(P), O.Layer.polys(R), want_area=True)
    assert got.stats["candidates"] == exp["candidates"]
    _check_rows(got, exp, area=True)
    assert len(got.pairs) == P.n   # every parcel lies in exactly one region


@pytest.mark.parametrize("upper", [False, True])
def test_mesh_self_relate(ctx, upper):
    M = _mesh(120)
    L = ctx.polys(M)
    got = ctx.join(L, L, want_area=True, want_relate=True, upper=upper)
    OM = O.Layer.polys(M)
    exp = O.join(OM, OM, want_area=True, want_relate=True, upper=upper)
    assert got.stats["candidates"] == exp["candidates"]
    _check_rows(got, exp, area=True, code=True)


@pytest.mark.parametrize("want_area,want_relate", [(True, True), (True, False), (False, True)])
def test_every_work_class_matches_oracle(ctx, want_area, want_relate):
    # median 24 stored vertices: pairs in the 8-lane (<= 16), 16-lane (<= 32), 32-lane (<= 64) and wave (larger) classes
    M = _mesh(70, seed=8, median=24.0)
    L = ctx.polys(M)
    got = ctx.join(L, L, want_area=want_area, want_relate=want_relate, upper=True)
    assert min(got.stats["n_small"], got.stats["n_medium"], got.stats["n_large"]) > 1000 and got.stats["n_tiny"] > 100, got.stats
    OM = O.Layer.polys(M)
    exp = O.join(OM, OM, want_area=want_area, want_relate=want_relate, upper=True)
    assert got.stats["candidates"] == exp["candidates"]
    _check_rows(got, exp, area=want_area, code=want_relate)


@pytest.mark.parametrize("maxv", [8, 64])
def test_mesh_subdivided_relate(ctx, maxv):
    M = gsj.subdivide(_mesh(100, seed=5), maxv)
    assert M.parent is not None and M.n > 100 * 100 or maxv == 64
    L = ctx.polys(M)
    got = ctx.join(L, L, want_area=True, want_relate=True, upper=True)
    OM = O.Layer.polys(M)
    exp = O.join(OM, OM, want_area=True, want_relate=True, upper=True)
    assert got.stats["candidates"] == exp["candidates"]
    _check_rows(got, exp, area=True, code=True)


def _holed_and_crowded():
    """a square with a square hole (120 vertices per side), and a ring whose vertices crowd a point"""
    t = np.linspace(0, 1, 121)[:-1]
    outer = np.concatenate([np.c_[t, 0 * t], np.c_[1 + 0 * t, t], np.c_[1 - t, 1 + 0 * t], np.c_[0 * t, 1 - t]])
    inner = 0.3 + 0.4 * outer[::-1]
    a = np.linspace(0, 2 * np.pi, 400, endpoint=False)
    crowd = np.c_[0.3 + 1e-9 * np.cos(a), 0.2 + 1e-9 * np.sin(a)]
    spiky = np.concatenate([[[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]], crowd])
    rings = [np.vstack([outer, outer[:1]]), np.vstack([inner, inner[:1]]), np.vstack([spiky, spiky[:1]])]
    xy = np.concatenate(rings) + np.array([-95.0, 38.0])
    rvo = np.concatenate([[0], np.cumsum([len(r) for r in rings])])
    return gsj.HostSet(np.array([0, 2, 3]), rvo, xy)


@pytest.mark.parametrize("maxv", [12, 32])
def test_subdivided_rings_relate(ctx, maxv):
    # pieces with synthetic edges and cut points on outer and inner rings, so the per-vertex synthetic / cut flags
    # wrap around every ring, joined with a parcel mesh over the same square and with themselves
    H = gsj.subdivide(_holed_and_crowded(), maxv)
    M = gsj.generate("parcels_mesh", nx=24, ny=24, x0=-95.0, y0=38.0, w=1 / 24, h=1 / 24, seed=3)
    dH, dM = ctx.polys(H), ctx.polys(M)
    oH, oM = O.Layer.polys(H), O.Layer.polys(M)
    for dP, dR, oP, oR, kw in ((dM, dH, oM, oH, {}), (dH, dH, oH, oH, {"upper": True})):
        got = ctx.join(dP, dR, want_area=True, want_relate=True, **kw)
        exp = O.join(oP, oR, want_area=True, want_relate=True, **kw)
        assert got.stats["candidates"] == exp["candidates"] > 0
        _check_rows(got, exp, area=True, code=True)


def test_intersects_only_and_regions_relate(ctx):
    P, R = _cfg1(60, segs=300, k=3)
    R = gsj.subdivide(R, 64)
    dP, dR = ctx.polys(P, build_bvh=False), ctx.polys(R)
    oP, oR = O.Layer.polys(P), O.Layer.polys(R)
    _check_rows(ctx.join(dP, dR), O.join(oP, oR, want_area=False))
    _check_rows(ctx.join(dP, dR, want_relate=True), O.join(oP, oR, want_area=False, want_relate=True), code=True)


def test_batching_is_invisible(ctx):
    M = gsj.subdivide(_mesh(80, seed=7), 12)
    L = ctx.polys(M)
    ref = ctx.join(L, L, want_area=True, want_relate=True, upper=True)
    small = ctx.join(L, L, want_area=True, want_relate=True, upper=True, batch_q=333, max_pairs=1000)
    assert small.stats["batches"] > 3
    _check_rows(small, {"pairs": ref.pairs, "area": ref.area, "code": ref.code}, area=True, code=True)


@pytest.mark.parametrize("subdiv", [0, 40])
def test_row_copy_paths_agree(ctx, subdiv):
    """Sub-batches of at most 1M candidates copy their rows in one host wait (the copy sized by the candidate count);
    larger ones read the row count first. A layer of more than 1M candidates run as one sub-batch and as sub-batches
    of 2^19 gives the same rows, areas and codes (subdivided: both fold pieces into parents, at both sizes)."""
    M = _mesh(560, seed=11)
    if subdiv:
        M = gsj.subdivide(M, subdiv)
    L = ctx.polys(M)
    whole = ctx.join(L, L, want_area=True, want_relate=True, upper=True)
    assert whole.stats["candidates"] > (1 << 20) and whole.stats["batches"] == 1
    parts = ctx.join(L, L, want_area=True, want_relate=True, upper=True, max_pairs=1 << 19)
    assert parts.stats["batches"] > 2
    _check_rows(parts, {"pairs": whole.pairs, "area": whole.area, "code": whole.code}, area=True, code=True)


def test_points_in_polygons(ctx):
    M = _mesh(80, seed=3)
    x0, y0 = -95.5, 38.5
    xy = gsj.generate_points(n=50000, x0=x0, y0=y0, w=80 / 1549, h=80 / 1549, clusters=30, sigma_m=30.0, seed=2)
    # points exactly on vertices and on edge midpoints of the layer: boundary cases
    xy = np.concatenate([xy, M.xy[::7], 0.5 * (M.xy[:-1:5] + M.xy[1::5])])
    pts = ctx.points(xy)
    for layer in (M, gsj.subdivide(M, 10)):
        got = ctx.points_in_polygons(pts, ctx.polys(layer))
        exp = O.points_join(O.Layer.points_layer(xy), O.Layer.polys(layer))
        g = sorted_rows(got.pairs, got.where)
        e = sorted_rows(exp["pairs"], exp["where"])
        assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1])
        assert (e[1] == 2).sum() > 0   # boundary cases exercised


def test_cluster(ctx):
    for M in (_mesh(90, seed=11), gsj.subdivide(_mesh(50, seed=12), 10)):
        L = ctx.polys(M)
        root, st = ctx.cluster(L)
        exp = O.cluster(O.Layer.polys(M))
        assert np.array_equal(root, exp["root"])


def test_empty_and_tiny_layers(ctx):
    one = gsj.HostSet(np.array([0, 1]), np.array([0, 5]), np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0, 0]], float))
    L1 = ctx.polys(one)
    r = ctx.join(L1, L1, want_area=True, want_relate=True)
    assert r.pairs.tolist() == [[0, 0]] and r.area[0] == 1.0 and gsj.de9im_str(r.code[0]) == "2FFF1FFF2"
    empty = gsj.HostSet(np.array([0]), np.array([0]), np.zeros((0, 2)))
    L0 = ctx.polys(empty)
    assert len(ctx.join(L0, L1).pairs) == 0 and len(ctx.join(L1, L0).pairs) == 0
