"""GPU: gsj_join_noalloc (fixed scratch, no allocation) returns the same rows as gsj_join, bit for bit, on two calls in
a row (the second smaller, so it runs over the first call's leftover scratch contents), for a subdivided relate with
areas and a two-layer intersects join; sizes beyond its limits fail with a clear error."""
import numpy as np
import pytest

import gsj
from conftest import sorted_rows

pytestmark = pytest.mark.gpu


def _mesh(n, seed, **kw):
    return gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=seed, **kw)


def _same(a, b, area, code):
    ga = sorted_rows(a.pairs, a.area if area else None, a.code if code else None)
    gb = sorted_rows(b.pairs, b.area if area else None, b.code if code else None)
    assert ga[0].shape == gb[0].shape and len(ga[0]) > 0
    assert np.array_equal(ga[0], gb[0])
    if area:
        assert np.array_equal(ga[1], gb[1])
    if code:
        assert np.array_equal(ga[2], gb[2])
    assert a.stats["candidates"] == b.stats["candidates"] and a.stats["hits"] == b.stats["hits"]


def test_noalloc_join_equals_join_on_two_calls(ctx):
    big = gsj.subdivide(_mesh(220, seed=5, median=20.0), 256)   # ~48k parcels: several sub-batches, overflow pairs
    small = gsj.subdivide(_mesh(60, seed=6), 16)
    for hs in (big, small):
        L = ctx.polys(hs)
        kw = dict(want_area=True, want_relate=True, upper=True)
        ref = ctx.join(L, L, **kw)
        got = ctx.join(L, L, noalloc=True, **kw)
        _same(got, ref, area=True, code=True)
        if hs is big:
            assert got.stats["batches"] > 1 and got.stats["n_over"] > 0
    P = gsj.generate("parcels_grid", nx=80, ny=80, seed=0)
    R = gsj.generate("regions_grid", nx=80, ny=80, kx=4, ky=4, edge_segs=500, seed=0)
    dP, dR = ctx.polys(P, build_bvh=False), ctx.polys(R)
    _same(ctx.join(dP, dR, noalloc=True, want_area=True), ctx.join(dP, dR, want_area=True), area=True, code=False)


def test_noalloc_join_rejects_sizes_beyond_its_limits(ctx):
    L = ctx.polys(_mesh(10, seed=1))
    for kw in (dict(max_pairs=1 << 20), dict(batch_q=1 << 20)):
        with pytest.raises(gsj.GsjError) as e:
            ctx.join(L, L, noalloc=True, **kw)
        assert e.value.code == -4 and "exceeds" in str(e.value)
    with pytest.raises(gsj.GsjError) as e:
        ctx.join(L, L, noalloc=True, window=(-96, 38, -95, 39))
    assert e.value.code == -1
