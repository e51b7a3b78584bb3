"""Layer preparation on the device: gsj_dedup_device must return exactly what the host dedup (pinned against the
oracle's restatement of engine/dedup.h:51-124 in the CPU tests) returns: the distinct set in first-occurrence order,
bit for bit, and the member lists in index order."""
import numpy as np
import pytest

import gsj

pytestmark = pytest.mark.gpu


def _with_duplicates(M, seed, extra=0.3):
    """M with a share of its geometries repeated (identical rings and coordinates) at shuffled positions"""
    rng = np.random.default_rng(seed)
    idx = np.concatenate([np.arange(M.n), rng.choice(M.n, int(extra * M.n))])
    rng.shuffle(idx)
    gro, rvo, xy = [0], [0], []
    for g in idx:
        for r in range(int(M.gro[g]), int(M.gro[g + 1])):
            xy.append(M.xy[int(M.rvo[r]):int(M.rvo[r + 1])])
            rvo.append(rvo[-1] + len(xy[-1]))
        gro.append(len(rvo) - 1)
    return gsj.HostSet(np.array(gro), np.array(rvo), np.concatenate(xy))


def _same(a, b):
    (ua, ma, ia), (ub, mb, ib) = a, b
    assert ua.n == ub.n
    assert np.array_equal(ua.gro, ub.gro) and np.array_equal(ua.rvo, ub.rvo)
    assert np.array_equal(ua.xy.view(np.uint64), ub.xy.view(np.uint64))
    assert np.array_equal(ma, mb) and np.array_equal(ia, ib)


@pytest.mark.parametrize("seed", [0, 1])
def test_device_dedup_equals_host_dedup(ctx, seed):
    M = gsj.generate("parcels_mesh", nx=80, ny=80, w=1 / 1549, h=1 / 1549, seed=seed)
    S = _with_duplicates(M, seed)
    _same(gsj.dedup(S, ctx=ctx), gsj.dedup(S))
    u, moff, midx = gsj.dedup(S, ctx=ctx)
    assert u.n == M.n and moff[-1] == S.n   # every source geometry once, every row a member


def test_device_dedup_edge_cases(ctx):
    R = gsj.generate("regions_grid", nx=20, ny=20, kx=2, ky=2, edge_segs=50, seed=0)   # few large polygons
    _same(gsj.dedup(R, ctx=ctx), gsj.dedup(R))
    one = _with_duplicates(gsj.generate("parcels_grid", nx=1, ny=1, seed=0), 3, extra=5.0)   # one geometry, six times
    u, moff, midx = gsj.dedup(one, ctx=ctx)
    assert u.n == 1 and list(moff) == [0, one.n] and list(midx) == list(range(one.n))
    _same((u, moff, midx), gsj.dedup(one))
    # same vertex multiset, different ring split: not duplicates
    P = gsj.generate("parcels_grid", nx=3, ny=1, seed=2)
    _same(gsj.dedup(P, ctx=ctx), gsj.dedup(P))


def _same_pieces(a, b):
    assert a.n == b.n
    for f in ("gro", "rvo", "parent", "syn"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(a.xy.view(np.uint64), b.xy.view(np.uint64))


def _holed_and_clustered():
    """a square with a square hole (many vertices per side), and a ring whose vertices crowd a point (cut 24 times)"""
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


@pytest.mark.parametrize("maxv", [8, 12, 32, 256])
def test_device_subdivide_equals_host_subdivide(ctx, maxv):
    M = gsj.generate("parcels_mesh", nx=60, ny=60, w=1 / 1549, h=1 / 1549, median=14.0, seed=3)
    _same_pieces(gsj.subdivide(M, maxv, ctx=ctx), gsj.subdivide(M, maxv))


def test_device_subdivide_large_holed_and_degenerate(ctx):
    R = gsj.generate("regions_grid", nx=30, ny=30, kx=3, ky=2, edge_segs=200, seed=1)   # ~800-vertex regions
    for maxv in (16, 64, 300):
        _same_pieces(gsj.subdivide(R, maxv, ctx=ctx), gsj.subdivide(R, maxv))
    H = _holed_and_clustered()
    for maxv in (20, 100):
        d, h = gsj.subdivide(H, maxv, ctx=ctx), gsj.subdivide(H, maxv)
        _same_pieces(d, h)
        assert d.n > 2 and d.syn.any()
