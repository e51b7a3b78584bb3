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
