"""GPU: the launch-shape table. Joins, windowed joins and distance joins give identical rows under the built-in
table and under tables that move pairs between work classes (no tiny or medium class, other tiny caps), change the
screens' register budget, the grid caps and the traversal block size, with a table whose buckets switch shapes
inside one workload; the extreme table is also checked against the oracle. A malformed table is refused and leaves
the one in effect; schedulers launch with their context's table."""
import numpy as np
import pytest

import gsj
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

EXTREME = ("bucket np>=0 tcap=0 mcap=0 screen_occ=4 screen_bpc=1 edge_bpc=1 mixed_bpc=1 large_bpc=1 dw_bpc=1 trav_block=64 ovf_lds=0 fork=0 mixed=0 large_cap=0 size_sort=0 dw_size_sort=1\n")
SWITCHING = ("bucket np>=0 tcap=8 mcap=48 screen_occ=4 screen_bpc=2 edge_bpc=64 trav_block=128\n"
             "bucket np>=2000 tcap=12 mcap=0 screen_occ=5 screen_bpc=64 edge_bpc=3 mixed_bpc=5 large_bpc=7 dw_bpc=2 large_cap=64 mixed=0 ovf_lds=2\n"
             "bucket np>=20000 tcap=16 mcap=64\n")


@pytest.fixture(scope="module")
def layer(ctx):
    # median 24 stored vertices, subdivided at 128: pairs in every class (tiny, small, medium, mixed, large)
    M = gsj.subdivide(gsj.generate("parcels_mesh", nx=80, ny=80, w=1 / 1549, h=1 / 1549, seed=8, median=24.0), 128)
    return M, ctx.polys(M)


def run_all(ctx, L, M):
    x0, y0 = M.xy.min(axis=0)
    x1, y1 = M.xy.max(axis=0)
    w = (x0 + 0.3 * (x1 - x0), y0 + 0.2 * (y1 - y0), x0 + 0.6 * (x1 - x0), y0 + 0.45 * (y1 - y0))
    out = [ctx.join(L, L, want_area=True, want_relate=True, upper=True),
           ctx.join(L, L, want_area=True, want_relate=True, upper=True, window=w),
           ctx.join(L, L, want_area=True, upper=True, batch_q=700),   # sub-batches of ~2-3k pairs: the middle bucket
           ctx.dwithin(L, L, 7.5, same=True, upper=True, report_distance=True),
           ctx.dwithin(L, L, 3.0, same=True, upper=True, window=w)]
    return out


def same_rows(a, b):
    ga = sorted_rows(a.pairs, a.area, a.code, a.dist)
    gb = sorted_rows(b.pairs, b.area, b.code, b.dist)
    for x, y in zip(ga, gb):
        if (x is None) != (y is None):
            return False
        if x is not None and not np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)):
            return False
    return True


def test_rows_are_identical_under_every_table(ctx, layer):
    M, L = layer
    before = ctx.tuning
    ctx.set_tuning("bucket np>=0\n")   # the built-in shapes
    try:
        base = run_all(ctx, L, M)
        st = base[0].stats
        assert min(st["n_tiny"], st["n_small"], st["n_medium"], st["n_mixed"] + st["n_large"]) > 0, st
        for table in (EXTREME, SWITCHING):
            ctx.set_tuning(table)
            assert ctx.tuning.splitlines()[1].startswith("bucket np>=0 tcap=" + table.split("tcap=")[1].split()[0])
            got = run_all(ctx, L, M)
            for i, (g, b) in enumerate(zip(got, base)):
                assert g.stats["candidates"] == b.stats["candidates"], (table, i)
                assert same_rows(g, b), (table, i)
            if table is EXTREME:
                s = got[0].stats
                assert s["n_tiny"] == 0 and s["n_medium"] == 0 and s["n_small"] > 0, s
    finally:
        ctx.set_tuning("\n".join(l for l in before.splitlines() if not l.startswith("#")) + "\n")


def test_extreme_table_matches_the_oracle(ctx, layer):
    M, L = layer
    before = ctx.tuning
    ctx.set_tuning(EXTREME)
    try:
        got = ctx.join(L, L, want_area=True, want_relate=True, upper=True)
    finally:
        ctx.set_tuning("\n".join(l for l in before.splitlines() if not l.startswith("#")) + "\n")
    OM = O.Layer.polys(M)
    exp = O.join(OM, OM, want_area=True, want_relate=True, upper=True)
    g = sorted_rows(got.pairs, got.area, got.code)
    e = sorted_rows(exp["pairs"], exp["area"], exp["code"])
    assert np.array_equal(g[0], e[0]) and np.array_equal(g[2], e[2])
    assert np.array_equal(g[1].view(np.uint64), e[1].view(np.uint64)), "areas not bit-identical"


def _circles(k=6, seed=5):
    """k x k overlapping circles of 120 to 600 vertices: large-class pairs whose lists in the overlap box run from a
    few dozen to several hundred edges (past the 96 of the large class, the 256 of the list classes' own overflow
    launch and the 448 of the LDS overflow path)"""
    rng = np.random.default_rng(seed)
    rings = []
    for i in range(k * k):
        n = (120, 200, 300, 420, 600)[i % 5]
        t = rng.uniform(0, 2 * np.pi) + np.arange(n) * (2 * np.pi / n)
        r = 1e-3 * rng.uniform(0.9, 1.1)
        ring = np.c_[-95 + 1.3e-3 * (i % k) + r * np.cos(t), 38 + 1.3e-3 * (i // k) + r * np.sin(t)]
        rings.append(np.vstack([ring, ring[:1]]))
    xy = np.concatenate(rings)
    rvo = np.concatenate([[0], np.cumsum([len(r) for r in rings])])
    return gsj.HostSet(np.arange(k * k + 1), rvo, xy)


@pytest.mark.parametrize("fork", [1, 0])
def test_list_class_overflow_on_its_own_chain(ctx, fork):
    """ovf_lds=2: the list classes' overflow pairs are refined right behind the large kernel (queue length read on the
    device), the pairs whose lists pass 256 edges handed on to the host-read queue. Rows equal the oracle's and the
    built-in shapes', and every overflow pair is counted once."""
    M = _circles()
    L = ctx.polys(M)
    before = ctx.tuning
    try:
        ctx.set_tuning("bucket np>=0\n")
        base = ctx.join(L, L, want_area=True, want_relate=True)
        for cap in (96, 0):
            ctx.set_tuning(f"bucket np>=0 ovf_lds=2 fork={fork} large_cap={cap}\n")
            got = ctx.join(L, L, want_area=True, want_relate=True)
            assert same_rows(got, base), cap
            n = got.stats["n_over"]
            assert (n == base.stats["n_over"] if cap else n <= got.stats["n_large"]) and n > 0, (cap, got.stats)
    finally:
        ctx.set_tuning("\n".join(l for l in before.splitlines() if not l.startswith("#")) + "\n")
    OM = O.Layer.polys(M)
    exp = O.join(OM, OM, want_area=True, want_relate=True)
    g = sorted_rows(got.pairs, got.area, got.code)
    e = sorted_rows(exp["pairs"], exp["area"], exp["code"])
    assert np.array_equal(g[0], e[0]) and np.array_equal(g[2], e[2])
    assert np.array_equal(g[1].view(np.uint64), e[1].view(np.uint64)), "areas not bit-identical"


def test_shipped_table_is_read_at_creation_and_gives_the_same_rows(ctx, layer):
    """a new context reads postgis-cuda_amd/tuning/<arch>.txt (when present) and its rows equal the built-in shapes'"""
    import glob
    import os

    M, L = layer
    shipped = glob.glob(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "postgis-cuda_amd", "tuning", "gfx950.txt"))
    c2 = gsj.Context(0)
    try:
        src = c2.tuning.splitlines()[0]
        if shipped:
            assert src.startswith("# source: ") and src.endswith("tuning/gfx950.txt"), src
        else:
            assert src == "# source: built-in"
        L2 = c2.polys(M)
        got = run_all(c2, L2, M)
    finally:
        c2.close()
    ctx.set_tuning("bucket np>=0\n")
    base = run_all(ctx, L, M)
    for i, (g, b) in enumerate(zip(got, base)):
        assert same_rows(g, b), i


def test_malformed_table_is_refused_and_the_table_kept(ctx):
    before = ctx.tuning
    with pytest.raises(gsj.GsjError) as ei:
        ctx.set_tuning("bucket np>=0 tcap=99\n")
    assert "tcap outside" in str(ei.value)
    assert ctx.tuning == before


def test_scheduler_lanes_use_the_context_table(ctx, layer):
    M, L = layer
    before = ctx.tuning
    ctx.set_tuning(EXTREME)
    try:
        s = gsj.Scheduler(ctx, lanes=2, batch=1)
        try:
            h = s.submit(gsj.REQ_JOIN, L, L, want_area=True, want_relate=True, upper=True)
            r = s.wait(h)
        finally:
            s.close()
    finally:
        ctx.set_tuning("\n".join(l for l in before.splitlines() if not l.startswith("#")) + "\n")
    assert r.stats["n_tiny"] == 0 and r.stats["n_medium"] == 0, r.stats
    direct = ctx.join(L, L, want_area=True, want_relate=True, upper=True)
    assert same_rows(r, direct)
