"""Served path (continuous batching over lanes) through libgsj's C ABI: every request's rows are the rows of the same
direct call, alone and in a mix that overlaps on the device; layers stay valid while requests hold them; errors
stay with their request."""
import numpy as np
import pytest

import gsj
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

CELL = 0.6 / 1095


def _mesh(n, seed):
    return gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=seed)


def _same(a, b):
    ka = sorted_rows(a.pairs, a.area, a.code, a.where, a.dist)
    kb = sorted_rows(b.pairs, b.area, b.code, b.where, b.dist)
    for x, y in zip(ka, kb):
        assert (x is None) == (y is None)
        if x is not None:
            assert np.array_equal(x, y)


def _workload(ctx):
    M1 = ctx.polys(_mesh(60, 1))
    M2 = ctx.polys(gsj.subdivide(_mesh(50, 2), 10))
    P = ctx.polys(gsj.generate("parcels_grid", nx=40, ny=40, seed=0), build_bvh=False)
    R = ctx.polys(gsj.generate("regions_grid", nx=40, ny=40, kx=2, ky=2, edge_segs=200, seed=0))
    D = ctx.polys(gsj.generate("parcels_mesh", nx=40, ny=40, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=2))
    pts = ctx.points(gsj.generate_points(n=20000, x0=-95.5, y0=38.5, w=60 / 1549, h=60 / 1549, clusters=20, seed=2))
    reqs = [
        (gsj.REQ_JOIN, M1, M1, dict(want_relate=True, upper=True)),
        (gsj.REQ_JOIN, P, R, dict(want_area=True)),
        (gsj.REQ_JOIN, M2, M2, dict(want_area=True, want_relate=True, upper=True)),
        (gsj.REQ_DWITHIN, D, D, dict(d=5.0, same=True, report_distance=True, upper=True)),
        (gsj.REQ_POINTS_IN_POLYGONS, pts, M1, {}),
        (gsj.REQ_POINTS_DWITHIN, pts, M2, dict(d=3.0, report_distance=True)),
        (gsj.REQ_JOIN, M2, M1, dict(want_area=True)),
        (gsj.REQ_CLUSTER, M2, None, {}),
        (gsj.REQ_INTERSECTS_BUFFERED, D, D, dict(ra=1.0, rb=1.5, same=True, upper=True)),
        (gsj.REQ_POINTS_INTERSECTS_BUFFERED, pts, M2, dict(rb=2.0)),
    ]
    return reqs, [M1, M2, P, R, D, pts]


def _direct(ctx, kind, A, B, kw):
    kw = dict(kw)
    if kind == gsj.REQ_JOIN:
        return ctx.join(A, B, **kw)
    if kind == gsj.REQ_POINTS_IN_POLYGONS:
        return ctx.points_in_polygons(A, B, **kw)
    if kind == gsj.REQ_DWITHIN:
        return ctx.dwithin(A, B, kw.pop("d"), **kw)
    if kind == gsj.REQ_POINTS_DWITHIN:
        return ctx.points_dwithin(A, B, kw.pop("d"), **kw)
    if kind == gsj.REQ_INTERSECTS_BUFFERED:
        return ctx.intersects_buffered(A, B, kw.pop("ra"), kw.pop("rb"), **kw)
    if kind == gsj.REQ_POINTS_INTERSECTS_BUFFERED:
        return ctx.points_intersects_buffered(A, B, kw.pop("rb"), **kw)
    root, st = ctx.cluster(A)
    return gsj.Rows(np.zeros((0, 2), np.uint32), None, None, None, st, None, root)


@pytest.mark.parametrize("lanes,budget", [(4, 0), (3, 1 << 20)])
def test_mix_equals_direct_calls(ctx, lanes, budget):
    reqs, _layers = _workload(ctx)
    ref = [_direct(ctx, k, A, B, kw) for k, A, B, kw in reqs]
    s = gsj.Scheduler(ctx, lanes=lanes, scratch_budget=budget)
    try:
        for rnd in range(2):   # every request twice, all in flight together
            tickets = [s.submit(k, A, B, **kw) for k, A, B, kw in reqs] + [s.submit(k, A, B, **kw) for k, A, B, kw in reqs]
            got = [s.wait(t) for t in tickets]
            for i, g in enumerate(got):
                r = ref[i % len(reqs)]
                if reqs[i % len(reqs)][0] == gsj.REQ_CLUSTER:
                    assert np.array_equal(g.root, r.root)
                else:
                    _same(g, r)
                assert g.stats["times"]["total_ms"] >= g.stats["times"]["run_ms"] >= 0
            lanes_used = {g.stats["times"]["lane"] for g in got}
            # no budget: requests spread over the lanes; a budget below every estimate admits one request at a time
            assert len(lanes_used) > 1 if budget == 0 else max(g.stats["times"]["queue_ms"] for g in got) > 0
    finally:
        s.close()


def test_alone_and_in_the_mix_are_identical(ctx):
    reqs, _layers = _workload(ctx)
    s = gsj.Scheduler(ctx, lanes=4)
    try:
        alone = []
        for k, A, B, kw in reqs:
            alone.append(s.wait(s.submit(k, A, B, **kw)))
        tickets = [s.submit(k, A, B, **kw) for _ in range(3) for k, A, B, kw in reqs]
        for i, t in enumerate(tickets):
            g, a = s.wait(t), alone[i % len(reqs)]
            if reqs[i % len(reqs)][0] == gsj.REQ_CLUSTER:
                assert np.array_equal(g.root, a.root)
            else:
                _same(g, a)
    finally:
        s.close()


def test_layers_outlive_their_owner_handle_and_errors_stay_with_the_request(ctx):
    M = ctx.polys(_mesh(40, 5))
    ref = ctx.join(M, M, want_area=True, upper=True)
    s = gsj.Scheduler(ctx, lanes=2)
    try:
        tickets = [s.submit(gsj.REQ_JOIN, M, M, want_area=True, upper=True) for _ in range(4)]
        M.free()   # the queued requests still hold the layer
        for t in tickets:
            _same(s.wait(t), ref)
        N = ctx.polys(_mesh(10, 6), build_bvh=False)
        bad = s.submit(gsj.REQ_JOIN, N, N)   # build layer without a BVH: this request fails, the lanes keep going
        good = s.submit(gsj.REQ_JOIN, N, ctx.polys(_mesh(10, 6)))
        with pytest.raises(gsj.GsjError):
            s.wait(bad)
        assert len(s.wait(good).pairs) > 0
    finally:
        s.close()


def test_done_callback_once_per_request_before_wait(ctx):
    """gsj_request.done_cb: called once per request, on a lane's thread, before wait returns for it; a request that
    fails on its lane gets it too"""
    import threading
    M = ctx.polys(_mesh(30, 7))
    N = ctx.polys(_mesh(10, 6), build_bvh=False)
    calls, lock, main = {}, threading.Lock(), threading.get_ident()

    def mark(i):
        def f():
            with lock:
                calls.setdefault(i, []).append(threading.get_ident())
        return f

    s = gsj.Scheduler(ctx, lanes=2)
    try:
        good = [s.submit(gsj.REQ_JOIN, M, M, upper=True, on_done=mark(i)) for i in range(6)]
        bad = s.submit(gsj.REQ_JOIN, N, N, on_done=mark("bad"))
        for i, t in enumerate(good):
            r = s.wait(t)
            assert len(calls.get(i, [])) == 1 and calls[i][0] != main and len(r.pairs) > 0, (i, calls.get(i))
        with pytest.raises(gsj.GsjError):
            s.wait(bad)
        assert len(calls["bad"]) == 1
    finally:
        s.close()


def _rows_in_order(a, b):
    """identical rows in identical order (batching keeps every request's own row order)"""
    assert np.array_equal(a.pairs, b.pairs)
    for f in ("area", "code"):
        x, y = getattr(a, f), getattr(b, f)
        if x is None or y is None:   # an empty result may carry no column
            assert (x is None or len(x) == 0) and (y is None or len(y) == 0) or (x is None and y is None)
        else:
            assert np.array_equal(x, y)


def test_batched_windowed_joins_equal_their_own_calls(ctx):
    """Windowed joins queued behind a long request run as batches (times["batch"] > 1); each one's rows, in order,
    and its candidates and hits equal its direct call, across area / relate groups, overlapping, empty, point-sized and
    layer-covering windows, and tiny probe batches and sub-batches that cut through the requests."""
    M = _mesh(70, 7)
    L = ctx.polys(M)
    big = ctx.polys(gsj.subdivide(_mesh(150, 8), 10))
    x0, y0 = M.xy[:, 0].min(), M.xy[:, 1].min()
    W, H = M.xy[:, 0].max() - x0, M.xy[:, 1].max() - y0
    rng = np.random.default_rng(11)
    wins = []
    for i in range(44):
        f = rng.uniform(0, 1, 2)
        s = rng.uniform(0.02, 0.4)
        wins.append((x0 + f[0] * W, y0 + f[1] * H, x0 + f[0] * W + s * W, y0 + f[1] * H + s * H))
    wins[3] = (x0 - 1, y0 - 1, x0 + W + 1, y0 + H + 1)                 # every probe
    wins[7] = (x0 + W + 1, y0 + H + 1, x0 + W + 2, y0 + H + 2)         # none
    wins[12] = (x0 + W / 2, y0 + H / 2, x0 + W / 2, y0 + H / 2)        # a point
    wins[20] = wins[21]                                                # twice the same window in one batch
    kws = []
    for i, w in enumerate(wins):
        kw = dict(window=w, upper=True)
        if i % 3 == 0:
            kw["want_area"] = True
        elif i % 3 == 1:
            kw["want_relate"] = True
        if i >= 36:
            kw.update(batch_q=5, max_pairs=40)
        kws.append(kw)
    s = gsj.Scheduler(ctx, lanes=1, batch=16)
    try:
        t0 = s.submit(gsj.REQ_JOIN, big, big, want_relate=True, want_area=True, upper=True)   # holds the lane
        tickets = [s.submit(gsj.REQ_JOIN, L, L, **kw) for kw in kws]
        s.wait(t0)
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    batched = [g.stats["times"]["batch"] for g in got]
    assert max(batched) > 1, "no request ran in a batch"
    assert sum(b > 1 for b in batched) >= len(got) // 2
    for kw, g in zip(kws, got):
        d = ctx.join(L, L, **kw)
        _rows_in_order(g, d)
        assert g.stats["candidates"] == d.stats["candidates"] and g.stats["hits"] == d.stats["hits"]
    assert len(got[3].pairs) > 0 and len(got[7].pairs) == 0


def test_batching_off_and_unbatchable_requests_run_alone(ctx):
    M = _mesh(30, 9)
    L = ctx.polys(M)
    S = ctx.polys(gsj.subdivide(_mesh(30, 10), 8))
    x0, y0 = M.xy[:, 0].min(), M.xy[:, 1].min()
    w = (x0, y0, x0 + 0.01, y0 + 0.01)
    for batch, reqs in ((1, [(L, dict(window=w, upper=True))] * 6),
                        (16, [(L, dict(upper=True))] * 3 + [(S, dict(window=w, upper=True))] * 3)):
        s = gsj.Scheduler(ctx, lanes=1, batch=batch)
        try:
            got = [s.wait(t) for t in [s.submit(gsj.REQ_JOIN, A, A, **kw) for A, kw in reqs]]
        finally:
            s.close()
        assert all(g.stats["times"]["batch"] == 1 for g in got)
        for (A, kw), g in zip(reqs, got):
            _rows_in_order(g, ctx.join(A, A, **kw))
    with pytest.raises(gsj.GsjError, match="batch"):
        gsj.Scheduler(ctx, lanes=1, batch=33)


def test_batched_windowed_dwithin_equal_their_own_calls(ctx):
    """Windowed distance joins of one layer pair batch across different distances (each member's search box and each
    candidate's threshold are its own request's); rows, order, distances, candidates and hits equal the direct calls."""
    M = gsj.generate("parcels_mesh", nx=60, ny=60, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=12)
    L = ctx.polys(M)
    big = ctx.polys(gsj.subdivide(_mesh(150, 8), 10))
    x0, y0 = M.xy[:, 0].min(), M.xy[:, 1].min()
    W, H = M.xy[:, 0].max() - x0, M.xy[:, 1].max() - y0
    rng = np.random.default_rng(13)
    reqs = []
    for i in range(40):
        f = rng.uniform(0, 1, 2)
        s = rng.uniform(0.02, 0.35)
        w = (x0 + f[0] * W, y0 + f[1] * H, x0 + f[0] * W + s * W, y0 + f[1] * H + s * H)
        d = 0.0 if i % 10 == 0 else float(rng.uniform(1.0, 40.0))
        reqs.append(dict(d=d, same=True, upper=True, report_distance=i < 30, window=w))
        if 20 <= i < 30:   # probe batches and candidate sub-batches that cut through the batched requests
            reqs[-1].update(batch_q=9, max_pairs=50)
    reqs[5]["window"] = (x0 - 1, y0 - 1, x0 + W + 1, y0 + H + 1)
    reqs[6]["window"] = (x0 + W + 1, y0 + H + 1, x0 + W + 2, y0 + H + 2)
    s = gsj.Scheduler(ctx, lanes=1, batch=16)
    try:
        t0 = s.submit(gsj.REQ_JOIN, big, big, want_relate=True, want_area=True, upper=True)
        tickets = [s.submit(gsj.REQ_DWITHIN, L, L, **kw) for kw in reqs]
        s.wait(t0)
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    batched = [g.stats["times"]["batch"] for g in got]
    assert max(batched) > 1 and sum(b > 1 for b in batched) >= len(got) // 2
    for kw, g in zip(reqs, got):
        kw = dict(kw)
        d = ctx.dwithin(L, L, kw.pop("d"), **kw)
        assert np.array_equal(g.pairs, d.pairs)
        if kw["report_distance"] and len(d.pairs):   # an empty result may carry no column
            assert np.array_equal(g.dist, d.dist)
        assert g.stats["candidates"] == d.stats["candidates"] and g.stats["hits"] == d.stats["hits"]
    assert len(got[5].pairs) > 0 and len(got[6].pairs) == 0


def test_batched_windowed_erosion_equal_their_own_calls(ctx):
    """Windowed eroded-intersects requests batch across different radii (the prefilter, the pass-2 filter and the
    erosion read each pair's own request's radii); rows, order, candidates and hits equal the direct calls."""
    M = gsj.generate("parcels_mesh", nx=50, ny=50, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=14)
    L = ctx.polys(M)
    big = ctx.polys(gsj.subdivide(_mesh(150, 8), 10))
    x0, y0 = M.xy[:, 0].min(), M.xy[:, 1].min()
    W, H = M.xy[:, 0].max() - x0, M.xy[:, 1].max() - y0
    rng = np.random.default_rng(15)
    reqs = []
    for i in range(36):
        f = rng.uniform(0, 1, 2)
        s = rng.uniform(0.03, 0.4)
        w = (x0 + f[0] * W, y0 + f[1] * H, x0 + f[0] * W + s * W, y0 + f[1] * H + s * H)
        ra, rb = (float(v) for v in rng.uniform(0.5, 3.0, 2))
        if i % 9 == 0:
            ra = rb = 0.0        # plain intersects
        elif i % 9 == 1:
            rb = 0.0             # one side eroded
        reqs.append(dict(ra=ra, rb=rb, same=True, upper=True, window=w))
        if 24 <= i < 34:   # probe batches and candidate sub-batches that cut through the batched requests
            reqs[-1].update(batch_q=7, max_pairs=40)
    reqs[4]["window"] = (x0 - 1, y0 - 1, x0 + W + 1, y0 + H + 1)
    reqs[7]["window"] = (x0 + W + 1, y0 + H + 1, x0 + W + 2, y0 + H + 2)
    s = gsj.Scheduler(ctx, lanes=1, batch=16)
    try:
        t0 = s.submit(gsj.REQ_JOIN, big, big, want_relate=True, want_area=True, upper=True)
        tickets = [s.submit(gsj.REQ_INTERSECTS_BUFFERED, L, L, **kw) for kw in reqs]
        s.wait(t0)
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    batched = [g.stats["times"]["batch"] for g in got]
    assert max(batched) > 1 and sum(b > 1 for b in batched) >= len(got) // 2
    for kw, g in zip(reqs, got):
        kw = dict(kw)
        d = ctx.intersects_buffered(L, L, kw.pop("ra"), kw.pop("rb"), **kw)
        assert np.array_equal(g.pairs, d.pairs)
        assert g.stats["candidates"] == d.stats["candidates"] and g.stats["hits"] == d.stats["hits"]
    assert len(got[4].pairs) > 0 and len(got[7].pairs) == 0


def test_listed_self_pairs_equal_the_flags_and_batch(ctx):
    """Eroded self-intersects with the self-pair rule as a list of parents (need_list) give the rows of the same rule
    as one flag per parent (need_self), alone and, with an empty list (no parent's pair with itself), in batches of
    windowed requests; a request naming both forms is refused."""
    M = gsj.generate("parcels_mesh", nx=40, ny=40, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=21)
    L = ctx.polys(M)
    B = ctx.polys(gsj.subdivide(_mesh(150, 8), 10))
    flags = (np.arange(M.n) % 5 == 0).astype(np.uint8)
    listed = np.flatnonzero(flags).astype(np.uint32)
    x0, y0 = M.xy[:, 0].min(), M.xy[:, 1].min()
    W, H = M.xy[:, 0].max() - x0, M.xy[:, 1].max() - y0
    rng = np.random.default_rng(22)
    wins = []
    for _ in range(12):
        f, sz = rng.uniform(0, 0.7, 2), rng.uniform(0.1, 0.3)
        wins.append((x0 + f[0] * W, y0 + f[1] * H, x0 + (f[0] + sz) * W, y0 + (f[1] + sz) * H))
    s = gsj.Scheduler(ctx, lanes=1, batch=16)
    try:
        with pytest.raises(gsj.GsjError):
            s.wait(s.submit(gsj.REQ_INTERSECTS_BUFFERED, L, L, ra=1.0, rb=1.0, same=True, upper=True, need_self=flags, need_list=listed))
        whole = s.wait(s.submit(gsj.REQ_INTERSECTS_BUFFERED, L, L, ra=1.0, rb=1.5, same=True, upper=True, need_list=listed))
        big = s.submit(gsj.REQ_JOIN, B, B, want_relate=True, want_area=True, upper=True)   # keeps the lane busy: the rest queue up
        tickets = [s.submit(gsj.REQ_INTERSECTS_BUFFERED, L, L, ra=1.0, rb=1.0, same=True, upper=True, window=w, need_list=[]) for w in wins]
        s.wait(big)
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    exp = ctx.intersects_buffered(L, L, 1.0, 1.5, same=True, need_self=flags, upper=True)
    assert np.array_equal(whole.pairs, exp.pairs) and len(whole.pairs) > 0
    selfs = [a for a, b in whole.pairs.tolist() if a == b]
    assert selfs and all(flags[a] for a in selfs)
    assert max(g.stats["times"]["batch"] for g in got) > 1
    none = np.zeros(M.n, np.uint8)
    for w, g in zip(wins, got):
        d = ctx.intersects_buffered(L, L, 1.0, 1.0, same=True, need_self=none, upper=True, window=w)
        assert np.array_equal(g.pairs, d.pairs)
        assert not any(a == b for a, b in g.pairs.tolist())


def test_windowed_mix_over_lanes_with_batching(ctx):
    """Windowed joins, relates, distance joins, eroded intersects and clusters of one resident layer, submitted at once
    to three lanes that batch: every request's rows (in order) and counts equal its own call, whatever lane and batch
    it ran in."""
    M = gsj.generate("parcels_mesh", nx=70, ny=70, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=16)
    L = ctx.polys(M)
    x0, y0 = M.xy[:, 0].min(), M.xy[:, 1].min()
    W, H = M.xy[:, 0].max() - x0, M.xy[:, 1].max() - y0
    rng = np.random.default_rng(17)
    reqs = []
    for i in range(60):
        f = rng.uniform(0, 1, 2)
        s = rng.uniform(0.02, 0.3)
        w = (x0 + f[0] * W, y0 + f[1] * H, x0 + f[0] * W + s * W, y0 + f[1] * H + s * H)
        k = i % 5
        if k == 0:
            reqs.append((gsj.REQ_JOIN, dict(want_area=True, upper=True, window=w)))
        elif k == 1:
            reqs.append((gsj.REQ_JOIN, dict(want_relate=True, upper=True, window=w)))
        elif k == 2:
            reqs.append((gsj.REQ_DWITHIN, dict(d=float(rng.uniform(1, 30)), same=True, upper=True, report_distance=True, window=w)))
        elif k == 3:
            reqs.append((gsj.REQ_INTERSECTS_BUFFERED, dict(ra=float(rng.uniform(0.5, 3)), rb=float(rng.uniform(0.5, 3)), same=True, upper=True,
                                                           window=w)))
        else:
            reqs.append((gsj.REQ_CLUSTER_ROWS, dict(window=w)))
    s = gsj.Scheduler(ctx, lanes=3, batch=8)
    try:
        tickets = [s.submit(k, L, L, **kw) for k, kw in reqs]
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    assert sum(g.stats["times"]["batch"] > 1 for g in got) > 0
    for (k, kw), g in zip(reqs, got):
        kw = dict(kw)
        if k == gsj.REQ_JOIN:
            d = ctx.join(L, L, **kw)
            _rows_in_order(g, d)
        elif k == gsj.REQ_DWITHIN:
            d = ctx.dwithin(L, L, kw.pop("d"), **kw)
            assert np.array_equal(g.pairs, d.pairs)
            if len(d.pairs):
                assert np.array_equal(g.dist, d.dist)
        elif k == gsj.REQ_INTERSECTS_BUFFERED:
            d = ctx.intersects_buffered(L, L, kw.pop("ra"), kw.pop("rb"), **kw)
            assert np.array_equal(g.pairs, d.pairs)
        else:
            d = ctx.cluster_rows(L, **kw)
            assert np.array_equal(g.pairs, d.pairs)
        assert g.stats["candidates"] == d.stats["candidates"] and g.stats["hits"] == d.stats["hits"]


def test_budget_admits_windowed_requests_by_their_windows(ctx):
    """Under a scratch budget a windowed request is estimated from the probes of the tiles its window meets, not from
    the whole layer: two small windows fit the budget together and run side by side; two windows covering the layer
    do not, so the second waits until the first has finished (each still runs: a request alone is always admitted)."""
    import time

    hs = _mesh(150, 4)
    L = ctx.polys(hs)
    x0, y0 = hs.xy[:, 0].min(), hs.xy[:, 1].min()
    x1, y1 = hs.xy[:, 0].max(), hs.xy[:, 1].max()
    small = [(x0, y0, x0 + 0.2 * (x1 - x0), y0 + 0.2 * (y1 - y0)), (x1 - 0.2 * (x1 - x0), y1 - 0.2 * (y1 - y0), x1, y1)]
    full = [(x0, y0, x1, y1)] * 2
    s = gsj.Scheduler(ctx, lanes=2, scratch_budget=16 << 20, batch=1)   # small windows ~5 MB each, full ~22 MB

    def spans(wins):
        t_sub, tickets = [], []
        for w in wins:
            t_sub.append(time.perf_counter())
            tickets.append(s.submit(gsj.REQ_JOIN, L, L, want_relate=True, upper=True, window=w))
        res = [s.wait(t) for t in tickets]
        assert all(len(r.pairs) > 0 for r in res)
        return [(t + r.stats["times"]["queue_ms"] / 1e3, t + (r.stats["times"]["queue_ms"] + r.stats["times"]["run_ms"]) / 1e3)
                for t, r in zip(t_sub, res)]

    try:
        # side by side: admitted together, so overlapping unless the first ends before the second is even submitted
        # (a request takes ~2 ms here); five tries
        small_runs = [spans(small) for _ in range(5)]
        assert any(b0 < a1 for (a0, a1), (b0, b1) in small_runs), small_runs
        # one after the other, every time: admission holds the second until the first has released its scratch
        for _ in range(3):
            (a0, a1), (b0, b1) = spans(full)
            assert b0 >= a1 - 2e-4, (a0, a1, b0, b1)
    finally:
        s.close()
