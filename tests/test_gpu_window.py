"""Query windows (an extension of the join options for windowed serving requests; the reference has none): a windowed
request probes only the probe-side parents with a piece box meeting the window, so its rows must be exactly the
unwindowed rows (pinned against the oracle elsewhere, and here for one case) whose probe parent meets it.

Bar: row sets, areas, DE-9IM codes and distances identical to the filtered unwindowed result."""
import numpy as np
import pytest

import gsj
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

CELL = 1 / 1549


def _mesh(n=60, seed=1, **kw):
    return gsj.generate("parcels_mesh", nx=n, ny=n, w=CELL, h=CELL, seed=seed, **kw)


def _members(layer, hs, w):
    """probe ids meeting window w: parents of a polygon layer (any piece box meets it), points of a point layer"""
    bb = layer.export()["bb"].astype(np.float64)
    m = (bb[:, 0] <= w[2]) & (bb[:, 2] >= w[0]) & (bb[:, 1] <= w[3]) & (bb[:, 3] >= w[1])
    if hs is None or hs.parent is None:
        return m
    par = np.zeros(int(hs.parent.max()) + 1, bool)
    np.logical_or.at(par, hs.parent, m)
    return par


def _window(hs, fx0, fy0, fx1, fy1):
    x0, y0 = hs.xy[:, 0].min(), hs.xy[:, 1].min()
    x1, y1 = hs.xy[:, 0].max(), hs.xy[:, 1].max()
    return (x0 + fx0 * (x1 - x0), y0 + fy0 * (y1 - y0), x0 + fx1 * (x1 - x0), y0 + fy1 * (y1 - y0))


def _same(got, full, keep, cols=()):
    sel = keep[full.pairs[:, 0]]
    e = sorted_rows(full.pairs[sel], *[getattr(full, c)[sel] for c in cols])
    g = sorted_rows(got.pairs, *[getattr(got, c) for c in cols])
    assert g[0].shape == e[0].shape, f"row count {len(g[0])} != filtered {len(e[0])}"
    assert np.array_equal(g[0], e[0]), "pair sets differ"
    for k in range(len(cols)):
        assert np.array_equal(g[1 + k], e[1 + k]), f"{cols[k]} differs"
    return len(g[0])


WINDOWS = [(0.3, 0.3, 0.6, 0.55), (0.0, 0.0, 0.05, 1.0), (0.49, 0.49, 0.49, 0.49)]


@pytest.mark.parametrize("w", WINDOWS)
def test_self_relate_and_area_in_window(ctx, w):
    S = gsj.subdivide(_mesh(60, seed=5, median=14.0), 12)   # parents with many pieces straddle the window edges
    L = ctx.polys(S)
    win = _window(S, *w)
    keep = _members(L, S, win)
    assert 0 < keep.sum() < len(keep)
    for kw, cols in ((dict(want_relate=True), ("code",)), (dict(want_area=True), ("area",))):
        full = ctx.join(L, L, upper=True, **kw)
        got = ctx.join(L, L, upper=True, window=win, **kw)
        assert _same(got, full, keep, cols) > 0 or keep.sum() < 3
        # small probe batches and sub-batches walk the member list the same way
        _same(ctx.join(L, L, upper=True, window=win, batch_q=7, max_pairs=60, **kw), full, keep, cols)


def test_two_layer_join_in_window_matches_oracle(ctx):
    P = gsj.generate("parcels_grid", nx=60, ny=60, seed=0)
    R = gsj.generate("regions_grid", nx=60, ny=60, kx=3, ky=3, edge_segs=200, seed=0)
    LP, LR = ctx.polys(P), ctx.polys(R)
    win = _window(P, 0.2, 0.1, 0.7, 0.4)
    keep = _members(LP, P, win)
    got = ctx.join(LP, LR, want_area=True, window=win)
    exp = O.join(O.Layer.polys(P), O.Layer.polys(R), want_area=True)
    sel = keep[exp["pairs"][:, 0]]
    g = sorted_rows(got.pairs, got.area)
    e = sorted_rows(exp["pairs"][sel], exp["area"][sel])
    assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1])
    assert len(g[0]) > 0


def test_points_and_distance_joins_in_window(ctx):
    M = _mesh(50, seed=3)
    L = ctx.polys(M)
    xy = gsj.generate_points(n=20000, x0=M.xy[:, 0].min(), y0=M.xy[:, 1].min(), w=50 * CELL, h=50 * CELL, clusters=10, sigma_m=40.0, seed=2)
    pts = ctx.points(xy)
    win = _window(M, 0.25, 0.25, 0.75, 0.6)
    kp = _members(pts, None, win)
    kl = _members(L, M, win)
    _same(ctx.points_in_polygons(pts, L, window=win), ctx.points_in_polygons(pts, L), kp, ("where",))
    _same(ctx.points_dwithin(pts, L, 8.0, report_distance=True, window=win), ctx.points_dwithin(pts, L, 8.0, report_distance=True), kp, ("dist",))
    _same(ctx.points_intersects_buffered(pts, L, 1.5, window=win), ctx.points_intersects_buffered(pts, L, 1.5), kp)
    _same(ctx.dwithin(L, L, 5.0, same=True, report_distance=True, upper=True, window=win),
          ctx.dwithin(L, L, 5.0, same=True, report_distance=True, upper=True), kl, ("dist",))
    _same(ctx.intersects_buffered(L, L, 1.0, 1.0, same=True, upper=True, window=win), ctx.intersects_buffered(L, L, 1.0, 1.0, same=True, upper=True), kl)


def test_window_edges_cases(ctx):
    M = _mesh(30, seed=4)
    L = ctx.polys(M)
    full = ctx.join(L, L, want_relate=True, upper=True)
    x0, y0, x1, y1 = _window(M, 0, 0, 1, 1)
    everything = ctx.join(L, L, want_relate=True, upper=True, window=(x0 - 1, y0 - 1, x1 + 1, y1 + 1))
    _same(everything, full, np.ones(M.n, bool), ("code",))
    nothing = ctx.join(L, L, want_relate=True, upper=True, window=(x1 + 1, y1 + 1, x1 + 2, y1 + 2))
    assert len(nothing.pairs) == 0 and nothing.stats["candidates"] == 0
    with pytest.raises(gsj.GsjError, match="window"):
        ctx.join(L, L, upper=True, window=(x1, y0, x0, y1))
    with pytest.raises(gsj.GsjError, match="window"):
        ctx.cluster(L, window=(x0, y0, x1, y1))


def test_window_through_the_scheduler(ctx):
    M = _mesh(40, seed=6)
    L = ctx.polys(M)
    win = _window(M, 0.1, 0.5, 0.9, 0.9)
    s = gsj.Scheduler(ctx, lanes=2)
    try:
        t = s.submit(gsj.REQ_JOIN, L, L, want_area=True, upper=True, window=win)
        got = s.wait(t)
    finally:
        s.close()
    _same(got, ctx.join(L, L, want_area=True, upper=True), _members(L, M, win), ("area",))


def _components(n, edges):
    """smallest member id of each vertex's connected component (breadth-first search)"""
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    root = np.full(n, -1, np.int64)
    for s in range(n):
        if root[s] >= 0:
            continue
        comp, stack = [s], [s]
        root[s] = s
        while stack:
            v = stack.pop()
            for w in adj[v]:
                if root[w] < 0:
                    root[w] = s
                    comp.append(w)
                    stack.append(w)
        root[comp] = min(comp)
    return root


@pytest.mark.parametrize("w", WINDOWS)
def test_windowed_cluster_rows_are_the_sub_layer_components(ctx, w):
    """cluster_intersecting of the parcels meeting a window: components over the intersecting pairs of two such
    parcels (the intersects rows of the full self-join), root = smallest parcel id; rows for those parcels only, in id
    order. Without a window, the rows are gsj_cluster's roots."""
    M = _mesh(60, seed=5, median=14.0)
    L = ctx.polys(M)
    win = _window(M, *w)
    keep = _members(L, None, win)
    full = ctx.join(L, L, upper=True).pairs
    inside = full[keep[full[:, 0]] & keep[full[:, 1]]]
    root = _components(M.n, inside)
    ids = np.flatnonzero(keep)
    got = ctx.cluster_rows(L, window=win)
    assert np.array_equal(got.pairs[:, 0], ids)
    assert np.array_equal(got.pairs[:, 1], root[ids])
    roots, _ = ctx.cluster(L)
    allr = ctx.cluster_rows(L)
    assert np.array_equal(allr.pairs[:, 0], np.arange(M.n)) and np.array_equal(allr.pairs[:, 1], roots)
    s = gsj.Scheduler(ctx, lanes=2)
    try:
        sv = s.wait(s.submit(gsj.REQ_CLUSTER_ROWS, L, window=win))
    finally:
        s.close()
    assert np.array_equal(sv.pairs, got.pairs)


def test_windowed_cluster_needs_unsubdivided_layers(ctx):
    S = ctx.polys(gsj.subdivide(_mesh(20, seed=2, median=14.0), 12))
    with pytest.raises(gsj.GsjError, match="windowed cluster"):
        ctx.cluster_rows(S, window=(-1e3, -1e3, 1e3, 1e3))
