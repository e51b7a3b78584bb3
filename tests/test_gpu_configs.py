"""GPU, through the served path: the batched-requests and serving configurations' request streams, every request's rows
checked against the oracle (not only against the engine's own direct calls).

* Batched requests: 64 requests drawn exactly as the benchmark draws them (config-1 parcels and regions, sizes
  lognormal in [1e3, 2e4] polygons, 40% intersects + area, 20% relate, 20% dwithin self-joins with d ~ U(1, 50) m, 20%
  eroded intersects with ra, rb ~ U(0.5, 3) m), all in flight at once on a four-lane scheduler.
* Serving stream: 48 windowed requests drawn as the benchmark's stream draws them (joins, relate, dwithin, eroded
  intersects, cluster_intersecting of a window's parcels), over a 200k-parcel layer of the national generator, four
  lanes with batching on. A windowed request's rows are the oracle's unwindowed rows whose probe parcel's float box
  meets the window; cluster rows are the connected components (smallest id as root) of the window's parcels over the
  oracle's intersecting pairs of two such parcels.

Bar: pair sets, DE-9IM codes and areas bit-identical to the oracle; candidate counts equal."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import bench
import gsj
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu


def _check(got, exp, cols=()):
    g = sorted_rows(got.pairs, *[getattr(got, c) for c in cols])
    e = sorted_rows(exp["pairs"], *[exp[c] for c in cols])
    assert g[0].shape == e[0].shape, f"row count {len(g[0])} != oracle {len(e[0])}"
    assert np.array_equal(g[0], e[0]), "pair sets differ"
    for k, c in enumerate(cols):
        assert np.array_equal(g[1 + k], e[1 + k]), f"{c} differs"


def test_batched_requests_match_the_oracle(ctx):
    mix = bench.config4_mix(64, 2e4)
    assert {q["kind"] for q in mix} == {"area", "relate", "dwithin", "erode"}
    host, subs = [], []
    for q in mix:
        P, R, LP, LR = bench.config4_layers(gsj, ctx, q)
        host.append((P, R))
        subs.append(bench.config4_submit_args(gsj, q, LP, LR))
    s = gsj.Scheduler(ctx, lanes=4)
    try:
        tickets = [s.submit(k, A, B, **kw) for k, A, B, kw in subs]
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    for q, (P, R), g in zip(mix, host, got):
        OP = O.Layer.polys(P)
        OR = O.Layer.polys(R) if R is not None else OP
        if q["kind"] == "area":
            exp = O.join(OP, OR, want_area=True)
            _check(g, exp, ("area",))
        elif q["kind"] == "relate":
            exp = O.join(OP, OR, want_area=False, want_relate=True)
            _check(g, exp, ("code",))
        elif q["kind"] == "dwithin":
            exp = O.dwithin(OP, OP, q["d"], same=True, upper=True)
            _check(g, exp)
        else:
            exp = O.erode(OP, OR, q["ra"], q["rb"])
            _check(g, exp)
        assert g.stats["candidates"] == exp["candidates"], q


NATIONAL_CELL = 60.0 / 19737


def test_serving_stream_matches_the_oracle(ctx):
    n = 447   # 200k parcels of the national generator
    x0, y0 = -125.0, 24.0
    M = gsj.generate("parcels_mesh", nx=n, ny=n, x0=x0, y0=y0, w=NATIONAL_CELL, h=NATIONAL_CELL, median=10.0, kmax=256, seed=4)
    L = ctx.polys(M)
    OM = O.Layer.polys(M)
    bb = OM.export()["bb"].astype(np.float64)
    stream = bench.window_stream(48, x0, y0, n * NATIONAL_CELL, n * NATIONAL_CELL)
    assert {name for name, _ in stream} == {"area", "relate", "dwithin", "erode", "cluster"}
    s = gsj.Scheduler(ctx, lanes=4, batch=16)
    try:
        tickets = [s.submit(bench.window_submit_kind(gsj, name), L, L, **kw) for name, kw in stream]
        got = [s.wait(t) for t in tickets]
    finally:
        s.close()
    assert any(g.stats["times"]["batch"] > 1 for g in got)   # requests were batched
    full = {"area": O.join(OM, OM, want_area=True, upper=True), "relate": O.join(OM, OM, want_area=False, want_relate=True, upper=True)}
    for (name, kw), g in zip(stream, got):
        w = kw["window"]
        keep = (bb[:, 0] <= w[2]) & (bb[:, 2] >= w[0]) & (bb[:, 1] <= w[3]) & (bb[:, 3] >= w[1])
        if name == "cluster":
            pr = full["area"]["pairs"]
            inside = pr[keep[pr[:, 0]] & keep[pr[:, 1]]]
            graph = coo_matrix((np.ones(len(inside)), (inside[:, 0], inside[:, 1])), shape=(M.n, M.n))
            _, lab = connected_components(graph, directed=False)
            root = np.full(lab.max() + 1, M.n)
            np.minimum.at(root, lab[keep], np.flatnonzero(keep))
            ids = np.flatnonzero(keep)
            assert np.array_equal(g.pairs[:, 0], ids) and np.array_equal(g.pairs[:, 1], root[lab[ids]]), kw
            continue
        if name in full:
            exp = full[name]
        elif name == "dwithin":
            exp = O.dwithin(OM, OM, kw["d"], same=True, upper=True)
        else:
            exp = O.erode(OM, OM, kw["ra"], kw["rb"], same=True, upper=True)
        sel = keep[exp["pairs"][:, 0]]
        cols = {"area": ("area",), "relate": ("code",)}.get(name, ())
        _check(g, {"pairs": exp["pairs"][sel], **{c: exp[c][sel] for c in cols}}, cols)
