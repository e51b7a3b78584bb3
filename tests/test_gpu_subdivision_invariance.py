"""GPU: the engine's piece -> parent aggregation is invariant under subdivision on the same cases as the oracle's
(tests/invariance.py: parent pairs and DE-9IM codes identical, areas within the cut-point rounding bound, outside the
reference's two documented exceptions), and the join overflow kernel (pairs whose edge lists outgrow the 96-edge LDS
lists) is exercised in both of its block shapes, bit-exact against the oracle."""
import json

import numpy as np
import pytest

import gsj
import invariance as I
from oracle import oracle as O
from conftest import sorted_rows

pytestmark = pytest.mark.gpu

CASES = [(ci, name, maxv) for ci, (name, _, _, mvs) in enumerate(I.cases()) for maxv in mvs]


@pytest.fixture(scope="module")
def cases():
    return I.cases()


@pytest.mark.parametrize("ci,name,maxv", CASES, ids=[f"{n.split()[0]}-{n.split()[1]}-maxv{m}" for _, n, m in CASES])
def test_engine_is_invariant_under_subdivision(ctx, cases, ci, name, maxv):
    _, M, R, _ = cases[ci]

    def engine_join(P, Q, upper):
        dP = ctx.polys(P)
        dQ = dP if Q is P else ctx.polys(Q)
        r = ctx.join(dP, dQ, want_area=True, want_relate=True, upper=upper)
        return I.rows({"pairs": r.pairs, "area": r.area, "code": r.code})

    rep = I.report(engine_join, M, maxv, R=R)
    assert I.failures(rep["one_sided"]) == 0, rep
    if "self" in rep:
        assert I.failures(rep["self"]) == 0, rep


@pytest.mark.parametrize("n,few,maxv", [(8, True, 0), (40, False, 0), (8, True, 400), (40, False, 400)])
def test_join_overflow_block_shapes(ctx, n, few, maxv):
    """Parcels of median 150 vertices: their self pairs (and some neighbours) collect more than 96 edges per side and
    go to the overflow kernel, eight waves per pair when a launch holds at most one pair per CU (n = 8), one wave
    per pair otherwise (n = 40). Unsubdivided, the largest lists exceed 448 edges and the launch keeps the lists in
    global memory; subdivided at 400 vertices they fit and sit in LDS."""
    M = gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=9, median=150.0)
    if maxv:
        M = gsj.subdivide(M, maxv)
    L = ctx.polys(M)
    cu = json.loads(ctx.status())["cu"]
    got = ctx.join(L, L, want_area=True, want_relate=True, upper=True)
    assert got.stats["n_over"] > 0
    assert (got.stats["n_over"] <= cu) == few, got.stats["n_over"]
    OM = O.Layer.polys(M)
    exp = O.join(OM, OM, want_area=True, want_relate=True, upper=True)
    assert got.stats["candidates"] == exp["candidates"]
    g = sorted_rows(got.pairs, got.area, got.code)
    e = sorted_rows(exp["pairs"], exp["area"], exp["code"])
    assert np.array_equal(g[0], e[0]) and np.array_equal(g[2], e[2])
    assert np.array_equal(g[1], e[1]), "areas not bit-identical"
