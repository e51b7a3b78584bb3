"""CPU tests: the oracle's candidate enumeration, clustering, subdivision and dedup against brute force / invariants;
the engine's host-side preparation (files, subdivide, dedup, generators) against the oracle; the C ABI's exports."""
import os
import re
import struct
import tempfile
from collections import deque
from fractions import Fraction

import numpy as np
import pytest

import gsj
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def mesh(n=60, seed=1, **kw):
    return gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=seed, **kw)


# ------------------------------------------------------------------------------------------- ABI
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "gsj.h")).read()
    declared = set(re.findall(r"\b(gsj_[a-z0-9_]+)\s*\(", hdr)) - {"gsj_rows_cb", "gsj_points_cb"}
    lib = gsj.load()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(gsj.EXPORTS) | {"gsj_version"}


def test_engine_fails_loudly_without_a_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(gsj.GsjError):
        gsj.Context(0)


# ------------------------------------------------------------------------------------------- candidates
def test_candidates_equal_brute_force():
    M = mesh(40, seed=3)
    L = O.Layer.polys(M)
    bb = L.export()["bb"]
    for upper in (False, True):
        got = {tuple(p) for p in O.candidates(L, L, upper=upper).tolist()}
        x0, y0, x1, y1 = (bb[:, k][:, None] for k in range(4))
        ov = ~((bb[:, 0][None, :] > x1) | (bb[:, 2][None, :] < x0) | (bb[:, 1][None, :] > y1) | (bb[:, 3][None, :] < y0))
        q, r = np.nonzero(ov)
        if upper:
            keep = r >= q
            q, r = q[keep], r[keep]
        assert got == set(zip(q.tolist(), r.tolist()))


def test_point_candidates_equal_brute_force_at_the_extent_border():
    """Point boxes touching the far border of the build layer's extent (a regression: the grid enumeration dropped
    them when the query's first cell index rounded to one past the last cell)."""
    rng = np.random.default_rng(4)
    for n, seed in ((40, 6), (25, 1), (33, 2), (18, 3)):
        M = gsj.generate("parcels_mesh", nx=n, ny=n, x0=-95.2, y0=38.6, w=0.6 / 1095, h=0.6 / 1095, seed=seed)
        L = O.Layer.polys(M)
        bb = L.export()["bb"]
        edge = np.nonzero(bb[:, 2] == bb[:, 2].max())[0]
        xy = np.concatenate([M.xy[::3], np.stack([bb[edge, 2].astype(np.float64), 0.5 * (bb[edge, 1] + bb[edge, 3]).astype(np.float64)], 1),
                             np.stack([rng.uniform(bb[:, 0].min(), bb[:, 2].max(), 50), np.full(50, bb[:, 3].max())], 1)])
        P = O.Layer.points_layer(xy)
        pb = P.export()["bb"]
        got = {tuple(p) for p in O.candidates(P, L).tolist()}
        ov = ~((bb[None, :, 0] > pb[:, None, 2]) | (bb[None, :, 2] < pb[:, None, 0]) | (bb[None, :, 1] > pb[:, None, 3]) |
               (bb[None, :, 3] < pb[:, None, 1]))
        q, r = np.nonzero(ov)
        assert got == set(zip(q.tolist(), r.tolist())), (n, seed)


# ------------------------------------------------------------------------------------------- cluster
def test_cluster_equals_bfs_components():
    M = gsj.subdivide(mesh(50, seed=4), 10)
    L = O.Layer.polys(M)
    res = O.cluster(L)
    n = L.nparent
    adj = [[] for _ in range(n)]
    for a, b in res["pairs"].tolist():
        adj[a].append(b)
        adj[b].append(a)
    root = [-1] * n
    for s in range(n):
        if root[s] >= 0:
            continue
        comp, dq = [s], deque([s])
        root[s] = s
        while dq:
            u = dq.popleft()
            for v in adj[u]:
                if root[v] < 0:
                    root[v] = s
                    comp.append(v)
                    dq.append(v)
        m = min(comp)
        for v in comp:
            root[v] = m
    assert res["root"].tolist() == root
    assert len(set(root)) < n   # neighbours that touch do merge


# ------------------------------------------------------------------------------------------- subdivide
def _shoelace_exact(hs, g):
    s = Fraction(0)
    for r in range(int(hs.gro[g]), int(hs.gro[g + 1])):
        xy = hs.xy[int(hs.rvo[r]):int(hs.rvo[r + 1])]
        for i in range(len(xy) - 1):
            s += Fraction(xy[i, 0]) * Fraction(xy[i + 1, 1]) - Fraction(xy[i + 1, 0]) * Fraction(xy[i, 1])
    return s / 2


@pytest.mark.parametrize("maxv", [6, 16, 64])
def test_subdivide_engine_equals_oracle_and_invariants(maxv):
    """Pieces partition their parent up to the rounding of cut points. At maxv=6 some parents hit the depth-24 cap
    (engine/subdivide.h:61) and zero-area slivers are dropped (subdivide.h:45), so the area check runs at >= 16."""
    M = mesh(30, seed=5)
    S = gsj.subdivide(M, maxv)
    og, orv, oxy, opar, osyn = O.subdivide(M, maxv)
    assert np.array_equal(S.gro, og) and np.array_equal(S.rvo, orv) and np.array_equal(S.xy, oxy)
    assert np.array_equal(S.parent, opar) and np.array_equal(S.syn, osyn)
    # pieces partition each parent: areas sum to the parent's (up to the cut points' rounding)
    piece_of = {}
    for g, p in enumerate(S.parent.tolist()):
        piece_of.setdefault(p, []).append(g)
    for p in (list(piece_of)[:200] if maxv >= 16 else []):
        whole = _shoelace_exact(M, p)
        parts = sum(_shoelace_exact(S, g) for g in piece_of[p])
        assert abs(float(parts - whole)) <= 1e-10 * float(whole)
    # vertex bound (depth limit aside) and synthetic edges lie on axis-parallel cut lines
    nv = S.rvo[S.gro[1:]] - S.rvo[S.gro[:-1]]
    if maxv >= 16:
        assert (nv[np.isin(S.parent, [p for p, v in piece_of.items() if len(v) > 1])] <= maxv).all()
    v = np.nonzero(S.syn)[0]
    v = v[v + 1 < len(S.xy)]
    assert ((S.xy[v, 0] == S.xy[v + 1, 0]) | (S.xy[v, 1] == S.xy[v + 1, 1])).all()


def test_dedup_engine_equals_oracle():
    M = mesh(20, seed=6)
    # duplicate some geometries (stacked parcels) in shuffled positions
    idx = np.concatenate([np.arange(M.n), np.arange(0, M.n, 7), np.arange(3, M.n, 11)])
    rng = np.random.default_rng(0)
    rng.shuffle(idx)
    gro, rvo, xy = [0], [0], []
    for g in idx:
        for r in range(int(M.gro[g]), int(M.gro[g + 1])):
            xy.append(M.xy[int(M.rvo[r]):int(M.rvo[r + 1])])
            rvo.append(rvo[-1] + len(xy[-1]))
        gro.append(len(rvo) - 1)
    D = gsj.HostSet(np.array(gro), np.array(rvo), np.concatenate(xy))
    U, moff, midx = gsj.dedup(D)
    first, omoff, omidx = O.dedup(D)
    assert U.n == len(first) == M.n
    assert np.array_equal(moff, omoff) and np.array_equal(midx, omidx)
    # members partition the rows and are exactly equal to their representative
    assert sorted(midx.tolist()) == list(range(D.n))
    for u in range(0, U.n, 13):
        for k in range(int(moff[u]), int(moff[u + 1])):
            g = int(midx[k])
            assert np.array_equal(D.xy[int(D.rvo[D.gro[g]]):int(D.rvo[D.gro[g + 1]])], U.xy[int(U.rvo[U.gro[u]]):int(U.rvo[U.gro[u + 1]])])


# ------------------------------------------------------------------------------------------- formats
def test_gset_pset_layout_and_round_trip():
    M = mesh(10, seed=7)
    pts = gsj.generate_points(n=1000, seed=2, x0=-95.5, y0=38.5, w=0.01, h=0.01, clusters=5)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "a.gset")
        gsj.write_gset(p, M)
        raw = open(p, "rb").read()
        # layout of scripts/dump_gset.py: '<4sIQQQ' header, u64 offsets, f64 xy
        magic, ver, n, nr, nv = struct.unpack_from("<4sIQQQ", raw, 0)
        assert (magic, ver, n, nr, nv) == (b"GSET", 1, M.n, M.nr, M.nv)
        off = 32
        assert np.array_equal(np.frombuffer(raw, "<u8", n + 1, off), M.gro)
        off += 8 * (n + 1)
        assert np.array_equal(np.frombuffer(raw, "<u8", nr + 1, off), M.rvo)
        off += 8 * (nr + 1)
        assert np.array_equal(np.frombuffer(raw, "<f8", 2 * nv, off).reshape(-1, 2), M.xy)
        assert len(raw) == off + 16 * nv
        R = gsj.read_gset(p)
        assert np.array_equal(R.xy, M.xy) and len(R.ids) == M.n
        q = os.path.join(d, "a.pset")
        gsj.write_pset(q, pts)
        raw = open(q, "rb").read()
        assert struct.unpack_from("<4sIQ", raw, 0) == (b"PSET", 1, len(pts)) and len(raw) == 16 + 16 * len(pts)
        assert np.array_equal(gsj.read_pset(q), pts)
        open(os.path.join(d, "bad.gset"), "wb").write(b"XXXX" + raw[4:40])
        with pytest.raises(gsj.GsjError):
            gsj.read_gset(os.path.join(d, "bad.gset"))


def _gset_bytes(gro, rvo, xy, n=None, nr=None, nv=None):
    gro, rvo, xy = np.asarray(gro, "<u8"), np.asarray(rvo, "<u8"), np.asarray(xy, "<f8")
    hdr = struct.pack("<4sIQQQ", b"GSET", 1, len(gro) - 1 if n is None else n, len(rvo) - 1 if nr is None else nr,
                      len(xy) // 2 if nv is None else nv)
    return hdr + gro.tobytes() + rvo.tobytes() + xy.tobytes()


def test_malformed_files_are_rejected_before_use():
    """Header counts bounded by the file size (no overflow, no allocation beyond the file), offsets checked for
    monotonicity and range: a corrupt or crafted file is an argument error, never an out-of-range read."""
    sq = [[0, 0], [1, 0], [1, 1], [0, 1], [0, 0]]
    good = _gset_bytes([0, 1], [0, 5], np.ravel(sq))
    bad = {
        "truncated": good[:-8],
        "trailing bytes": good + b"\0" * 8,
        "huge n": _gset_bytes([0, 1], [0, 5], np.ravel(sq), n=2**64 - 1),
        "n past the file": _gset_bytes([0, 1], [0, 5], np.ravel(sq), n=1000),
        "huge nv": _gset_bytes([0, 1], [0, 5], np.ravel(sq), nv=2**63),
        "gro decreasing": _gset_bytes([0, 2, 1], [0, 5], np.ravel(sq + sq)[:10]),
        "rvo decreasing": _gset_bytes([0, 2], [0, 4, 3], np.ravel(sq[:3])),
        "rvo past nv": _gset_bytes([0, 1], [0, 7], np.ravel(sq)),
        "gro[0] != 0": _gset_bytes([1, 1], [0, 5], np.ravel(sq)),
        "header only": good[:20],
    }
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "ok.gset")
        open(p, "wb").write(good)
        assert gsj.read_gset(p).n == 1
        for name, raw in bad.items():
            p = os.path.join(d, "bad.gset")
            open(p, "wb").write(raw)
            with pytest.raises(gsj.GsjError) as e:
                gsj.read_gset(p)
            assert e.value.code == -1, (name, str(e.value))
        for name, raw in {"truncated": struct.pack("<4sIQ", b"PSET", 1, 3) + b"\0" * 40,
                          "huge n": struct.pack("<4sIQ", b"PSET", 1, 2**62) + b"\0" * 16}.items():
            p = os.path.join(d, "bad.pset")
            open(p, "wb").write(raw)
            with pytest.raises(gsj.GsjError) as e:
                gsj.read_pset(p)
            assert e.value.code == -1, (name, str(e.value))
        with pytest.raises(gsj.GsjError) as e:
            gsj.read_gset(os.path.join(d, "missing.gset"))
        assert e.value.code == -6


def test_prep_entry_points_check_offsets():
    """gro / rvo must start at 0 and never decrease (the arrays keep their lengths, so every read stays in them)"""
    M = mesh(6, seed=2)
    rvo_swapped = M.rvo.copy()
    rvo_swapped[[2, 3]] = rvo_swapped[[3, 2]]
    gro_dip = M.gro.copy()
    gro_dip[5] = gro_dip[4] - 1
    for gro, rvo in ((M.gro[::-1].copy(), M.rvo), (gro_dip, M.rvo), (M.gro, rvo_swapped), (M.gro + 1, M.rvo)):
        H = gsj.HostSet(gro, rvo, M.xy)
        for f in (lambda: gsj.subdivide(H, 8), lambda: gsj.dedup(H)):
            with pytest.raises(gsj.GsjError) as e:
                f()
            assert e.value.code == -1


# ------------------------------------------------------------------------------------------- generators
def test_generators_are_deterministic_and_match_their_spec():
    a, b = mesh(50, seed=9), mesh(50, seed=9)
    assert np.array_equal(a.xy, b.xy) and np.array_equal(a.rvo, b.rvo)
    assert not np.array_equal(mesh(50, seed=10).xy[:100], a.xy[:100])
    # config 1: every parcel strictly inside one region, one candidate per parcel (closed form area = shoelace)
    P = gsj.generate("parcels_grid", nx=100, ny=100, seed=0)
    R = gsj.generate("regions_grid", nx=100, ny=100, seed=0)
    k = np.diff(P.rvo) - 1
    assert k.min() >= 4 and k.max() <= 12 and R.n == 16 and (np.diff(R.rvo) == 2001).all()
    res = O.join(O.Layer.polys(P), O.Layer.polys(R), want_area=True)
    assert res["candidates"] == P.n and len(res["pairs"]) == P.n and len(np.unique(res["pairs"][:, 0])) == P.n
    # config 2: shared edges ~30%, overlaps ~5% of interior sides; vertex counts lognormal, median ~7 (+closing)
    M = mesh(200, seed=1)
    r = O.join(O.Layer.polys(M), O.Layer.polys(M), want_area=False, want_relate=True, upper=True)
    codes = [gsj.de9im_str(c) for c in r["code"]]
    sides = 2 * 200 * 199
    assert 0.27 < codes.count("FF2F11212") / sides < 0.33
    assert 0.04 < codes.count("212101212") / sides < 0.06
    assert codes.count("2FFF1FFF2") == M.n
    nv = np.diff(M.rvo) - 1
    assert 6 <= np.median(nv) <= 9
