"""Subdivision invariance of the piece -> parent aggregation (shared by the CPU and GPU tests and by the planted-mistake
check, which runs it in a child process against another build of the oracle).

Subdividing a layer changes how a parent pair's evidence is computed (per piece pair, then summed, engine/engine.cuh:
963-1042) but not what it measures: a parent pair intersects, and has the same DE-9IM matrix and intersection area,
with or without subdivision. The reference's rules meet this exactly except in two cases that follow from them:

* dropped slivers: a clipped ring whose shoelace sum in world coordinates rounds to exactly 0 is dropped
  (engine/subdivide.h:51-53), so deep cuts of a spiky parcel can lose pieces of ~1e-13 square degrees. Such parents
  are found independently here: their pieces' areas do not add up to the parent's.
* doubly cut shared edges: where two parents share an edge and both are cut across it at different points, each
  piece edge ends at a rounded cut point that is not an endpoint of the other piece's edge, so neither edge has both
  endpoints on the other's line (engine/engine.cuh:437-442) and the two collinear edges are not seen as coincident; a
  sliver of interior overlap can follow. Such parent pairs are found independently here: both parents were cut and
  they share an edge exactly.

Everywhere else: the same parent pairs, the same DE-9IM codes, and areas equal within AREA_BOUND times the smaller
parent's area. The bound covers the rounding of cut points: a cut point lies within an ulp or so of the edge it cuts,
so every cut moves the boundary by about 1e-14 degrees along a piece edge of at most 1e-3 degrees; the largest
difference measured on the config-2 generator's layers is 1.1e-11 of the smaller parent (at 8 vertices per piece).
"""
import numpy as np

import gsj

AREA_BOUND = 1e-10
LOSS_REL = 1e-10


def mesh(n, seed=1, **kw):
    return gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=seed, **kw)


def shoelace_sums(hs, owner, origin):
    """Area of every geometry of hs summed into owner[geometry], each ring's shoelace taken relative to
    origin[owner] (the parent's first vertex, so that pieces and parents are measured in the same frame)."""
    nr = len(hs.rvo) - 1
    ring_geom = np.repeat(np.arange(hs.n), np.diff(hs.gro.astype(np.int64)))
    vert_ring = np.repeat(np.arange(nr), np.diff(hs.rvo.astype(np.int64)))
    xy = hs.xy - origin[owner[ring_geom[vert_ring]]]
    x, y = xy[:, 0], xy[:, 1]
    term = x[:-1] * y[1:] - x[1:] * y[:-1]
    same = vert_ring[:-1] == vert_ring[1:]          # consecutive vertices of one ring form an edge
    out = np.zeros(int(owner.max()) + 1 if len(owner) else 0)
    np.add.at(out, owner[ring_geom[vert_ring[:-1][same]]], 0.5 * term[same])
    return out


def parent_areas(M):
    origin = M.xy[M.rvo[M.gro[:-1]].astype(np.int64)]
    return shoelace_sums(M, np.arange(M.n), origin), origin


def lossy_parents(M, S, A=None, origin=None):
    """parents whose pieces' areas do not add up to the parent's (a piece was dropped). Rounded cut points alone
    leave sums within 2e-11 of the parent's area on the config-2 generator's layers; a dropped sliver takes away more
    than 8e-10 there, so LOSS_REL separates the two."""
    if A is None:
        A, origin = parent_areas(M)
    if S.parent is None:
        return np.zeros(M.n, bool)
    ps = np.zeros(M.n)
    ps[: len(shoelace_sums(S, S.parent.astype(np.int64), origin))] = shoelace_sums(S, S.parent.astype(np.int64), origin)
    return np.abs(ps - A) > LOSS_REL * np.abs(A)


def shared_edge_pairs(M):
    """parent pairs (a < b) holding the same edge with opposite directions (bit-identical endpoints)"""
    nr = len(M.rvo) - 1
    ring_geom = np.repeat(np.arange(M.n), np.diff(M.gro.astype(np.int64)))
    vert_ring = np.repeat(np.arange(nr), np.diff(M.rvo.astype(np.int64)))
    ok = vert_ring[:-1] == vert_ring[1:]
    g = ring_geom[vert_ring[:-1][ok]]
    b = np.ascontiguousarray(M.xy).view(np.uint64).reshape(-1, 2)
    fwd = np.concatenate([b[:-1][ok], b[1:][ok]], axis=1)
    rev = np.concatenate([b[1:][ok], b[:-1][ok]], axis=1)
    index = {}
    for k, row in enumerate(map(bytes, fwd)):
        index.setdefault(row, []).append(int(g[k]))
    out = set()
    for k, row in enumerate(map(bytes, rev)):
        for h in index.get(row, ()):
            if h != g[k]:
                out.add((min(h, int(g[k])), max(h, int(g[k]))))
    return out


def cut_parents(M, S):
    if S.parent is None:
        return np.zeros(M.n, bool)
    return np.bincount(S.parent.astype(np.int64), minlength=M.n) > 1


def rows(r, area=True, code=True):
    p = np.asarray(r["pairs"]).astype(np.int64).reshape(-1, 2)
    a = r["area"] if area else np.zeros(len(p))
    c = r["code"] if code else np.zeros(len(p), np.uint32)
    return {(int(x), int(y)): (float(a[i]), int(c[i])) for i, (x, y) in enumerate(p.tolist())}


def compare(base, sub, exempt, scale):
    """Differences between the unsubdivided rows `base` and the subdivided rows `sub` outside `exempt(a, b)`:
    missing / extra parent pairs, DE-9IM code differences, areas off by more than AREA_BOUND x scale(a, b)."""
    out = {"rows": len(base), "missing": 0, "extra": 0, "code": 0, "area": 0, "exempt": 0, "max_area_rel": 0.0}
    for k in base.keys() | sub.keys():
        if exempt(*k):
            out["exempt"] += 1
            continue
        if k not in sub:
            out["missing"] += 1
        elif k not in base:
            out["extra"] += 1
        else:
            (a0, c0), (a1, c1) = base[k], sub[k]
            out["code"] += int(c0 != c1)
            d = abs(a0 - a1) / scale(*k)
            out["max_area_rel"] = max(out["max_area_rel"], d)
            out["area"] += int(d > AREA_BOUND)
    return out


def failures(rep):
    return rep["missing"] + rep["extra"] + rep["code"] + rep["area"]


def report(join, M, maxv, self_join=True, R=None):
    """Invariance of `join` (a callable (P, R, upper) -> rows dict, over host sets) on M subdivided at maxv.

    One-sided: sub(M) x R vs M x R, R = M unless given (every shared edge is cut on one side at most).
    Self (R None and self_join): sub(M) x sub(M) vs M x M, each unordered pair once."""
    S = gsj.subdivide(M, maxv)
    A, origin = parent_areas(M)
    lossy = lossy_parents(M, S, A, origin)
    R = M if R is None else R
    AR = A if R is M else parent_areas(R)[0]
    scale = lambda a, b: min(abs(A[a]), abs(AR[b]))
    out = {"maxv": maxv, "pieces": int(S.n), "lossy_parents": int(lossy.sum())}
    out["one_sided"] = compare(join(M, R, False), join(S, R, False), lambda a, b: bool(lossy[a]), scale)
    if self_join and R is M:
        cut = cut_parents(M, S)
        dbl = {k for k in shared_edge_pairs(M) if cut[k[0]] and cut[k[1]]}
        out["double_cut_pairs"] = len(dbl)
        out["self"] = compare(join(M, M, True), join(S, S, True), lambda a, b: bool(lossy[a] or lossy[b]) or (a, b) in dbl, scale)
    return out


def oracle_join(P, R, upper):
    from oracle import oracle as O

    OP = O.Layer.polys(P)
    OR = OP if R is P else O.Layer.polys(R)
    return rows(O.join(OP, OR, want_area=True, want_relate=True, upper=upper))


def holed_and_crowded():
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


def square_and_rectangles(k=60):
    """A unit square with k vertices per side (so that it is cut, at dyadic midpoints), and rectangles whose sides lie on
    the 1/8-degree lines strictly inside the square or cross its boundary (never run along it, which would give an
    edge more overlaps than the reference's per-edge caps hold): many rectangle edges run exactly along the square's
    cut lines, where a real edge lies on the boundary between two pieces of the other polygon. Returns the square, the
    rectangles and the exact area of each rectangle's overlap with the square."""
    t = np.linspace(0, 1, k + 1)[:-1]
    sq = np.concatenate([np.c_[t, 0 * t], np.c_[1 + 0 * t, t], np.c_[1 - t, 1 + 0 * t], np.c_[0 * t, 1 - t]])
    sq = np.vstack([sq, sq[:1]]) + np.array([-95.0, 38.0])
    S = gsj.HostSet(np.array([0, 1]), np.array([0, len(sq)]), sq)
    lines = [-3 / 16] + [i / 8 for i in range(1, 8)] + [19 / 16]
    rects, exact = [], []
    for i in range(len(lines) - 1):
        for j in range(len(lines) - 2):
            x0, x1 = lines[i], lines[min(i + 1 + (i % 2), len(lines) - 1)]
            y0, y1 = lines[j], lines[j + 2]
            rects.append(np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1], [x0, y0]]) + np.array([-95.0, 38.0]))
            exact.append((min(x1, 1) - max(x0, 0)) * (min(y1, 1) - max(y0, 0)))
    xy = np.concatenate(rects)
    return S, gsj.HostSet(np.arange(len(rects) + 1), np.arange(len(rects) + 1) * 5, xy), np.array(exact)


def cases():
    """(name, M, R or None, maxv list): the config-2 generator (lognormal vertex counts, shared edges, overlaps),
    config-1 regions (2000 vertices each) under parcels, and holed / crowded rings under a parcel mesh"""
    P1 = gsj.generate("parcels_grid", nx=60, ny=60, seed=0)
    R1 = gsj.generate("regions_grid", nx=60, ny=60, kx=3, ky=3, edge_segs=300, seed=0)
    H = holed_and_crowded()
    HM = gsj.generate("parcels_mesh", nx=24, ny=24, x0=-95.0, y0=38.0, w=1 / 24, h=1 / 24, seed=3)
    SQ, RC, _ = square_and_rectangles()
    return [("config-2 mesh 160x160", mesh(160), None, (8, 16, 64, 256)),
            ("dyadic rectangles along a subdivided square's cut lines", SQ, RC, (8, 16, 32)),
            ("config-1 regions under parcels", R1, P1, (64, 256)),
            ("holed and crowded rings under a mesh", H, HM, (12, 32))]


def total_failures(join, which=None):
    """sum of failures over the cases (or the (case index, maxv) pairs in `which`)"""
    n = 0
    for ci, (name, M, R, mvs) in enumerate(cases()):
        for maxv in mvs:
            if which is None or (ci, maxv) in which:
                rep = report(join, M, maxv, R=R)
                n += failures(rep["one_sided"]) + (failures(rep["self"]) if "self" in rep else 0)
    return n


if __name__ == "__main__":   # child process of the planted-mistake test: GSJ_ORACLE_LIB names the build to check
    print("failures", total_failures(oracle_join, which={(0, 16), (1, 16)}), flush=True)
