"""ctypes wrapper of the oracle (test infrastructure only): the CPU restatement of the reference engine that the HIP
engine is checked against. Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GSJ_ORACLE_LIB names another build of the same sources (the planted-mistake builds of the invariance tests)
LIB_PATH = os.environ.get("GSJ_ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")
FLOPS_LIB_PATH = os.path.join(_HERE, "build", "liboracle_flops.so")
_lib = None
_flib = None


def build():
    subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)


def load(flops=False):
    global _lib, _flib
    if (_flib if flops else _lib) is not None:
        return _flib if flops else _lib
    path = FLOPS_LIB_PATH if flops else LIB_PATH
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    d, u64, u32, i32, vp = C.c_double, C.c_uint64, C.c_uint32, C.c_int, C.c_void_p
    dp, u64p, u32p, u8p, fp = C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.POINTER(C.c_float)
    sig = {
        "orc_orient": (d, [d] * 6), "orc_left_of": (i32, [d] * 6), "orc_seg_isect": (i32, [d] * 8),
        "orc_f_down": (C.c_float, [d]), "orc_f_up": (C.c_float, [d]), "orc_de9im_code": (u32, [dp, u32, d, d]),
        "orc_set_threads": (None, [i32]), "orc_flops": (None, [u64p, i32]),
        "orc_layer_polys": (vp, [u64p, u64p, dp, u64, u64, u64, u32p, u8p]), "orc_layer_points": (vp, [dp, u64]),
        "orc_layer_free": (None, [vp]), "orc_layer_export": (None, [vp, dp, dp, fp, dp, dp]), "orc_layer_nparent": (u32, [vp]),
        "orc_result_new": (vp, []), "orc_result_free": (None, [vp]), "orc_result_rows": (u64, [vp]), "orc_result_candidates": (u64, [vp]),
        "orc_result_copy": (None, [vp, u32p, dp, u32p, u8p]), "orc_result_nroot": (u64, [vp]), "orc_result_roots": (None, [vp, u32p]),
        "orc_candidates": (u64, [vp, vp, i32, u32p, u64]),
        "orc_refine": (None, [vp, vp, u32p, u64, i32, i32, i32, u8p, dp, dp, u32p]),
        "orc_join": (None, [vp, vp, i32, i32, i32, i32, vp]),
        "orc_pip_pair": (C.c_uint8, [vp, vp, u32, u32]), "orc_points_join": (None, [vp, vp, vp]), "orc_cluster": (None, [vp, vp]),
        "orc_subdivide": (vp, [u64p, u64p, dp, u64, i32]), "orc_flat_sizes": (None, [vp, u64p, u64p, u64p]),
        "orc_flat_copy": (None, [vp, u64p, u64p, dp, u8p, u32p]), "orc_flat_free": (None, [vp]),
        "orc_dedup": (vp, [u64p, u64p, dp, u64]), "orc_dedup_n": (u64, [vp]), "orc_dedup_copy": (None, [vp, u32p, u64p, u32p]),
        "orc_dedup_free": (None, [vp]),
        # distance joins and their geodesy (buffer.cpp)
        "orc_gm_sin": (d, [d]), "orc_gm_cos": (d, [d]), "orc_gm_tan": (d, [d]), "orc_gm_atan": (d, [d]), "orc_gm_atan2": (d, [d, d]),
        "orc_geo_m_per_deg": (None, [d, dp, dp]), "orc_geo_deg_cover": (None, [d, d, d, i32, dp, dp]),
        "orc_geo_utm_k": (d, [d] * 4), "orc_geo_vincenty": (d, [d] * 4), "orc_geo_geog_dist": (d, [d] * 8 + [i32]),
        "orc_geo_frame": (None, [d, d, dp]), "orc_geo_band": (d, [d] * 5), "orc_geo_seg_seg_d2": (d, [d] * 8),
        "orc_dw_pair": (i32, [vp, vp, u32, u32, d, i32, dp]),
        "orc_dw_new": (vp, []), "orc_dw_free": (None, [vp]), "orc_dw_rows": (u64, [vp]), "orc_dw_candidates": (u64, [vp]),
        "orc_dw_copy": (None, [vp, u32p, dp]), "orc_dwithin": (None, [vp, vp, d, i32, i32, i32, vp]),
        "orc_points_dwithin": (None, [vp, vp, d, i32, vp]),
        "orc_er_pair": (i32, [vp, vp, u32, u32, d, d]), "orc_erode": (None, [vp, vp, d, d, i32, i32, u8p, vp]),
        "orc_points_erode": (None, [vp, vp, d, vp]),
        # binary COPY / EWKB parse (ewkb.cpp)
        "orc_copy_parse": (vp, [u8p, u64, i32, i32]), "orc_copy_error": (C.c_char_p, [vp]), "orc_copy_sizes": (None, [vp, u64p]),
        "orc_copy_arrays": (None, [vp] + [C.c_void_p] * 7), "orc_copy_free": (None, [vp]), "orc_ring_twice_area": (d, [dp, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    if flops:
        lib.orc_set_threads(1)   # counters are per thread
        _flib = lib
    else:
        _lib = lib
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def set_threads(n: int):
    load().orc_set_threads(int(n))


def orient(ax, ay, bx, by, cx, cy):
    return load().orc_orient(ax, ay, bx, by, cx, cy)


def left_of(ax, ay, bx, by, px, py):
    return bool(load().orc_left_of(ax, ay, bx, by, px, py))


def seg_isect(*a):
    return bool(load().orc_seg_isect(*a))


def f_down(v):
    return load().orc_f_down(v)


def f_up(v):
    return load().orc_f_up(v)


def de9im_code(ev, bits, la, lb):
    e = np.ascontiguousarray(ev, dtype=np.float64)
    return int(load().orc_de9im_code(_p(e, C.c_double), bits, la, lb))


class Layer:
    def __init__(self, h, n, points=False):
        self.h, self.n, self.points = h, n, points
        self._keep = None

    @classmethod
    def polys(cls, hs):
        gro = np.ascontiguousarray(hs.gro, np.uint64)
        rvo = np.ascontiguousarray(hs.rvo, np.uint64)
        xy = np.ascontiguousarray(hs.xy, np.float64).reshape(-1, 2)
        par = None if hs.parent is None else np.ascontiguousarray(hs.parent, np.uint32)
        syn = None if hs.syn is None else np.ascontiguousarray(hs.syn, np.uint8)
        h = load().orc_layer_polys(_p(gro, C.c_uint64), _p(rvo, C.c_uint64), _p(xy, C.c_double), len(gro) - 1, len(rvo) - 1, len(xy),
                                   _p(par, C.c_uint32), _p(syn, C.c_uint8))
        L = cls(h, len(gro) - 1)
        L.nv = len(xy)
        return L

    @classmethod
    def points_layer(cls, xy):
        xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
        return cls(load().orc_layer_points(_p(xy, C.c_double), len(xy)), len(xy), points=True)

    @property
    def nparent(self):
        return int(load().orc_layer_nparent(self.h))

    def export(self):
        if self.points:
            bb = np.zeros((self.n, 4), np.float32)
            load().orc_layer_export(self.h, None, None, _p(bb, C.c_float), None, None)
            return {"bb": bb}
        origin, ext, bb = np.zeros((self.n, 2)), np.zeros((self.n, 2)), np.zeros((self.n, 4), np.float32)
        q, blen = np.zeros((self.nv, 2)), np.zeros(self.nparent)
        load().orc_layer_export(self.h, _p(origin, C.c_double), _p(ext, C.c_double), _p(bb, C.c_float), _p(q, C.c_double), _p(blen, C.c_double))
        return {"origin": origin, "ext": ext, "bb": bb, "q": q, "blen": blen}

    def __del__(self):
        try:
            if self.h:
                load().orc_layer_free(self.h)
        except Exception:
            pass


def candidates(P: Layer, R: Layer, upper=False):
    lib = load()
    n = lib.orc_candidates(P.h, R.h, int(upper), None, 0)
    out = np.zeros((max(n, 1), 2), np.uint32)
    lib.orc_candidates(P.h, R.h, int(upper), _p(out, C.c_uint32), n)
    return out[:n]


def refine(P: Layer, R: Layer, pairs, want_area=True, want_relate=False, snap_half_ulp=False):
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    n = len(pairs)
    isect, area, ev, evb = np.zeros(n, np.uint8), np.zeros(n), np.zeros((n, 4)), np.zeros(n, np.uint32)
    load().orc_refine(P.h, R.h, _p(pairs, C.c_uint32), n, int(want_area), int(want_relate), int(snap_half_ulp), _p(isect, C.c_uint8),
                      _p(area, C.c_double), _p(ev, C.c_double), _p(evb, C.c_uint32))
    return isect, area, ev, evb


def _result(r, area=False, code=False, where=False):
    lib = load()
    n = lib.orc_result_rows(r)
    pairs = np.zeros((max(n, 1), 2), np.uint32)
    a = np.zeros(max(n, 1)) if area else None
    c = np.zeros(max(n, 1), np.uint32) if code else None
    w = np.zeros(max(n, 1), np.uint8) if where else None
    lib.orc_result_copy(r, _p(pairs, C.c_uint32), _p(a, C.c_double), _p(c, C.c_uint32), _p(w, C.c_uint8))
    out = {"pairs": pairs[:n], "candidates": int(lib.orc_result_candidates(r))}
    if area:
        out["area"] = a[:n]
    if code:
        out["code"] = c[:n]
    if where:
        out["where"] = w[:n]
    return out


def join(P: Layer, R: Layer, want_area=True, want_relate=False, upper=False, no_agg=False):
    lib = load()
    r = lib.orc_result_new()
    try:
        lib.orc_join(P.h, R.h, int(want_area), int(want_relate), int(upper), int(no_agg), r)
        return _result(r, area=want_area, code=want_relate)
    finally:
        lib.orc_result_free(r)


def points_join(pts: Layer, R: Layer):
    lib = load()
    r = lib.orc_result_new()
    try:
        lib.orc_points_join(pts.h, R.h, r)
        return _result(r, where=True)
    finally:
        lib.orc_result_free(r)


def pip_pair(pts: Layer, R: Layer, i: int, g: int) -> int:
    return int(load().orc_pip_pair(pts.h, R.h, i, g))


def cluster(P: Layer):
    lib = load()
    r = lib.orc_result_new()
    try:
        lib.orc_cluster(P.h, r)
        n = lib.orc_result_nroot(r)
        root = np.zeros(max(n, 1), np.uint32)
        lib.orc_result_roots(r, _p(root, C.c_uint32))
        out = _result(r)
        out["root"] = root[:n]
        return out
    finally:
        lib.orc_result_free(r)


def subdivide(hs, maxv):
    lib = load()
    gro, rvo = np.ascontiguousarray(hs.gro, np.uint64), np.ascontiguousarray(hs.rvo, np.uint64)
    xy = np.ascontiguousarray(hs.xy, np.float64).reshape(-1, 2)
    h = lib.orc_subdivide(_p(gro, C.c_uint64), _p(rvo, C.c_uint64), _p(xy, C.c_double), len(gro) - 1, int(maxv))
    try:
        n, nr, nv = C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib.orc_flat_sizes(h, C.byref(n), C.byref(nr), C.byref(nv))
        og, orv = np.zeros(n.value + 1, np.uint64), np.zeros(nr.value + 1, np.uint64)
        oxy, osyn, opar = np.zeros((max(nv.value, 1), 2)), np.zeros(max(nv.value, 1), np.uint8), np.zeros(max(n.value, 1), np.uint32)
        lib.orc_flat_copy(h, _p(og, C.c_uint64), _p(orv, C.c_uint64), _p(oxy, C.c_double), _p(osyn, C.c_uint8), _p(opar, C.c_uint32))
        return og, orv, oxy[: nv.value], opar[: n.value], osyn[: nv.value]
    finally:
        lib.orc_flat_free(h)


def dedup(hs):
    lib = load()
    gro, rvo = np.ascontiguousarray(hs.gro, np.uint64), np.ascontiguousarray(hs.rvo, np.uint64)
    xy = np.ascontiguousarray(hs.xy, np.float64).reshape(-1, 2)
    n = len(gro) - 1
    h = lib.orc_dedup(_p(gro, C.c_uint64), _p(rvo, C.c_uint64), _p(xy, C.c_double), n)
    try:
        un = lib.orc_dedup_n(h)
        first, moff, midx = np.zeros(max(un, 1), np.uint32), np.zeros(un + 1, np.uint64), np.zeros(max(n, 1), np.uint32)
        lib.orc_dedup_copy(h, _p(first, C.c_uint32), _p(moff, C.c_uint64), _p(midx, C.c_uint32))
        return first[:un], moff, midx[:n]
    finally:
        lib.orc_dedup_free(h)


FLOP_KINDS = ("add", "mul", "fma", "div", "sqrt")


def refine_flops(hs_P, hs_R, pairs, want_area=True, want_relate=False):
    """Algorithmic FLOPs of refining `pairs` (counting build of the oracle, one thread): dict of add/sub, mul, fma,
    div, sqrt counts and `flops` = add + mul + 2 fma + div + sqrt."""
    lib = load(flops=True)
    P = _FlopLayer(lib, hs_P)
    R = P if hs_R is hs_P else _FlopLayer(lib, hs_R)
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    n = len(pairs)
    isect, area, ev, evb = np.zeros(n, np.uint8), np.zeros(n), np.zeros((n, 4)), np.zeros(n, np.uint32)
    cnt = np.zeros(11, np.uint64)
    lib.orc_flops(_p(cnt, C.c_uint64), 1)
    lib.orc_refine(P.h, R.h, _p(pairs, C.c_uint32), n, int(want_area), int(want_relate), 0, _p(isect, C.c_uint8), _p(area, C.c_double),
                   _p(ev, C.c_double), _p(evb, C.c_uint32))
    lib.orc_flops(_p(cnt, C.c_uint64), 1)
    out = dict(zip(FLOP_KINDS, (int(v) for v in cnt[:5])))
    out["flops"] = out["add"] + out["mul"] + 2 * out["fma"] + out["div"] + out["sqrt"]
    out["pairs"] = n
    # the edge-contribution phase alone, and the pairs that reached it
    e = dict(zip(FLOP_KINDS, (int(v) for v in cnt[5:10])))
    out["edge"] = e
    out["edge_flops"] = e["add"] + e["mul"] + 2 * e["fma"] + e["div"] + e["sqrt"]
    out["edge_pairs"] = int(cnt[10])
    return out


class _FlopLayer:
    def __init__(self, lib, hs):
        self.lib = lib
        gro, rvo = np.ascontiguousarray(hs.gro, np.uint64), np.ascontiguousarray(hs.rvo, np.uint64)
        xy = np.ascontiguousarray(hs.xy, np.float64).reshape(-1, 2)
        par = None if hs.parent is None else np.ascontiguousarray(hs.parent, np.uint32)
        syn = None if hs.syn is None else np.ascontiguousarray(hs.syn, np.uint8)
        self.h = lib.orc_layer_polys(_p(gro, C.c_uint64), _p(rvo, C.c_uint64), _p(xy, C.c_double), len(gro) - 1, len(rvo) - 1, len(xy),
                                     _p(par, C.c_uint32), _p(syn, C.c_uint8))

    def __del__(self):
        try:
            self.lib.orc_layer_free(self.h)
        except Exception:
            pass


# ------------------------------------------------------------------------------------------- distance joins
def gm(name, *a):
    """Portable elementary function of the engine (sin, cos, tan, atan, atan2) as built on the host."""
    return getattr(load(), "orc_gm_" + name)(*a)


def m_per_deg(lat):
    kx, ky = C.c_double(), C.c_double()
    load().orc_geo_m_per_deg(lat, C.byref(kx), C.byref(ky))
    return kx.value, ky.value


def deg_cover(y0, y1, d, larger=True):
    dx, dy = C.c_double(), C.c_double()
    load().orc_geo_deg_cover(y0, y1, d, int(larger), C.byref(dx), C.byref(dy))
    return dx.value, dy.value


def utm_k(glon, glat, lon, lat):
    return load().orc_geo_utm_k(glon, glat, lon, lat)


def vincenty(lon1, lat1, lon2, lat2):
    return load().orc_geo_vincenty(lon1, lat1, lon2, lat2)


def geo_frame(lat_lo, lat_hi):
    """metric frame of a window's latitude range: (kx, ky, rel, curv)"""
    out = np.zeros(4)
    load().orc_geo_frame(lat_lo, lat_hi, _p(out, C.c_double))
    return tuple(out.tolist())


def geog_band(lat_lo, lat_hi, d, La2, Lb2=0.0):
    return float(load().orc_geo_band(lat_lo, lat_hi, d, La2, Lb2))


def seg_seg_d2(a0, a1, b0, b1):
    return float(load().orc_geo_seg_seg_d2(*a0, *a1, *b0, *b1))


def geog_dist(a0, a1, b0, b1=None):
    """Geography distance (m) of edge a0-a1 from edge b0-b1, or from point b0 when b1 is None."""
    pt = b1 is None
    b1 = b0 if pt else b1
    return load().orc_geo_geog_dist(a0[0], a0[1], a1[0], a1[1], b0[0], b0[1], b1[0], b1[1], int(pt))


def dw_pair(P: Layer, R: Layer, a: int, b: int, d: float, report=True):
    dist = C.c_double()
    hit = load().orc_dw_pair(P.h, R.h, a, b, d, int(report), C.byref(dist))
    return bool(hit), dist.value


def _dw_result(r):
    lib = load()
    n = lib.orc_dw_rows(r)
    pairs = np.zeros((n, 2), np.uint32)
    dist = np.zeros(n, np.float64)
    lib.orc_dw_copy(r, _p(pairs, C.c_uint32), _p(dist, C.c_double))
    out = {"pairs": pairs, "dist": dist, "candidates": lib.orc_dw_candidates(r)}
    lib.orc_dw_free(r)
    return out


def dwithin(P: Layer, R: Layer, d: float, same=False, upper=False, report=False):
    lib = load()
    r = lib.orc_dw_new()
    lib.orc_dwithin(P.h, R.h, d, int(same), int(upper), int(report), r)
    return _dw_result(r)


def points_dwithin(pts: Layer, R: Layer, d: float, report=False):
    lib = load()
    r = lib.orc_dw_new()
    lib.orc_points_dwithin(pts.h, R.h, d, int(report), r)
    return _dw_result(r)


def er_pair(P: Layer, R: Layer, a: int, b: int, ra: float, rb: float) -> bool:
    """Eroded intersects of parent a of P and parent b of R (all their pieces)."""
    return bool(load().orc_er_pair(P.h, R.h, a, b, ra, rb))


def erode(P: Layer, R: Layer, ra: float, rb: float, same=False, upper=False, need_self=None):
    """ST_Intersects(ST_Buffer(A, -ra), ST_Buffer(B, -rb)) rows (parent pairs)."""
    lib = load()
    r = lib.orc_dw_new()
    ns = None if need_self is None else np.ascontiguousarray(need_self, np.uint8)
    lib.orc_erode(P.h, R.h, ra, rb, int(same), int(upper), _p(ns, C.c_uint8), r)
    out = _dw_result(r)
    del out["dist"]
    return out


def points_erode(pts: Layer, R: Layer, r: float):
    """Points inside ST_Buffer(polygon, -r): rows (point index, parent polygon)."""
    lib = load()
    h = lib.orc_dw_new()
    lib.orc_points_erode(pts.h, R.h, r, h)
    out = _dw_result(h)
    del out["dist"]
    return out



# ---------------------------------------------------------------------------------------------- binary COPY / EWKB parse
def copy_parse(buf: bytes, header=True, kind=0):
    """The layer loader's parse of a binary COPY stream of (id, geom) rows (oracle/ewkb.cpp): dict of the counters
    (rows, nulls, bad, empty, polys, points), consumed / done, the polygon arrays gro, rvo, xy (n, 2), the points pxy
    (n, 2), ids (str per accepted row) and row (stream row of each accepted row). kind: the layer kind an earlier piece
    of the stream decided (1 polygons, 2 points)."""
    lib = load()
    raw = C.create_string_buffer(bytes(buf), max(len(buf), 1))
    h = lib.orc_copy_parse(C.cast(raw, C.POINTER(C.c_uint8)), len(buf), int(header), int(kind))
    try:
        err = lib.orc_copy_error(h).decode()
        if err:
            raise ValueError(err)
        sz = np.zeros(12, np.uint64)
        lib.orc_copy_sizes(h, _p(sz, C.c_uint64))
        rows, nulls, bad, empty, polys, points, consumed, done, n, nr, nv, npts = (int(v) for v in sz)
        gro, rvo, xy, pxy, ido, ids, row = (C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint64)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)(),
                                            C.POINTER(C.c_uint64)(), C.c_char_p(), C.POINTER(C.c_uint64)())
        lib.orc_copy_arrays(h, C.byref(gro), C.byref(rvo), C.byref(xy), C.byref(pxy), C.byref(ido), C.byref(ids), C.byref(row))
        nacc = polys + points

        def arr(p, cnt, dt):
            return np.ctypeslib.as_array(p, (cnt,)).copy() if cnt else np.zeros(0, dt)

        id_off = arr(ido, nacc + 1, np.uint64)
        blob = C.string_at(ids, int(id_off[-1])) if nacc else b""
        return {"rows": rows, "nulls": nulls, "bad": bad, "empty": empty, "polys": polys, "points": points, "consumed": consumed,
                "done": bool(done), "gro": arr(gro, n + 1, np.uint64), "rvo": arr(rvo, nr + 1, np.uint64),
                "xy": arr(xy, 2 * nv, np.float64).reshape(-1, 2), "pxy": arr(pxy, 2 * npts, np.float64).reshape(-1, 2),
                "ids": [blob[int(id_off[k]):int(id_off[k + 1])].decode() for k in range(nacc)], "row": arr(row, nacc, np.uint64),
                "stats": {"rows": rows, "nulls": nulls, "bad": bad, "empty": empty, "polys": polys, "points": points}}
    finally:
        lib.orc_copy_free(h)


def ring_twice_area(xy):
    """twice the signed area of a closed ring (n, 2) as the loader orients rings (oracle/ewkb.cpp twice_area)"""
    a = np.ascontiguousarray(xy, np.float64).reshape(-1)
    return load().orc_ring_twice_area(_p(a, C.c_double), len(a) // 2)
