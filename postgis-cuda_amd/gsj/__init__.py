"""Python bindings (ctypes) for libgsj, the MI355X spatial join engine.

Mirrors the engine entry points of the reference (landstats-us/postgis-cuda engine/engine.cuh):
``Context.join`` ~ join_stream, ``Context.points_in_polygons`` ~ join_points_stream, ``Context.cluster`` ~
cluster_intersecting; host-side preparation ``subdivide`` ~ engine/subdivide.h, ``dedup`` ~ engine/dedup.h,
``read_gset``/``read_pset`` ~ engine/gset.h. The compute path is the HIP library only: if ``libgsj.so`` is not
built, importing the bindings raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libgsj.so")

GSJ_OK = 0
ERRORS = {-1: "ARG", -2: "HIP", -3: "NOMEM", -4: "LIMIT", -5: "CANCEL", -6: "IO", -7: "INTERNAL"}
DE9IM_PT_INTERIOR = 1 | 3 << 12 | 2 << 14 | 3 << 16
DE9IM_PT_BOUNDARY = 1 << 2 | 3 << 12 | 2 << 14 | 3 << 16


class GsjError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gsj error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class JoinOpts(C.Structure):
    _fields_ = [("want_area", C.c_int), ("want_relate", C.c_int), ("upper", C.c_int), ("snap_half_ulp", C.c_int),
                ("batch_q", C.c_uint32), ("max_pairs", C.c_uint64), ("no_agg", C.c_int), ("cancel", C.POINTER(C.c_int)),
                ("has_window", C.c_int), ("window", C.c_double * 4)]


class JoinStats(C.Structure):
    _fields_ = [("traverse_ms", C.c_double), ("refine_ms", C.c_double), ("compact_ms", C.c_double), ("copy_ms", C.c_double),
                ("candidates", C.c_uint64), ("hits", C.c_uint64), ("heavy", C.c_uint64), ("batches", C.c_uint64),
                ("k_small_ms", C.c_double), ("k_large_ms", C.c_double), ("k_over_ms", C.c_double),
                ("n_small", C.c_uint64), ("n_large", C.c_uint64), ("n_over", C.c_uint64), ("launches", C.c_uint64),
                ("k_medium_ms", C.c_double), ("n_medium", C.c_uint64), ("k_screen_ms", C.c_double), ("n_small_edge", C.c_uint64),
                ("n_medium_edge", C.c_uint64), ("k_mixed_ms", C.c_double), ("n_mixed", C.c_uint64),
                ("k_tiny_ms", C.c_double), ("n_tiny", C.c_uint64), ("n_tiny_edge", C.c_uint64), ("k_traverse_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LayerInfo(C.Structure):
    _fields_ = [("n", C.c_uint64), ("nr", C.c_uint64), ("nv", C.c_uint64), ("nparent", C.c_uint32), ("is_points", C.c_int),
                ("has_bvh", C.c_int), ("device_bytes", C.c_uint64), ("bvh_ms", C.c_double)]


ROWS_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_size_t)
POINTS_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.c_size_t)
DIST_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.c_size_t)
DONE_CB = C.CFUNCTYPE(None, C.c_void_p)

REQ_JOIN, REQ_POINTS_IN_POLYGONS, REQ_CLUSTER, REQ_DWITHIN, REQ_POINTS_DWITHIN = 1, 2, 3, 4, 5
REQ_INTERSECTS_BUFFERED, REQ_POINTS_INTERSECTS_BUFFERED, REQ_CLUSTER_ROWS = 6, 7, 8


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("P", C.c_void_p), ("R", C.c_void_p), ("opts", JoinOpts), ("d", C.c_double), ("same", C.c_int),
                ("report_distance", C.c_int), ("ra", C.c_double), ("rb", C.c_double), ("need_self", C.POINTER(C.c_uint8)),
                ("rows_cb", ROWS_CB), ("points_cb", POINTS_CB), ("dist_cb", DIST_CB), ("user", C.c_void_p),
                ("root", C.POINTER(C.c_uint32)), ("nroot", C.c_uint64), ("need_listed", C.c_int),
                ("need_list", C.POINTER(C.c_uint32)), ("n_need", C.c_uint64), ("done_cb", DONE_CB), ("done_user", C.c_void_p)]


class RequestTimes(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double), ("lane", C.c_int), ("batch", C.c_int)]


_lib = None

# every symbol include/gsj.h declares (tests check the library exports all of them)
EXPORTS = [
    "gsj_last_error", "gsj_version", "gsj_ctx_create", "gsj_ctx_destroy", "gsj_ctx_status_json", "gsj_ctx_trim", "gsj_ctx_set_tuning", "gsj_ctx_tuning",
    "gsj_layer_create_polys", "gsj_layer_create_points", "gsj_layer_free", "gsj_bvh_build", "gsj_layer_info_get", "gsj_layer_export",
    "gsj_join", "gsj_join_noalloc", "gsj_points_in_polygons", "gsj_cluster", "gsj_cluster_rows", "gsj_dwithin", "gsj_points_dwithin", "gsj_intersects_buffered",
    "gsj_points_intersects_buffered",
    "gsj_sched_create", "gsj_sched_destroy", "gsj_sched_set_batch", "gsj_submit", "gsj_wait", "gsj_poll", "gsj_de9im_str", "gsj_de9im_transpose", "gsj_de9im_match",
    "gsj_hostset_free", "gsj_hostset_sizes", "gsj_hostset_arrays", "gsj_gset_read", "gsj_gset_write", "gsj_pset_read", "gsj_pset_write",
    "gsj_free", "gsj_subdivide", "gsj_dedup", "gsj_dedup_device", "gsj_subdivide_device", "gsj_generate", "gsj_generate_points",
    "gsj_copy_parse_device", "gsj_copy_write",
]


def load(path: str | None = None):
    """Load libgsj.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("GSJ_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise ImportError(f"libgsj.so not found at {p}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    vp, u64, u32, i32, dp = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_double)
    u64p, u32p, u8p, fp = C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.POINTER(C.c_float)
    sig = {
        "gsj_last_error": (C.c_char_p, []), "gsj_version": (C.c_char_p, []),
        "gsj_ctx_create": (i32, [i32, C.POINTER(vp)]), "gsj_ctx_destroy": (None, [vp]), "gsj_ctx_status_json": (C.c_char_p, [vp]),
        "gsj_ctx_trim": (i32, [vp]), "gsj_ctx_set_tuning": (i32, [vp, C.c_char_p]), "gsj_ctx_tuning": (C.c_char_p, [vp]),
        "gsj_layer_create_polys": (i32, [vp, u64p, u64p, dp, u64, u64, u64, u32p, u8p, C.POINTER(vp)]),
        "gsj_layer_create_points": (i32, [vp, dp, u64, C.POINTER(vp)]), "gsj_layer_free": (None, [vp]), "gsj_bvh_build": (i32, [vp]),
        "gsj_layer_info_get": (i32, [vp, C.POINTER(LayerInfo)]), "gsj_layer_export": (i32, [vp, dp, dp, fp, dp, dp]),
        "gsj_join": (i32, [vp, vp, vp, C.POINTER(JoinOpts), ROWS_CB, vp, C.POINTER(JoinStats)]),
        "gsj_join_noalloc": (i32, [vp, vp, vp, C.POINTER(JoinOpts), ROWS_CB, vp, C.POINTER(JoinStats)]),
        "gsj_points_in_polygons": (i32, [vp, vp, vp, C.POINTER(JoinOpts), POINTS_CB, vp, C.POINTER(JoinStats)]),
        "gsj_cluster": (i32, [vp, vp, C.POINTER(JoinOpts), u32p, u64, C.POINTER(JoinStats)]),
        "gsj_cluster_rows": (i32, [vp, vp, C.POINTER(JoinOpts), ROWS_CB, vp, C.POINTER(JoinStats)]),
        "gsj_dwithin": (i32, [vp, vp, vp, C.c_double, i32, i32, C.POINTER(JoinOpts), DIST_CB, vp, C.POINTER(JoinStats)]),
        "gsj_points_dwithin": (i32, [vp, vp, vp, C.c_double, i32, C.POINTER(JoinOpts), DIST_CB, vp, C.POINTER(JoinStats)]),
        "gsj_intersects_buffered": (i32, [vp, vp, vp, C.c_double, C.c_double, i32, u8p, C.POINTER(JoinOpts), DIST_CB, vp, C.POINTER(JoinStats)]),
        "gsj_points_intersects_buffered": (i32, [vp, vp, vp, C.c_double, C.POINTER(JoinOpts), DIST_CB, vp, C.POINTER(JoinStats)]),
        "gsj_sched_create": (i32, [vp, i32, u64, C.POINTER(vp)]), "gsj_sched_destroy": (None, [vp]), "gsj_sched_set_batch": (i32, [vp, i32]),
        "gsj_submit": (i32, [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
        "gsj_wait": (i32, [vp, u64, C.POINTER(JoinStats), C.POINTER(RequestTimes)]), "gsj_poll": (i32, [vp, u64]),
        "gsj_de9im_str": (None, [u32, C.c_char_p]), "gsj_de9im_transpose": (u32, [u32]), "gsj_de9im_match": (i32, [u32, C.c_char_p]),
        "gsj_hostset_free": (None, [vp]), "gsj_hostset_sizes": (i32, [vp, u64p, u64p, u64p]),
        "gsj_hostset_arrays": (i32, [vp, C.POINTER(u64p), C.POINTER(u64p), C.POINTER(dp), C.POINTER(u32p), C.POINTER(u8p)]),
        "gsj_gset_read": (i32, [C.c_char_p, C.POINTER(vp)]), "gsj_gset_write": (i32, [C.c_char_p, u64p, u64p, dp, u64, u64, u64]),
        "gsj_pset_read": (i32, [C.c_char_p, C.POINTER(dp), u64p]), "gsj_pset_write": (i32, [C.c_char_p, dp, u64]),
        "gsj_free": (None, [vp]), "gsj_subdivide": (i32, [u64p, u64p, dp, u64, i32, C.POINTER(vp)]),
        "gsj_dedup": (i32, [u64p, u64p, dp, u64, C.POINTER(vp), u64p, u32p, u64p]),
        "gsj_dedup_device": (i32, [vp, u64p, u64p, dp, u64, C.POINTER(vp), u64p, u32p, u64p]),
        "gsj_subdivide_device": (i32, [vp, u64p, u64p, dp, u64, i32, C.POINTER(vp)]),
        "gsj_generate": (i32, [C.c_char_p, C.c_char_p, C.POINTER(vp)]), "gsj_generate_points": (i32, [C.c_char_p, C.POINTER(dp), u64p]),
        "gsj_copy_parse_device": (i32, [vp, C.c_void_p, u64, i32, i32, C.POINTER(vp), C.POINTER(dp), u64p, C.POINTER(C.c_void_p),
                                        C.POINTER(u64p), C.POINTER(CopyStats)]),
        "gsj_copy_write": (i32, [u64p, u64p, dp, u64, C.c_char_p, u64p, i32, C.POINTER(C.c_void_p), u64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != GSJ_OK:
        raise GsjError(rc, _lib.gsj_last_error().decode())


def _ptr(a, ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


# ------------------------------------------------------------------------------------------- host sets
@dataclass
class HostSet:
    """Flat polygon set in the .gset layout (engine/gset.h:5-11) plus optional subdivision metadata."""
    gro: np.ndarray
    rvo: np.ndarray
    xy: np.ndarray                  # (nv, 2) float64
    parent: np.ndarray | None = None
    syn: np.ndarray | None = None
    ids: list = field(default_factory=list)

    @property
    def n(self):
        return len(self.gro) - 1

    @property
    def nr(self):
        return len(self.rvo) - 1

    @property
    def nv(self):
        return len(self.xy)

    def normalized(self):
        self.gro = np.ascontiguousarray(self.gro, dtype=np.uint64)
        self.rvo = np.ascontiguousarray(self.rvo, dtype=np.uint64)
        self.xy = np.ascontiguousarray(np.asarray(self.xy, dtype=np.float64).reshape(-1, 2))
        if self.parent is not None:
            self.parent = np.ascontiguousarray(self.parent, dtype=np.uint32)
        if self.syn is not None:
            self.syn = np.ascontiguousarray(self.syn, dtype=np.uint8)
        return self

    def geometry(self, g):
        """rings of geometry g as a list of (k, 2) arrays (closed)"""
        return [self.xy[self.rvo[r]:self.rvo[r + 1]] for r in range(int(self.gro[g]), int(self.gro[g + 1]))]


class _NativeSet:
    """Owns a library-allocated host set; freed when the last array viewing it is gone."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        if self.h:
            load().gsj_hostset_free(self.h)
            self.h = None


class _NativeBuffer:
    """Owns a buffer the library allocated with malloc (gsj_free)."""

    def __init__(self, p):
        self.p = p

    def __del__(self):
        if self.p:
            load().gsj_free(self.p)
            self.p = None


class _View:
    """numpy view of library memory (the array's base keeps the owner alive)"""

    def __init__(self, owner, ptr, shape, typestr):
        self.owner = owner
        self.__array_interface__ = {"data": (C.cast(ptr, C.c_void_p).value, False), "shape": shape, "typestr": typestr, "version": 3}


def _take_hostset(h) -> HostSet:
    """wrap a library host set without copying it (national-scale layers are tens of GB)"""
    lib = load()
    owner = _NativeSet(h)
    n, nr, nv = C.c_uint64(), C.c_uint64(), C.c_uint64()
    _check(lib.gsj_hostset_sizes(h, C.byref(n), C.byref(nr), C.byref(nv)))
    gro, rvo, xy = C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint64)(), C.POINTER(C.c_double)()
    par, syn = C.POINTER(C.c_uint32)(), C.POINTER(C.c_uint8)()
    _check(lib.gsj_hostset_arrays(h, C.byref(gro), C.byref(rvo), C.byref(xy), C.byref(par), C.byref(syn)))

    def view(ptr, count, typestr, empty_dtype):
        return np.asarray(_View(owner, ptr, (count,), typestr)) if count and ptr else np.zeros(0, empty_dtype)

    out = HostSet(view(gro, n.value + 1, "<u8", np.uint64), view(rvo, nr.value + 1, "<u8", np.uint64),
                  view(xy, 2 * nv.value, "<f8", np.float64).reshape(-1, 2))
    if par:
        out.parent = view(par, n.value, "<u4", np.uint32)
        out.syn = view(syn, nv.value, "|u1", np.uint8)
    return out.normalized()


def _spec(kw):
    def fmt(v):
        if isinstance(v, (float, np.floating)):
            return repr(float(v))   # shortest text that reads back to the same double
        if isinstance(v, np.integer):
            return str(int(v))
        return str(v)

    return ";".join(f"{k}={fmt(v)}" for k, v in kw.items())


def generate(kind: str, **spec) -> HostSet:
    """Deterministic synthetic layer: kind in parcels_grid (config 1 parcels), regions_grid (config 1 regions),
    parcels_mesh (config 2 parcels)."""
    lib = load()
    h = C.c_void_p()
    _check(lib.gsj_generate(kind.encode(), _spec(spec).encode(), C.byref(h)))
    return _take_hostset(h)


def generate_points(**spec) -> np.ndarray:
    lib = load()
    p, n = C.POINTER(C.c_double)(), C.c_uint64()
    _check(lib.gsj_generate_points(_spec(spec).encode(), C.byref(p), C.byref(n)))
    if not n.value:
        lib.gsj_free(p)
        return np.zeros((0, 2))
    return np.asarray(_View(_NativeBuffer(p), p, (2 * n.value,), "<f8")).reshape(-1, 2)


def subdivide(hs: HostSet, maxv: int, ctx: "Context" = None) -> HostSet:
    """pieces of at most maxv stored vertices; with ctx, computed on its device (the same result)"""
    lib = load()
    hs = hs.normalized()
    h = C.c_void_p()
    args = (_ptr(hs.gro, C.c_uint64), _ptr(hs.rvo, C.c_uint64), _ptr(hs.xy, C.c_double), hs.n, maxv, C.byref(h))
    _check(lib.gsj_subdivide_device(ctx.h, *args) if ctx is not None else lib.gsj_subdivide(*args))
    return _take_hostset(h)


def dedup(hs: HostSet, ctx: "Context" = None):
    """(distinct set, moff[un+1], midx[n]); with ctx, computed on its device (the same result)"""
    lib = load()
    hs = hs.normalized()
    h = C.c_void_p()
    moff = np.zeros(hs.n + 1, np.uint64)
    midx = np.zeros(max(hs.n, 1), np.uint32)
    un = C.c_uint64()
    args = (_ptr(hs.gro, C.c_uint64), _ptr(hs.rvo, C.c_uint64), _ptr(hs.xy, C.c_double), hs.n, C.byref(h), _ptr(moff, C.c_uint64),
            _ptr(midx, C.c_uint32), C.byref(un))
    _check(lib.gsj_dedup_device(ctx.h, *args) if ctx is not None else lib.gsj_dedup(*args))
    return _take_hostset(h), moff[: un.value + 1].copy(), midx[: hs.n].copy()


class CopyStats(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("skipped_null", C.c_uint64), ("skipped_type", C.c_uint64), ("skipped_empty", C.c_uint64),
                ("polys", C.c_uint64), ("points", C.c_uint64), ("consumed", C.c_uint64), ("done", C.c_int), ("kind", C.c_int),
                ("scan_ms", C.c_double), ("emit_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


@dataclass
class CopyLayer:
    """a parsed binary COPY stream: polys (polygon layer) or pts (n, 2) (point layer), ids per accepted row"""
    polys: HostSet | None
    pts: np.ndarray | None
    ids: list
    stats: dict


def copy_parse(ctx: "Context", buf: bytes, header=True, kind=0, want_ids=True) -> CopyLayer:
    """gsj_copy_parse_device: the rows of a binary COPY stream of (id text, geom) parsed on ctx's device; stats carries
    the loader's counters, consumed / done, and the layer kind for the next piece of the same stream. want_ids=False
    leaves ids as None (no Python string per row)."""
    lib = load()
    raw = C.create_string_buffer(bytes(buf), max(len(buf), 1))
    h, p, n = C.c_void_p(), C.POINTER(C.c_double)(), C.c_uint64()
    ids, off, st = C.c_void_p(), C.POINTER(C.c_uint64)(), CopyStats()
    _check(lib.gsj_copy_parse_device(ctx.h, C.cast(raw, C.c_void_p), len(buf), int(header), int(kind), C.byref(h), C.byref(p), C.byref(n),
                                     C.byref(ids), C.byref(off), C.byref(st)))
    try:
        nacc = st.polys + st.points
        o = np.ctypeslib.as_array(off, (nacc + 1,)).copy()
        blob = C.string_at(ids.value, int(o[-1])) if nacc and o[-1] else b""
        idl = [blob[int(o[k]):int(o[k + 1])].decode() for k in range(nacc)] if want_ids else None
        pts = np.ctypeslib.as_array(p, (2 * n.value,)).copy().reshape(-1, 2) if n.value else (np.zeros((0, 2)) if p else None)
    finally:
        lib.gsj_free(ids)
        lib.gsj_free(C.cast(off, C.c_void_p))
        if p:
            lib.gsj_free(C.cast(p, C.c_void_p))
    polys = _take_hostset(h) if h.value else None
    return CopyLayer(polys, pts, idl, st.as_dict())


def copy_write(hs: HostSet, ids=None, srid=0) -> bytes:
    """gsj_copy_write: the set as a binary COPY stream of (id text, geom EWKB) rows (ids: str per geometry, default
    the row numbers)"""
    lib = load()
    hs = hs.normalized()
    out, ln = C.c_void_p(), C.c_uint64()
    if ids is not None:
        enc = [str(i).encode() for i in ids]
        off = np.zeros(len(enc) + 1, np.uint64)
        off[1:] = np.cumsum([len(e) for e in enc])
        blob = b"".join(enc)
        idp, offp = C.c_char_p(blob), _ptr(off, C.c_uint64)
    else:
        idp, offp = None, None
    _check(lib.gsj_copy_write(_ptr(hs.gro, C.c_uint64), _ptr(hs.rvo, C.c_uint64), _ptr(hs.xy, C.c_double), hs.n, idp, offp, int(srid),
                              C.byref(out), C.byref(ln)))
    try:
        return C.string_at(out.value, ln.value)
    finally:
        lib.gsj_free(out)


def read_gset(path: str) -> HostSet:
    lib = load()
    h = C.c_void_p()
    _check(lib.gsj_gset_read(path.encode(), C.byref(h)))
    hs = _take_hostset(h)
    ids = path + ".ids"
    if os.path.exists(ids):
        with open(ids) as f:
            hs.ids = f.read().splitlines()
    return hs


def write_gset(path: str, hs: HostSet, ids=None):
    lib = load()
    hs = hs.normalized()
    _check(lib.gsj_gset_write(path.encode(), _ptr(hs.gro, C.c_uint64), _ptr(hs.rvo, C.c_uint64), _ptr(hs.xy, C.c_double), hs.n, hs.nr, hs.nv))
    ids = ids if ids is not None else (hs.ids or [str(i) for i in range(hs.n)])
    with open(path + ".ids", "w") as f:
        f.write("\n".join(map(str, ids)) + "\n")


def read_pset(path: str) -> np.ndarray:
    lib = load()
    p, n = C.POINTER(C.c_double)(), C.c_uint64()
    _check(lib.gsj_pset_read(path.encode(), C.byref(p), C.byref(n)))
    try:
        return np.ctypeslib.as_array(p, (2 * n.value,)).copy().reshape(-1, 2) if n.value else np.zeros((0, 2))
    finally:
        lib.gsj_free(p)


def write_pset(path: str, xy: np.ndarray, ids=None):
    lib = load()
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    _check(lib.gsj_pset_write(path.encode(), _ptr(xy, C.c_double), len(xy)))
    ids = ids if ids is not None else [str(i) for i in range(len(xy))]
    with open(path + ".ids", "w") as f:
        f.write("\n".join(map(str, ids)) + "\n")


def de9im_str(code: int) -> str:
    return "".join("F012"[(int(code) >> (2 * k)) & 3] for k in range(9))


def de9im_transpose(code: int) -> int:
    return int(load().gsj_de9im_transpose(int(code)))


def de9im_match(code: int, pattern: str) -> bool:
    return bool(load().gsj_de9im_match(int(code), pattern.encode()))


# ------------------------------------------------------------------------------------------- device
class Layer:
    def __init__(self, ctx: "Context", handle, kind: str, hs: HostSet | None = None):
        self.ctx, self.h, self.kind, self.host = ctx, handle, kind, hs

    def build_bvh(self):
        _check(load().gsj_bvh_build(self.h))
        return self

    def info(self) -> LayerInfo:
        i = LayerInfo()
        _check(load().gsj_layer_info_get(self.h, C.byref(i)))
        return i

    def export(self):
        """prepared device arrays: origin (n,2), ext (n,2), bb (n,4) float32, q (nv,2), blen (nparent,)"""
        i = self.info()
        if i.is_points:
            bb = np.zeros((i.n, 4), np.float32)
            _check(load().gsj_layer_export(self.h, None, None, _ptr(bb, C.c_float), None, None))
            return {"bb": bb}
        origin, ext = np.zeros((i.n, 2)), np.zeros((i.n, 2))
        bb, q, blen = np.zeros((i.n, 4), np.float32), np.zeros((i.nv, 2)), np.zeros(i.nparent)
        _check(load().gsj_layer_export(self.h, _ptr(origin, C.c_double), _ptr(ext, C.c_double), _ptr(bb, C.c_float), _ptr(q, C.c_double),
                                       _ptr(blen, C.c_double)))
        return {"origin": origin, "ext": ext, "bb": bb, "q": q, "blen": blen}

    def free(self):
        if self.h and self.ctx.h:
            load().gsj_layer_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Rows:
    pairs: np.ndarray           # (n, 2) uint32 parent ids
    area: np.ndarray | None
    code: np.ndarray | None
    where: np.ndarray | None
    stats: dict
    dist: np.ndarray | None = None   # distance joins with report_distance: metres per row
    root: np.ndarray | None = None   # cluster: smallest parent index of each parent's component


def _opts(want_area=False, want_relate=False, upper=False, snap_half_ulp=False, batch_q=0, max_pairs=0, no_agg=False, window=None):
    """window = (x0, y0, x1, y1): probe only the parents with a piece box meeting the rectangle"""
    o = JoinOpts()
    o.want_area, o.want_relate, o.upper, o.snap_half_ulp = int(want_area), int(want_relate), int(upper), int(snap_half_ulp)
    o.batch_q, o.max_pairs, o.no_agg = batch_q, max_pairs, int(no_agg)
    if window is not None:
        o.has_window = 1
        o.window[:] = [float(v) for v in window]
    return o


class Context:
    """One HIP device + stream; layers and joins live on it."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        _check(lib.gsj_ctx_create(device, C.byref(h)))
        self.h = h
        self._layers = weakref.WeakSet()

    def status(self) -> str:
        return load().gsj_ctx_status_json(self.h).decode()

    def trim(self):
        """free the context's grow-only scratch (it grows again on the next call)"""
        _check(load().gsj_ctx_trim(self.h))

    @property
    def tuning(self) -> str:
        """the launch-shape table in effect (text form of gsj_ctx_set_tuning)"""
        return load().gsj_ctx_tuning(self.h).decode()

    def set_tuning(self, text: str):
        """replace the launch-shape table; the rows of every call are the same under any table"""
        _check(load().gsj_ctx_set_tuning(self.h, text.encode()))

    def polys(self, hs: HostSet, build_bvh=True) -> Layer:
        hs = hs.normalized()
        h = C.c_void_p()
        _check(load().gsj_layer_create_polys(self.h, _ptr(hs.gro, C.c_uint64), _ptr(hs.rvo, C.c_uint64), _ptr(hs.xy, C.c_double), hs.n, hs.nr,
                                            hs.nv, _ptr(hs.parent, C.c_uint32), _ptr(hs.syn, C.c_uint8), C.byref(h)))
        L = Layer(self, h, "polygon", hs)
        self._layers.add(L)
        return L.build_bvh() if build_bvh else L

    def points(self, xy: np.ndarray, build_bvh=False) -> Layer:
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        h = C.c_void_p()
        _check(load().gsj_layer_create_points(self.h, _ptr(xy, C.c_double), len(xy), C.byref(h)))
        L = Layer(self, h, "point")
        self._layers.add(L)
        return L.build_bvh() if build_bvh else L

    def join(self, P: Layer, R: Layer, collect=True, noalloc=False, **kw) -> Rows:
        """Rows of the join; collect=False delivers them to a no-op callback (counts and stats only); noalloc=True runs
        gsj_join_noalloc (fixed scratch, no allocation; same rows)."""
        o = _opts(**kw)
        chunks = []

        def cb(user, pairs, area, code, n):
            if not collect:
                return 0
            p = np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2)
            a = np.ctypeslib.as_array(area, (n,)).copy() if area else None
            c = np.ctypeslib.as_array(code, (n,)).copy() if code else None
            chunks.append((p, a, c))
            return 0

        st = JoinStats()
        entry = load().gsj_join_noalloc if noalloc else load().gsj_join
        _check(entry(self.h, P.h, R.h, C.byref(o), ROWS_CB(cb), None, C.byref(st)))
        pairs = np.concatenate([c[0] for c in chunks]) if chunks else np.zeros((0, 2), np.uint32)
        area = np.concatenate([c[1] for c in chunks]) if chunks and o.want_area else (np.zeros(0) if o.want_area else None)
        code = np.concatenate([c[2] for c in chunks]) if chunks and o.want_relate else (np.zeros(0, np.uint32) if o.want_relate else None)
        return Rows(pairs, area, code, None, st.as_dict())

    def points_in_polygons(self, pts: Layer, R: Layer, collect=True, **kw) -> Rows:
        o = _opts(**kw)
        chunks = []

        def cb(user, pairs, where, n):
            if collect:
                chunks.append((np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2), np.ctypeslib.as_array(where, (n,)).copy()))
            return 0

        st = JoinStats()
        _check(load().gsj_points_in_polygons(self.h, pts.h, R.h, C.byref(o), POINTS_CB(cb), None, C.byref(st)))
        pairs = np.concatenate([c[0] for c in chunks]) if chunks else np.zeros((0, 2), np.uint32)
        where = np.concatenate([c[1] for c in chunks]) if chunks else np.zeros(0, np.uint8)
        return Rows(pairs, None, None, where, st.as_dict())

    def _dist_rows(self, fn, *args, collect=True, **kw):
        o = _opts(**kw)
        chunks = []

        def cb(user, pairs, dist, n):
            if not collect:
                return 0
            p = np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2)
            chunks.append((p, np.ctypeslib.as_array(dist, (n,)).copy() if dist else None))
            return 0

        st = JoinStats()
        _check(fn(self.h, *args, C.byref(o), DIST_CB(cb), None, C.byref(st)))
        pairs = np.concatenate([c[0] for c in chunks]) if chunks else np.zeros((0, 2), np.uint32)
        dist = np.concatenate([c[1] for c in chunks]) if chunks and chunks[0][1] is not None else None
        return Rows(pairs, None, None, None, st.as_dict(), dist)

    def dwithin(self, P: Layer, R: Layer, d: float, same=False, report_distance=False, **kw) -> Rows:
        """ST_DWithin(geography) rows (parent pairs); dist per row in metres with report_distance."""
        r = self._dist_rows(load().gsj_dwithin, P.h, R.h, float(d), int(same), int(report_distance), **kw)
        if report_distance and r.dist is None:
            r.dist = np.zeros(0)
        return r

    def points_dwithin(self, pts: Layer, R: Layer, d: float, report_distance=False, **kw) -> Rows:
        """ST_DWithin(geography) of points against polygons: rows (point index, parent polygon)."""
        r = self._dist_rows(load().gsj_points_dwithin, pts.h, R.h, float(d), int(report_distance), **kw)
        if report_distance and r.dist is None:
            r.dist = np.zeros(0)
        return r

    def intersects_buffered(self, P: Layer, R: Layer, ra: float, rb: float, same=False, need_self=None, **kw) -> Rows:
        """ST_Intersects(ST_Buffer(a::geography, -ra), ST_Buffer(b::geography, -rb)) rows (parent pairs)."""
        ns = None if need_self is None else np.ascontiguousarray(need_self, np.uint8)
        return self._dist_rows(load().gsj_intersects_buffered, P.h, R.h, float(ra), float(rb), int(same), _ptr(ns, C.c_uint8), **kw)

    def points_intersects_buffered(self, pts: Layer, R: Layer, r: float, **kw) -> Rows:
        """Points inside ST_Buffer(polygon::geography, -r): rows (point index, parent polygon)."""
        return self._dist_rows(load().gsj_points_intersects_buffered, pts.h, R.h, float(r), **kw)

    def cluster(self, P: Layer, **kw):
        o = _opts(**kw)
        n = P.info().nparent
        root = np.zeros(max(n, 1), np.uint32)
        st = JoinStats()
        _check(load().gsj_cluster(self.h, P.h, C.byref(o), _ptr(root, C.c_uint32), n, C.byref(st)))
        return root[:n], st.as_dict()

    def cluster_rows(self, P: Layer, **kw) -> Rows:
        """(parent, component root) rows in ascending parent; with window=, the clusters of the parents meeting it."""
        o = _opts(**kw)
        chunks = []

        def cb(user, pairs, area, code, n):
            chunks.append(np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2))
            return 0

        st = JoinStats()
        _check(load().gsj_cluster_rows(self.h, P.h, C.byref(o), ROWS_CB(cb), None, C.byref(st)))
        return Rows(np.concatenate(chunks) if chunks else np.zeros((0, 2), np.uint32), None, None, None, st.as_dict())

    def close(self):
        if self.h:
            for L in list(self._layers):   # layers are device objects of this context: free them first
                L.free()
            load().gsj_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scheduler:
    """Served path: requests on resident layers of `ctx`, run by `lanes` lanes with continuous admission.

    submit() returns a ticket; wait(ticket) returns the request's Rows (as the direct call would) with
    `times` (queue / run / total ms, lane, batch) in its stats. collect=False passes no callback: rows are computed,
    counted (stats["hits"]) and copied to the host, but not delivered. batch: most windowed joins run in one pass
    (gsj_sched_set_batch; 1 = each alone)."""

    def __init__(self, ctx: "Context", lanes: int = 4, scratch_budget: int = 0, batch: int = 16):
        self.h = None
        h = C.c_void_p()
        _check(load().gsj_sched_create(ctx.h, lanes, scratch_budget, C.byref(h)))
        rc = load().gsj_sched_set_batch(h, batch)
        if rc != 0:
            msg = _lib.gsj_last_error().decode()
            load().gsj_sched_destroy(h)
            raise GsjError(rc, msg)
        self.h = h
        self.ctx = ctx
        self._pending = {}

    def submit(self, kind: int, P: "Layer", R: "Layer" = None, collect=True, d=0.0, same=False, report_distance=False, ra=0.0, rb=0.0,
               need_self=None, need_list=None, on_done=None, **kw) -> int:
        """need_list (buffered requests, instead of need_self): the parents whose pair with itself is tested, ascending
        (an empty list: none). on_done(): called once on the lane's thread when the request has finished, before wait
        returns for it (gsj_request.done_cb)."""
        import threading

        req = Request()
        req.kind, req.P, req.R, req.opts = kind, P.h, (R.h if R is not None else None), _opts(**kw)
        req.d, req.same, req.report_distance, req.ra, req.rb = float(d), int(same), int(report_distance), float(ra), float(rb)
        chunks, lock = [], threading.Lock()
        keep = [req]
        if need_self is not None:
            ns = np.ascontiguousarray(need_self, np.uint8)
            req.need_self = _ptr(ns, C.c_uint8)
            keep.append(ns)
        if need_list is not None:
            nl = np.ascontiguousarray(need_list, np.uint32)
            req.need_listed, req.need_list, req.n_need = 1, _ptr(nl, C.c_uint32), len(nl)
            keep.append(nl)
        root = None
        if kind == REQ_CLUSTER:
            n = P.info().nparent
            root = np.zeros(max(n, 1), np.uint32)
            req.root, req.nroot = _ptr(root, C.c_uint32), n
        elif collect and kind in (REQ_JOIN, REQ_CLUSTER_ROWS):
            def cb(user, pairs, area, code, n):
                p = np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2)
                with lock:
                    chunks.append((p, np.ctypeslib.as_array(area, (n,)).copy() if area else None,
                                   np.ctypeslib.as_array(code, (n,)).copy() if code else None))
                return 0
            req.rows_cb = ROWS_CB(cb)
            keep.append(req.rows_cb)
        elif collect and kind == REQ_POINTS_IN_POLYGONS:
            def cb(user, pairs, where, n):
                with lock:
                    chunks.append((np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2), np.ctypeslib.as_array(where, (n,)).copy()))
                return 0
            req.points_cb = POINTS_CB(cb)
            keep.append(req.points_cb)
        elif collect:
            def cb(user, pairs, dist, n):
                with lock:
                    chunks.append((np.ctypeslib.as_array(pairs, (2 * n,)).copy().reshape(-1, 2),
                                   np.ctypeslib.as_array(dist, (n,)).copy() if dist else None))
                return 0
            req.dist_cb = DIST_CB(cb)
            keep.append(req.dist_cb)
        if on_done is not None:
            req.done_cb = DONE_CB(lambda user: on_done())
            keep.append(req.done_cb)
        t = C.c_uint64()
        _check(load().gsj_submit(self.h, C.byref(req), C.byref(t)))
        self._pending[t.value] = (kind, chunks, keep, root, req.opts)
        return t.value

    def poll(self, ticket: int) -> bool:
        rc = load().gsj_poll(self.h, ticket)
        if rc < 0:
            _check(rc)
        return rc == 1

    def wait(self, ticket: int) -> Rows:
        kind, chunks, keep, root, o = self._pending.pop(ticket)
        st, tm = JoinStats(), RequestTimes()
        rc = load().gsj_wait(self.h, ticket, C.byref(st), C.byref(tm))
        _check(rc)
        stats = st.as_dict()
        stats["times"] = {"queue_ms": tm.queue_ms, "run_ms": tm.run_ms, "total_ms": tm.total_ms, "lane": tm.lane, "batch": tm.batch}
        pairs = np.concatenate([c[0] for c in chunks]) if chunks else np.zeros((0, 2), np.uint32)
        if kind == REQ_CLUSTER:
            return Rows(np.zeros((0, 2), np.uint32), None, None, None, stats, None, root[:keep[0].nroot])
        if kind == REQ_CLUSTER_ROWS:
            return Rows(pairs, None, None, None, stats)
        if kind == REQ_JOIN:
            area = np.concatenate([c[1] for c in chunks]) if chunks and o.want_area else None
            code = np.concatenate([c[2] for c in chunks]) if chunks and o.want_relate else None
            return Rows(pairs, area, code, None, stats)
        if kind == REQ_POINTS_IN_POLYGONS:
            where = np.concatenate([c[1] for c in chunks]) if chunks else np.zeros(0, np.uint8)
            return Rows(pairs, None, None, where, stats)
        dist = np.concatenate([c[1] for c in chunks]) if chunks and chunks[0][1] is not None else None
        return Rows(pairs, None, None, None, stats, dist)

    def close(self):
        if self.h:
            load().gsj_sched_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

