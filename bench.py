#!/usr/bin/env python3
"""Benchmark: refined pairs/s of the resident-layer spatial join on one MI355X.

Workload (default): the 2.4M-parcel layer of the config-2 generator (1549 x 1549 jittered lattice over 1 x 1 degree,
lognormal vertex counts, ~30% shared edges, ~5% overlaps, seed 1), subdivided at 256 vertices, self-joined with
full DE-9IM relate (each unordered pair once). One step = one request: traversal + FP64 refinement + compaction +
piece -> parent aggregation + DE-9IM codes + copy of the rows to pinned host memory, with the layer and its LBVH
resident in HBM before the timed region. `value` = refined (candidate) pairs per second.

Beside it: the roofline of the dominant refinement kernel (algorithmic FP64 FLOPs from the oracle's counting build
x pairs per launch / HIP-event duration), the oracle timed on the host cores on a bounded sample (cpu_baseline), and
the parity of the engine against the oracle on that sample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale S] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "postgis-cuda_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
METRIC = "refined pairs/s + p50/p99 request latency, served vs serial, 0 pair diffs"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mesh(gsj, n, seed=1, subdiv=256):
    M = gsj.generate("parcels_mesh", nx=n, ny=n, w=1 / 1549, h=1 / 1549, seed=seed)
    return M, gsj.subdivide(M, subdiv)


def pair_class(hs, pairs):
    """Refinement work class of each candidate pair of a self-join of hs (the engine's partition): 4 tiny (both pieces
    <= 16 stored vertices), 0 small (both <= 32), 1 medium (both <= 64), all three with <= 8 rings; 2 mixed (one piece
    <= 32 vertices); 3 large."""
    return pair_class2(hs, hs, pairs)


def pair_class2(hp, hr, pairs):
    """pair_class for a join of two sets (probe hp, build hr)"""
    def shape(hs):
        return np.diff(hs.gro.astype(np.int64)), hs.rvo[hs.gro[1:]].astype(np.int64) - hs.rvo[hs.gro[:-1]].astype(np.int64)
    (ra, va), (rb, vb) = shape(hp), shape(hr)
    va, vb = va[pairs[:, 0]], vb[pairs[:, 1]]
    rok = (ra[pairs[:, 0]] <= 8) & (rb[pairs[:, 1]] <= 8)
    tiny = rok & (va <= 16) & (vb <= 16)
    small = ~tiny & rok & (va <= 32) & (vb <= 32)
    med = ~tiny & ~small & rok & (va <= 64) & (vb <= 64)
    mixed = ~tiny & ~small & ~med & (np.minimum(va, vb) <= 32)
    return np.where(tiny, 4, np.where(small, 0, np.where(med, 1, np.where(mixed, 2, 3))))


def time_cpu(fn, seconds):
    """run fn() until about `seconds` of host time; returns (total seconds, repeats, last result)"""
    reps, t, r = 0, 0.0, None
    while reps == 0 or t < seconds:
        t1 = time.perf_counter()
        r = fn()
        t += time.perf_counter() - t1
        reps += 1
    return t, reps, r


def host_threads():
    return min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))


def parse_leg(gsj, ctx, M, args):
    """Layer load from a binary COPY stream, the loader's input: the config-2 parcels as (id text, EWKB with SRID) rows,
    parsed on the device (gsj_copy_parse_device: host framing, upload, row walk, ring records and oriented
    coordinates, download, ids). Roofline of the device passes: compulsory bytes = the rows read once + 16 B per vertex
    + 8 B per ring and per geometry written + 8 B per row of the row map, over the passes' HIP-event time. Beside it
    (cpu_baseline leg) the oracle's restatement of the loader's parse on one host thread, and parity with it."""
    buf = gsj.copy_write(M, srid=4326)
    d = gsj.copy_parse(ctx, buf)   # warm-up: the parse's buffers reach their size (and the ids, for parity)
    best = float("inf")
    for _ in range(3):   # timed: the C ABI call (framing, upload, passes, download, ids as one buffer)
        t0 = time.perf_counter()
        gsj.copy_parse(ctx, buf, want_ids=False)
        best = min(best, time.perf_counter() - t0)
    nr, nv = int(M.gro[-1]), len(M.xy)
    kms = d.stats["scan_ms"] + d.stats["emit_ms"]
    bytes_c = len(buf) + 16 * nv + 8 * (nr + 1) + 8 * (M.n + 1) + 8 * M.n
    gbs = bytes_c / (kms * 1e-3) / 1e9
    out = {"workload": f"config-2 parcels as a binary COPY stream: {M.n} rows, {nv} vertices, {len(buf) / 1e6:.1f} MB (EWKB, SRID 4326)",
           "value": round(M.n / best, 1), "unit": "rows/s", "ms": round(best * 1e3, 2),
           "kernel_ms": {"scan": round(d.stats["scan_ms"], 3), "emit": round(d.stats["emit_ms"], 3)},
           "roofline": {"kernel": "k_ewkb_scan + k_ewkb_rings + k_ewkb_coords", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None, "compulsory_bytes": bytes_c, "launch_ms": round(kms, 3)},
           "same_as_input": bool(np.array_equal(d.polys.xy.view(np.uint64), M.xy.view(np.uint64)) and np.array_equal(d.polys.rvo, M.rvo))}
    log(f"parse: {M.n} rows in {best * 1e3:.1f} ms end to end, device passes {kms:.2f} ms ({gbs:.0f} GB/s compulsory)")
    if not args.no_cpu:
        from oracle import oracle as O

        t0 = time.perf_counter()
        o = O.copy_parse(buf)
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(M.n / t_cpu, 1), "unit": "rows/s", "cores": 1, "kind": "port",
                               "sample": f"the whole stream once ({t_cpu:.2f} s)"}
        out["parity"] = {"rows_equal": o["polys"] == d.stats["polys"] and o["ids"] == d.ids,
                         "arrays_bit_identical": bool(np.array_equal(o["xy"].view(np.uint64), d.polys.xy.view(np.uint64))
                                                      and np.array_equal(o["gro"], d.polys.gro) and np.array_equal(o["rvo"], d.polys.rvo))}
        log(f"parse cpu: {t_cpu:.2f} s; parity {out['parity']}")
    return out


def config1_leg(gsj, ctx, args):
    """Parcels x regions intersects + intersection area: 100 x 100 parcels of 4-12 vertices under 16 regions of 2000
    vertices, seed 0. One step = one request (layers resident). The dominant kernel is the mixed class (a parcel
    against a region: 16-lane groups over the region's edges near the parcel); its FLOPs are the counting oracle's
    for every candidate pair of the request."""
    from oracle import oracle as O

    P = gsj.generate("parcels_grid", nx=100, ny=100, seed=0)
    R = gsj.generate("regions_grid", nx=100, ny=100, kx=4, ky=4, edge_segs=500, seed=0)
    LP, LR = ctx.polys(P, build_bvh=False), ctx.polys(R)
    for _ in range(3):
        ctx.join(LP, LR, want_area=True, collect=False)
    steps = max(10, 20 * args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.join(LP, LR, want_area=True, collect=False)
    wall = time.perf_counter() - t0
    st = r.stats
    # the longest kernel among the classes that had pairs (an empty class's launches still take a few microseconds)
    kname, kfield, kcount = max(((k, f, n) for k, f, n in (("k_refine_mixed", "k_mixed_ms", "n_mixed"), ("k_refine_large", "k_large_ms", "n_large"),
                                                           ("k_screen_small", "k_screen_ms", "n_small")) if st[n] > 0), key=lambda k: st[k[1]])
    OP, OR = O.Layer.polys(P), O.Layer.polys(R)
    cand = O.candidates(OP, OR)
    cls = pair_class2(P, R, cand)
    pick = cand[cls == {"k_refine_mixed": 2, "k_refine_large": 3, "k_screen_small": 0}[kname]]
    fl = O.refine_flops(P, R, pick, want_area=True)
    per_launch_ms = st[kfield] / max(st["launches"], 1)
    achieved = fl["flops"] * (st[kcount] / max(len(pick), 1)) / max(st["launches"], 1) / (per_launch_ms * 1e-3) / 1e12
    O.set_threads(host_threads())
    t_cpu, reps, exp = time_cpu(lambda: O.join(OP, OR, want_area=True), args.cpu_seconds / 2)
    got = ctx.join(LP, LR, want_area=True)
    g, e = np.lexsort((got.pairs[:, 1], got.pairs[:, 0])), np.lexsort((exp["pairs"][:, 1], exp["pairs"][:, 0]))
    parity = {"rows_gpu": len(got.pairs), "rows_oracle": len(exp["pairs"]), "pairs_equal": bool(np.array_equal(got.pairs[g], exp["pairs"][e])),
              "areas_bit_identical": bool(np.array_equal(got.area[g], exp["area"][e])), "candidates_equal": got.stats["candidates"] == exp["candidates"]}
    out = {"workload": "config 1: 10000 parcels (4-12 vertices) x 16 regions (2000 vertices), intersects + area, seed 0",
           "value": round(st["candidates"] * steps / wall, 1), "unit": "refined pairs/s", "ms_per_request": round(1e3 * wall / steps, 4),
           "candidates": st["candidates"], "rows": st["hits"],
           "roofline": {"kernel": kname, "bound": "valu", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / FP64_PEAK_TFLOPS, 6), "traffic": None, "flops_per_pair": round(fl["flops"] / max(len(pick), 1), 1),
                        "pairs_per_launch": int(st[kcount] / max(st["launches"], 1)), "launch_ms": round(per_launch_ms, 4),
                        "flops_counted": "the whole refinement of the class's pairs"},
           "cpu_baseline": {"value": round(exp["candidates"] * reps / t_cpu, 1), "unit": "refined pairs/s", "cores": host_threads(), "kind": "port",
                            "sample": f"the whole request ({exp['candidates']} candidates), run {reps} times in {t_cpu:.1f} s"},
           "parity": parity}
    LP.free()
    LR.free()
    return out


def config3_leg(gsj, ctx, args, L, S):
    """Points in polygons and geodesic dwithin. PIP: 5M points (half uniform over the 1 x 1 degree extent, half in 1000
    clusters of sigma 50 m, seed 2) against the resident 2.4M-parcel layer L (subdivided pieces S; rows are parents).
    dwithin: self-join of a 1.2M-parcel layer of the same generator over 0.6 x 0.6 degrees from (-95.2, 38.6), d = 5 m
    on WGS84, with the distance of every row. Rooflines: HBM, algorithmic bytes = what each candidate's test reads
    once (point 16 B + pair ids 8 B + the piece's per-geometry records 40 B + 16 B per stored vertex; dwithin: pair ids
    8 B + both pieces' records and vertices) and writes (1 B flag, 8 B distance), over the kernel's HIP-event time."""
    from oracle import oracle as O

    out = {}
    xy = gsj.generate_points(n=5_000_000, x0=-95.5, y0=38.5, w=1.0, h=1.0, clusters=1000, sigma_m=50.0, seed=2)
    pts = ctx.points(xy)
    for _ in range(2):
        ctx.points_in_polygons(pts, L, collect=False)
    steps = max(3, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.points_in_polygons(pts, L, collect=False)
    wall = time.perf_counter() - t0
    st = r.stats
    nv_piece = S.nv / S.n
    bytes_per = 16 + 8 + 40 + 16 * nv_piece + 1
    k_ms = st["k_small_ms"] / max(st["launches"], 1)
    gbs = bytes_per * st["candidates"] / max(st["launches"], 1) / (k_ms * 1e-3) / 1e9
    nsamp = 500_000
    O.set_threads(host_threads())
    OS = O.Layer.polys(S)
    Osamp = O.Layer.points_layer(xy[:nsamp])
    t_cpu, reps, exp = time_cpu(lambda: O.points_join(Osamp, OS), args.cpu_seconds / 2)
    got = ctx.points_in_polygons(ctx.points(xy[:nsamp]), L)
    g, e = np.lexsort((got.pairs[:, 1], got.pairs[:, 0])), np.lexsort((exp["pairs"][:, 1], exp["pairs"][:, 0]))
    out["points_in_polygons"] = {
        "workload": "config 3: 5M points (50% uniform, 50% in 1000 clusters of sigma 50 m, seed 2) in the 2.4M-parcel layer (subdivide 256)",
        "value": round(st["candidates"] * steps / wall, 1), "unit": "refined pairs/s", "ms_per_request": round(1e3 * wall / steps, 3),
        "candidates": st["candidates"], "rows": st["hits"],
        "roofline": {"kernel": "k_pip", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                     "traffic": None, "bytes_per_pair": round(bytes_per, 1), "launch_ms": round(k_ms, 4)},
        "cpu_baseline": {"value": round(exp["candidates"] * reps / t_cpu, 1), "unit": "refined pairs/s", "cores": host_threads(), "kind": "port",
                         "sample": f"the first {nsamp} points ({exp['candidates']} candidates), run {reps} times in {t_cpu:.1f} s"},
        "parity": {"rows_gpu": len(got.pairs), "rows_oracle": len(exp["pairs"]),
                   "rows_equal": bool(np.array_equal(got.pairs[g], exp["pairs"][e]) and np.array_equal(got.where[g], exp["where"][e]))}}
    pts.free()
    del Osamp, xy
    n12 = 1095
    M12 = gsj.generate("parcels_mesh", nx=n12, ny=n12, x0=-95.2, y0=38.6, w=0.6 / n12, h=0.6 / n12, seed=2)
    L12 = ctx.polys(M12)
    kw = dict(same=True, upper=True, report_distance=True, collect=False)
    ctx.dwithin(L12, L12, 5.0, **kw)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = ctx.dwithin(L12, L12, 5.0, **kw)
    wall = time.perf_counter() - t0
    st = r.stats
    kms = {k: st[f"k_{k}_ms"] for k in ("tiny", "small", "medium", "mixed", "large", "over")}
    kbig = max(kms, key=kms.get)
    nv_geom = M12.nv / M12.n
    bytes_dw = 8 + 2 * (40 + 16 * nv_geom) + 1 + 8
    n_k = {"tiny": "n_tiny", "small": "n_small", "medium": "n_medium", "mixed": "n_mixed", "large": "n_large", "over": "n_over"}[kbig]
    k_ms = kms[kbig] / max(st["launches"], 1)
    gbs = bytes_dw * st[n_k] / max(st["launches"], 1) / (k_ms * 1e-3) / 1e9
    OM = O.Layer.polys(M12)
    t_cpu, reps, exp = time_cpu(lambda: O.dwithin(OM, OM, 5.0, same=True, upper=True, report=True), args.cpu_seconds / 2)
    got = ctx.dwithin(L12, L12, 5.0, same=True, upper=True, report_distance=True)
    g, e = np.lexsort((got.pairs[:, 1], got.pairs[:, 0])), np.lexsort((exp["pairs"][:, 1], exp["pairs"][:, 0]))
    same_rows = bool(np.array_equal(got.pairs[g], exp["pairs"][e]))
    out["self_dwithin_5m"] = {
        "workload": f"config 3: self dwithin 5 m with distances, {M12.n} parcels over 0.6 x 0.6 degrees from (-95.2, 38.6), seed 2",
        "value": round(st["candidates"] * steps / wall, 1), "unit": "refined pairs/s", "ms_per_request": round(1e3 * wall / steps, 3),
        "candidates": st["candidates"], "rows": st["hits"], "kernel_ms": {k: round(v, 3) for k, v in kms.items()},
        "roofline": {"kernel": f"k_dw_lists ({kbig} class)" if kbig != "over" else "k_dw_overflow", "bound": "hbm", "achieved": round(gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None, "bytes_per_pair": round(bytes_dw, 1),
                     "launch_ms": round(k_ms, 4)},
        "cpu_baseline": {"value": round(exp["candidates"] * reps / t_cpu, 1), "unit": "refined pairs/s", "cores": host_threads(), "kind": "port",
                         "sample": f"the whole self-join ({exp['candidates']} candidates), run {reps} times in {t_cpu:.1f} s"},
        "parity": {"rows_gpu": len(got.pairs), "rows_oracle": len(exp["pairs"]), "rows_equal": same_rows,
                   "max_distance_diff_m": float(np.max(np.abs(got.dist[g] - exp["dist"][e]))) if same_rows and len(g) else None}}
    L12.free()
    return out


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "refine_relate_2.4M.md")


def summary_column(kernel, ncells, col):
    """Column `col` of the committed counter summary's table of `ncells` cells for `kernel`; a name without template
    arguments sums its instantiations (k_screen_small: the tiny, small and medium screens, timed together as the
    engine's k_screen_ms). None when the summary lacks the kernel."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    total, found = 0, False
    for line in open(TRAFFIC_SUMMARY):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) != ncells:
            continue
        name = cells[0].strip("`")
        if name == kernel or ("<" not in kernel and name.startswith(kernel + "<")):
            try:
                total += int(float(cells[col]))
                found = True
            except ValueError:
                return None
    return total if found else None


def hbm_traffic(kernel):
    """HBM bytes per dispatch of `kernel` on this workload, from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE passes, each scaled by the copy calibration recorded there); None when the summary lacks the kernel.
    Rows of the summary's traffic table read: | `kernel` | fetch KiB | write KiB | HBM bytes | (the trace table has 5
    columns)."""
    return summary_column(kernel, 4, 3)


def other_rooflines(n, st, bvh_ms):
    """HBM rooflines of the LBVH build and the traversal, with the compulsory bytes of each: build 72 B per primitive
    (read its 16 B box; key 8 + value 4, sorted id 4, node boxes 2 x 16, parents 2 x 4, children 8 written once);
    traversal 16 B per probe box + 8 B per candidate written + 44 B per primitive of tree read once. Times are the
    engine's HIP-event device times: the build's kernels (gsj_layer_info.bvh_ms), the walks of the step's candidate
    search (k_traverse_ms: count and write passes)."""
    build_b = 72.0 * n
    trav_b = 16.0 * n + 8.0 * st["candidates"] + 44.0 * n
    out = []
    for name, b, ms in (("lbvh_build", build_b, bvh_ms), ("traversal", trav_b, st["k_traverse_ms"])):
        gbs = b / max(ms * 1e-3, 1e-9) / 1e9
        out.append({"kernel": name, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 5), "compulsory_bytes": int(b), "ms": round(ms, 3), "timing": "HIP events"})
    return out


def issued_f64(kernel):
    """Issued FP64 lane-FLOPs per dispatch of `kernel` (64 x (ADD + MUL + 2 FMA) wave instructions, whatever the exec
    mask) from the committed counter summary's FP64 instruction table (7 cells a row); None when it lacks the kernel."""
    return summary_column(kernel, 7, 6)


def copy_calibration(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size through the HIP runtime (its copy kernel), so that the HBM byte
    counters of a profiling run can be scaled: FETCH_SIZE / WRITE_SIZE of these dispatches should read nbytes each."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {reps} device copies of {nbytes} bytes")


def config4_mix(n, smax, seed=3):
    """The batched-requests configuration's requests: sizes lognormal in [1e3, smax] polygons, kinds 40% intersects +
    area, 20% relate, 20% dwithin (d ~ U(1, 50) m), 20% eroded intersects (ra, rb ~ U(0.5, 3) m). One dict per request:
    kind ("area", "relate", "dwithin", "erode"), side (cells per side of its config-1 layers), layer seed, parameters."""
    rng = np.random.default_rng(seed)
    mu, sig = np.log(np.sqrt(1e3 * smax)), np.log(smax / 1e3) / 6
    sizes = np.clip(np.exp(rng.normal(mu, sig, n)), 1e3, smax).astype(np.int64)
    kinds = rng.choice(4, size=n, p=[0.4, 0.2, 0.2, 0.2])
    dists = rng.uniform(1.0, 50.0, n)
    radii = rng.uniform(0.5, 3.0, (n, 2))
    return [{"kind": ("area", "relate", "dwithin", "erode")[kinds[i]], "side": max(1, int(round(np.sqrt(sizes[i])))), "seed": 1000 * seed + i,
             "d": float(dists[i]), "ra": float(radii[i, 0]), "rb": float(radii[i, 1])} for i in range(n)]


def config4_layers(gsj, ctx, q):
    """device layers of one config-4 request: parcels (and regions unless it is a dwithin self-join)"""
    P = gsj.generate("parcels_grid", nx=q["side"], ny=q["side"], seed=q["seed"])
    if q["kind"] == "dwithin":
        return P, None, ctx.polys(P), None
    R = gsj.generate("regions_grid", nx=q["side"], ny=q["side"], kx=4, ky=4, edge_segs=500, seed=q["seed"])
    return P, R, ctx.polys(P, build_bvh=False), ctx.polys(R)


def config4_submit_args(gsj, q, LP, LR):
    """(kind, P, R, keyword arguments) of a config-4 request for gsj.Scheduler.submit"""
    if q["kind"] == "dwithin":
        return gsj.REQ_DWITHIN, LP, LP, dict(d=q["d"], same=True, upper=True)
    if q["kind"] == "erode":
        return gsj.REQ_INTERSECTS_BUFFERED, LP, LR, dict(ra=q["ra"], rb=q["rb"])
    return gsj.REQ_JOIN, LP, LR, dict(want_area=q["kind"] == "area", want_relate=q["kind"] == "relate")


def window_stream(nreq, x0, y0, W, H, seed=4):
    """The serving configuration's windowed requests over one resident parcel layer: windows of area lognormal in
    [1e-4, 10] square degrees (clipped to the extent), kinds 35% intersects + area, 20% relate, 20% dwithin
    (d ~ U(1, 50) m), 20% eroded intersects (ra, rb ~ U(0.5, 3) m) as self joins, 5% cluster_intersecting of the
    window's parcels. One (kind name, keyword arguments) per request."""
    rng = np.random.default_rng(seed)
    mu, sig = np.log(np.sqrt(1e-4 * 10.0)), np.log(10.0 / 1e-4) / 6
    area = np.clip(np.exp(rng.normal(mu, sig, nreq)), 1e-4, 10.0)
    side = np.minimum(np.sqrt(area), min(W, H))
    cx, cy = x0 + rng.uniform(0, 1, nreq) * (W - side) + side / 2, y0 + rng.uniform(0, 1, nreq) * (H - side) + side / 2
    kinds = rng.choice(5, size=nreq, p=[0.35, 0.2, 0.2, 0.2, 0.05])
    dists, radii = rng.uniform(1.0, 50.0, nreq), rng.uniform(0.5, 3.0, (nreq, 2))
    out = []
    for i in range(nreq):
        win = (cx[i] - side[i] / 2, cy[i] - side[i] / 2, cx[i] + side[i] / 2, cy[i] + side[i] / 2)
        if kinds[i] == 2:
            out.append(("dwithin", dict(d=float(dists[i]), same=True, upper=True, window=win)))
        elif kinds[i] == 3:
            out.append(("erode", dict(ra=float(radii[i, 0]), rb=float(radii[i, 1]), same=True, upper=True, window=win)))
        elif kinds[i] == 4:
            out.append(("cluster", dict(window=win)))
        else:
            out.append(("area" if kinds[i] == 0 else "relate", dict(want_area=kinds[i] == 0, want_relate=kinds[i] == 1, upper=True, window=win)))
    return out


def window_submit_kind(gsj, name):
    return {"dwithin": gsj.REQ_DWITHIN, "erode": gsj.REQ_INTERSECTS_BUFFERED, "cluster": gsj.REQ_CLUSTER_ROWS}.get(name, gsj.REQ_JOIN)


def serving_leg(gsj, ctx, args):
    """Served vs serial on a mixed stream of uneven requests (the batched-requests configuration: sizes lognormal in
    [1k, 2M] polygons of the config-1 generator on sqrt(s) x sqrt(s) cells, seed 3; 40% intersects + area and 20%
    relate of parcels x regions, 20% dwithin self-joins with d ~ U(1, 50) m, 20% eroded intersects of parcels x
    regions with ra, rb ~ U(0.5, 3) m). Every request arrives at t = 0; serial = one lane (one request at a time),
    served = `lanes` lanes with continuous admission. Latency = submit -> finished."""
    n = args.requests
    mix = config4_mix(n, args.smax)
    t0 = time.time()
    reqs, layers = [], []
    for q in mix:
        P, R, LP, LR = config4_layers(gsj, ctx, q)
        reqs.append(config4_submit_args(gsj, q, LP, LR))
        layers += [L for L in (LP, LR) if L is not None]
        del P, R
    sizes = np.array([q["side"] ** 2 for q in mix])
    log(f"serving mix: {n} requests, {int(sizes.sum())} polygons, generated + resident in {time.time() - t0:.1f}s")

    def run(lanes):
        s = gsj.Scheduler(ctx, lanes=lanes, batch=1 if lanes == 1 else 16)   # serial: one request at a time
        try:
            for k, A, B, kw in reqs:   # warm-up: every lane's scratch grows to its working size
                s.submit(k, A, B, collect=False, **kw)
            for t in list(s._pending):
                s.wait(t)
            tw = time.perf_counter()
            tickets = [s.submit(k, A, B, collect=False, **kw) for k, A, B, kw in reqs]
            res = [s.wait(t) for t in tickets]
            wall = time.perf_counter() - tw
        finally:
            s.close()
        cand = sum(r.stats["candidates"] for r in res)
        rows = sum(r.stats["hits"] for r in res)
        lat = np.array([r.stats["times"]["total_ms"] for r in res])
        return {"lanes": lanes, "wall_s": round(wall, 4), "refined_pairs_per_s": round(cand / wall, 1), "rows_per_s": round(rows / wall, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "candidates": int(cand), "rows": int(rows)}, [(r.stats["candidates"], r.stats["hits"]) for r in res]

    serial, cs = run(1)
    served, cv = run(args.lanes)
    assert cs == cv, "served and serial request results differ"
    out = {"workload": f"batched requests: {n} requests, sizes lognormal in [1e3, {args.smax:.0e}] polygons (config-1 generator), seed 3",
           "mix": "40% intersects+area, 20% relate, 20% dwithin d~U(1,50) m, 20% eroded intersects ra,rb~U(0.5,3) m",
           "arrivals": "all at t=0", "serial": serial, "served": served,
           "speedup_pairs_per_s": round(served["refined_pairs_per_s"] / serial["refined_pairs_per_s"], 3),
           "p99_ratio_served_over_serial": round(served["p99_ms"] / serial["p99_ms"], 3)}
    for L in layers:
        L.free()
    return out


NATIONAL_NX, NATIONAL_NY = 19737, 8224   # 162.3M cells of 60/19737 degrees over 60 x 25 degrees


def national_leg(gsj, ctx, args):
    """Resident national-scale layers and a windowed request stream (the serving configuration).

    Parcels: the config-2 lattice generator (cells share or overlap their sides, as real parcels do, so that distance
    joins find neighbours; median 10 vertices, at most 256 so no piece needs subdivision, seed 4) with cells of
    60/19737 degrees tiled from (-125, 24): 19737 x 8224 cells = 162.3M parcels at scale 1, both sides scaled by
    sqrt(scale). Points: 161M at scale 1, half uniform over the extent, half in clusters of sigma 50 m, seed 4.
    On the device: upload, LBVH, self-dwithin at 5 m of the whole parcel layer and points-in-polygons of every point
    (rows counted on the host). Then a request stream over the resident parcels with Poisson arrivals: self joins in
    query windows of area lognormal in [1e-4, 10] square degrees, mix 35% intersects + area, 20% relate, 20% dwithin
    d ~ U(1, 50) m, 20% eroded intersects ra, rb ~ U(0.5, 3) m, 5% cluster_intersecting of the window's parcels;
    served (`lanes` lanes, batching) vs serial (one lane, one request at a time) on the same arrival times, at `rate`
    and at `overload_rate`; latency = arrival -> finished."""
    f = args.national
    cell = 60.0 / NATIONAL_NX
    nx, ny = max(1, round(NATIONAL_NX * f ** 0.5)), max(1, round(NATIONAL_NY * f ** 0.5))
    x0, y0, W, H = -125.0, 24.0, nx * cell, ny * cell
    out = {"workload": f"resident parcels {nx}x{ny} cells of {cell:.6g} deg (config-2 generator, median 10, kmax 256, seed 4) + "
                       f"{int(161e6 * f)} clustered points; self-dwithin 5 m, points-in-polygons, windowed request stream"}
    t0 = time.time()
    M = gsj.generate("parcels_mesh", nx=nx, ny=ny, x0=x0, y0=y0, w=cell, h=cell, median=10.0, kmax=256, seed=4)
    t_gen = time.time() - t0
    log(f"national: generated {M.n} parcels, {M.nv} vertices in {t_gen:.1f}s")
    t0 = time.time()
    L = ctx.polys(M, build_bvh=False)
    t_up = time.time() - t0
    n_par, n_v = M.n, M.nv
    L.host = None
    del M
    t0 = time.time()
    L.build_bvh()
    t_bvh = time.time() - t0
    info = L.info()
    log(f"national: upload {t_up:.1f}s, bvh {t_bvh:.1f}s, {info.device_bytes / 2**30:.1f} GiB resident")
    npts = int(161e6 * f)
    t0 = time.time()
    xy = gsj.generate_points(n=npts, x0=x0, y0=y0, w=W, h=H, clusters=max(1, int(32000 * f)), sigma_m=50.0, seed=4)
    P = ctx.points(xy)
    del xy
    t_pts = time.time() - t0
    pinfo = P.info()
    out["resident"] = {"parcels": n_par, "vertices": n_v, "points": npts, "parcel_layer_gib": round(info.device_bytes / 2**30, 2),
                       "point_layer_gib": round(pinfo.device_bytes / 2**30, 2), "generate_s": round(t_gen, 1), "upload_s": round(t_up, 1),
                       "bvh_s": round(t_bvh, 2), "points_generate_upload_s": round(t_pts, 1)}
    log(f"national: points {npts} in {t_pts:.1f}s; {ctx.status()}")
    t0 = time.perf_counter()
    r = ctx.dwithin(L, L, 5.0, same=True, upper=True, collect=False)
    t_dw = time.perf_counter() - t0
    out["self_dwithin_5m"] = {"s": round(t_dw, 3), "candidates": r.stats["candidates"], "rows": r.stats["hits"],
                              "refined_pairs_per_s": round(r.stats["candidates"] / t_dw, 1)}
    log(f"national: self-dwithin 5 m {t_dw:.2f}s {r.stats}")
    t0 = time.perf_counter()
    r = ctx.points_in_polygons(P, L, collect=False)
    t_pip = time.perf_counter() - t0
    out["points_in_polygons"] = {"s": round(t_pip, 3), "candidates": r.stats["candidates"], "rows": r.stats["hits"],
                                 "refined_pairs_per_s": round(r.stats["candidates"] / t_pip, 1)}
    log(f"national: points-in-polygons {t_pip:.2f}s {r.stats}")
    P.free()

    # windowed request stream, Poisson arrivals
    nreq = args.stream_requests
    stream = window_stream(nreq, x0, y0, W, H)
    reqs = [(window_submit_kind(gsj, name), kw) for name, kw in stream]

    def run(lanes, arrivals):
        # serial: one request at a time, as the reference serves them; served: lanes plus batches of windowed joins
        s = gsj.Scheduler(ctx, lanes=lanes, batch=1 if lanes == 1 else args.batch)
        try:
            for k, kw in reqs[: min(nreq, 16)]:   # warm-up: lanes' scratch
                s.wait(s.submit(k, L, L, collect=False, **kw))
            t_start = time.perf_counter()
            tickets, sub = [], []
            for (k, kw), ta in zip(reqs, arrivals):
                dt = t_start + ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                sub.append(time.perf_counter())
                tickets.append(s.submit(k, L, L, collect=False, **kw))
            res = [s.wait(t) for t in tickets]
            wall = time.perf_counter() - t_start
        finally:
            s.close()
        cand = sum(r.stats["candidates"] for r in res)
        rows = sum(r.stats["hits"] for r in res)
        lat = np.array([r.stats["times"]["total_ms"] for r in res])
        nb = np.array([r.stats["times"]["batch"] for r in res])
        return {"lanes": lanes, "wall_s": round(wall, 3), "refined_pairs_per_s": round(cand / wall, 1), "rows_per_s": round(rows / wall, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3), "candidates": int(cand),
                "rows": int(rows), "batched_requests": int((nb > 1).sum()),
                "mean_batch": round(float(nb[nb > 1].mean()), 2) if (nb > 1).any() else 1.0}, [(r.stats["candidates"], r.stats["hits"]) for r in res]

    # the same requests at the configuration's top arrival rate and at an overload rate (both paths saturated: the
    # refined pairs/s are then each path's capacity)
    streams = []
    for rate in (args.rate, args.overload_rate):
        arrivals = np.cumsum(np.random.default_rng(40).exponential(1.0 / rate, nreq))
        arrivals -= arrivals[0]
        serial, cs = run(1, arrivals)
        served, cv = run(args.lanes, arrivals)
        assert cs == cv, "served and serial request results differ"
        streams.append({"requests": nreq, "arrivals": f"Poisson, {rate:g}/s", "window_area_deg2": "lognormal in [1e-4, 10]",
                        "mix": "35% intersects+area, 20% relate, 20% dwithin d~U(1,50) m, 20% eroded intersects ra,rb~U(0.5,3) m (self joins), "
                               "5% cluster_intersecting of the window's parcels",
                        "serial": serial, "served": served,
                        "speedup_pairs_per_s": round(served["refined_pairs_per_s"] / serial["refined_pairs_per_s"], 3),
                        "p99_ratio_served_over_serial": round(served["p99_ms"] / serial["p99_ms"], 3),
                        "reading": "capacity: both paths saturated, refined pairs/s is each path's throughput" if rate > args.rate else
                                   "the configuration's top rate: both paths keep up with the arrivals, so their refined pairs/s follow the "
                                   "arrivals; latency is the difference"})
        log(f"national stream {rate:g}/s: serial {serial}; served {served}")
    out["stream"], out["stream_overload"] = streams
    L.free()
    if not args.no_wire:
        ctx.trim()
        spec = (f"gen:parcels_mesh:nx={nx};ny={ny};x0={x0!r};y0={y0!r};w={cell!r};h={cell!r};median=10;kmax=256;seed=4")
        out["wire"] = wire_streams(args, spec, stream, (args.rate, args.overload_rate))
    return out


def wire_line(name, kw, layer="n"):
    """one request of the windowed stream as gpujoind arguments (self joins list each unordered pair once)"""
    from gsj import client as G

    w = "window=" + ",".join(repr(float(v)) for v in kw["window"])
    if name in ("area", "relate"):
        return (G.OP_JOIN, [layer, layer, "1", "distinct", w]) if name == "area" else (G.OP_RELATE, [layer, layer, "", "distinct", w])
    if name == "dwithin":
        return G.OP_DWITHIN, [layer, layer, "dwithin", repr(kw["d"]), "0", "distinct", w]
    if name == "erode":
        return G.OP_DWITHIN, [layer, layer, "buffer", repr(-kw["ra"]), repr(-kw["rb"]), "distinct", w]
    return G.OP_CLUSTER, [layer, w]


def wire_streams(args, spec, stream, rates):
    """The windowed stream through the daemon over its unix socket, as the extension's clients would send it: gpujoind
    with one lane and no batching (serial) and with `lanes` lanes and batching (served), each holding the national
    parcel layer (LOAD of the generator spec); gsjload sends every request on its own connection at its Poisson arrival
    time and measures its latency at the client, from the last request byte to the end-of-stream stats frame, with
    every row drained. Refined pairs/s = the daemon's reported candidates over the client's wall time."""
    import subprocess
    import tempfile
    from gsj import client as G

    daemon = os.path.join(ROOT, "postgis-cuda_amd", "bin", "gpujoind")
    load = os.path.join(ROOT, "postgis-cuda_amd", "bin", "gsjload")
    out = {"measured": "client side over the daemon's unix socket: last request byte -> end-of-stream stats frame, all rows drained",
           "self_pairs": "distinct (each unordered pair once)"}
    lines = [wire_line(name, kw) for name, kw in stream]
    for mode, lanes, batch in (("serial", 1, 1), ("served", args.lanes, args.batch)):
        d = tempfile.mkdtemp()
        sock = os.path.join(d, "gpujoind.sock")
        with open(os.path.join(d, "daemon.log"), "w") as logf:
            proc = subprocess.Popen([daemon, "--socket", sock, "--data-dir", d, "--lanes", str(lanes), "--batch", str(batch), "--threads", "8"],
                                    stdout=logf, stderr=logf)
            try:
                for _ in range(600):
                    if os.path.exists(sock) or proc.poll() is not None:
                        break
                    time.sleep(0.1)
                assert proc.poll() is None, open(os.path.join(d, "daemon.log")).read()
                t0 = time.time()
                info = G.Client(sock).call(G.OP_LOAD, "n", spec, "id", "geom", "", 0, "false")
                log(f"wire {mode}: layer loaded in {time.time() - t0:.1f}s ({info['rows']} rows)")
                for rate in rates:
                    arrivals = np.cumsum(np.random.default_rng(40).exponential(1.0 / rate, len(lines)))
                    arrivals -= arrivals[0]
                    text = "\n".join("\t".join([f"{t:.6f}", str(op)] + a) for t, (op, a) in zip(arrivals, lines)) + "\n"
                    r = subprocess.run([load, "--socket", sock], input=text, capture_output=True, text=True, timeout=900)
                    recs = [json.loads(x) for x in r.stdout.splitlines() if x.strip()]
                    summ = recs[-1]
                    cand = sum(int(x["stats"].get("candidates", 0)) for x in recs[:-1] if x.get("stats"))
                    # the daemon's own account, summed over requests: its time per request (receipt to stats frame) and
                    # the part of it spent encoding and sending rows
                    dsum = lambda k: round(sum(float(x["stats"].get(k, 0.0)) for x in recs[:-1] if x.get("stats")), 3)
                    res = {"lanes": lanes, "batch": batch, "wall_s": summ["wall_s"], "refined_pairs_per_s": round(cand / summ["wall_s"], 1),
                           "rows_per_s": round(summ["rows"] / summ["wall_s"], 1), "p50_ms": summ["p50_ms"], "p99_ms": summ["p99_ms"],
                           "candidates": cand, "rows": summ["rows"], "failed": summ["failed"],
                           "daemon_s": {"total": dsum("total_s"), "encode_send": dsum("host_encode_send_s")}}
                    out.setdefault(f"{rate:g}/s", {})[mode] = res
                    log(f"wire {mode} {rate:g}/s: {res}")
            finally:
                proc.terminate()
                proc.wait(timeout=120)
    for key, v in out.items():
        if isinstance(v, dict) and "serial" in v and "served" in v:
            v["speedup_pairs_per_s"] = round(v["served"]["refined_pairs_per_s"] / max(v["serial"]["refined_pairs_per_s"], 1e-9), 3)
            v["p99_ratio_served_over_serial"] = round(v["served"]["p99_ms"] / max(v["serial"]["p99_ms"], 1e-9), 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=1549, help="lattice cells per side (1549 = the 2.4M-parcel config)")
    ap.add_argument("--sample", type=int, default=400, help="cells per side of the CPU/parity sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline and parity legs (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="host work of the cpu_baseline leg (repeats of the join)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-vs-serial leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the parcels x regions and points / dwithin legs")
    ap.add_argument("--no-wire", action="store_true", help="skip the national stream through the daemon's socket")
    ap.add_argument("--no-prep", action="store_true", help="skip the host-vs-device layer preparation timings (profiling runs)")
    ap.add_argument("--requests", type=int, default=256, help="requests in the served-vs-serial mix")
    ap.add_argument("--smax", type=float, default=2e6, help="largest request (polygons) in the mix")
    ap.add_argument("--lanes", type=int, default=4, help="lanes of the served path")
    ap.add_argument("--batch", type=int, default=16, help="most windowed requests a served lane runs in one pass (national stream)")
    ap.add_argument("--calibrate-copy", action="store_true", help="also run 1 GiB device copies (HBM counter calibration)")
    ap.add_argument("--national", type=float, default=0.0,
                    help="run only the resident national-scale leg at this scale (1.0 = 162M parcels + 161M points)")
    ap.add_argument("--national-scale", type=float, default=1.0, help="scale of the national leg that the default run appends")
    ap.add_argument("--no-national", action="store_true", help="skip the national leg of the default run")
    ap.add_argument("--stream-requests", type=int, default=1000, help="requests in the national windowed stream")
    ap.add_argument("--rate", type=float, default=500.0, help="Poisson arrival rate (requests/s) of the national stream")
    ap.add_argument("--overload-rate", type=float, default=5000.0, help="second arrival rate of the national stream, above both paths' capacity")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU design is described, not built)",
              file=sys.stderr)
        sys.exit(2)

    import gsj

    if args.national > 0:
        ctx = gsj.Context(0)
        t0 = time.perf_counter()
        nat = national_leg(gsj, ctx, args)
        st = nat["stream_overload"]["served"]   # the served path's capacity on the windowed stream
        print(json.dumps({"metric": METRIC, "value": st["refined_pairs_per_s"], "unit": "refined pairs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
                          "ms_per_step": round(1e3 * (time.perf_counter() - t0), 1), "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic (deterministic generators)",
                          "config": {"workload": nat["workload"], "parallelism": "1 GPU"}, "national": nat}), flush=True)
        ctx.close()
        return

    t0 = time.time()
    M, S = mesh(gsj, args.scale)
    log(f"generated {M.n} parcels ({M.nv} vertices), {S.n} pieces after subdivide 256 in {time.time() - t0:.1f}s")
    ctx = gsj.Context(0)
    if args.calibrate_copy:
        copy_calibration()
    # layer preparation of the same 2.4M parcels: host restatement vs device (identical results, checked)
    prep = {}
    for name, fn in () if args.no_prep else (("subdivide_host", lambda: gsj.subdivide(M, 256)), ("subdivide_device", lambda: gsj.subdivide(M, 256, ctx=ctx)),
                     ("dedup_host", lambda: gsj.dedup(M)), ("dedup_device", lambda: gsj.dedup(M, ctx=ctx))):
        fn()
        t0 = time.perf_counter()
        r = fn()
        prep[name] = round((time.perf_counter() - t0) * 1e3, 1)
        if name == "subdivide_device":
            assert np.array_equal(r.xy.view(np.uint64), S.xy.view(np.uint64)) and np.array_equal(r.syn, S.syn), "device subdivision differs"
        del r   # a prep result is 2.4M parcels of host arrays: freed here, not when a later `r` is bound inside a timed step
    log(f"prep (ms, host arrays in and out): {prep}")
    parse = None if args.no_prep else parse_leg(gsj, ctx, M, args)
    t0 = time.time()
    L = ctx.polys(S, build_bvh=False)
    t_up = time.time() - t0
    t0 = time.time()
    L.build_bvh()
    t_bvh = time.time() - t0
    bvh_dev_ms = L.info().bvh_ms   # device time of the build (the layer is freed before the national leg)
    log(f"upload {t_up * 1e3:.0f} ms, bvh build {t_bvh * 1e3:.0f} ms ({bvh_dev_ms:.2f} ms of device time); {ctx.status()}")

    opts = dict(want_area=False, want_relate=True, upper=True)
    for _ in range(args.warmup):
        ctx.join(L, L, collect=False, **opts)
    times, stats = [], []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        r = ctx.join(L, L, collect=False, **opts)
        times.append(time.perf_counter() - t1)
        stats.append(r.stats)
    t_all = time.perf_counter() - t_all
    st = stats[-1]
    cand, rows = st["candidates"], st["hits"]
    ms_step = 1e3 * t_all / args.steps
    value = cand * args.steps / t_all
    log(f"step {ms_step:.2f} ms: {cand} candidates, {rows} rows; traverse {st['traverse_ms']:.2f} refine {st['refine_ms']:.2f} "
        f"compact {st['compact_ms']:.2f} copy {st['copy_ms']:.2f} ms; kernels screen {st['k_screen_ms']:.2f} tiny {st['k_tiny_ms']:.2f} "
        f"small {st['k_small_ms']:.2f} medium {st['k_medium_ms']:.2f} mixed {st['k_mixed_ms']:.2f} large {st['k_large_ms']:.2f} "
        f"over {st['k_over_ms']:.2f} ms (pairs {st['n_tiny']}/{st['n_small']}/{st['n_medium']}/{st['n_mixed']}/{st['n_large']}/{st['n_over']}, "
        f"edge work {st['n_tiny_edge']}/{st['n_small_edge']}/{st['n_medium_edge']}); steps (ms) {[round(t * 1e3, 2) for t in times]}")

    # dominant refinement kernel and its algorithmic FLOPs (counting oracle on a sample of the same generator):
    # (name, duration field, pair class of the sample, pairs field, FLOPs counted: edge phase / screen part / all)
    kernels = (("k_screen_small", "k_screen_ms", (4, 0, 1), ("n_tiny", "n_small", "n_medium"), "screen"),
               ("k_edges_small<8, 16, 0>", "k_tiny_ms", (4,), ("n_tiny_edge",), "edge"),
               ("k_edges_small<16, 32, 1>", "k_small_ms", (0,), ("n_small_edge",), "edge"),
               ("k_edges_small<32, 64, 2>", "k_medium_ms", (1,), ("n_medium_edge",), "edge"),
               ("k_refine_mixed", "k_mixed_ms", (2,), ("n_mixed",), "all"),
               ("k_refine_large", "k_large_ms", (3,), ("n_large",), "all"))
    kern = max(kernels, key=lambda k: st[k[1]])
    roofline = None
    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as O

        Ms, Ss = mesh(gsj, args.sample)
        Os = O.Layer.polys(Ss)
        cand_s = O.candidates(Os, Os, upper=True)
        pick = cand_s[np.isin(pair_class(Ss, cand_s), kern[2])]
        rng = np.random.default_rng(0)
        pick = pick[rng.choice(len(pick), size=min(len(pick), 40000), replace=False)]
        fl = O.refine_flops(Ss, Ss, pick, want_area=False, want_relate=True)
        if kern[4] == "edge":     # per pair that reaches the edge contributions
            flops_per_pair = fl["edge_flops"] / max(fl["edge_pairs"], 1)
        elif kern[4] == "screen":   # contact + containment: everything but the edge contributions
            flops_per_pair = (fl["flops"] - fl["edge_flops"]) / max(fl["pairs"], 1)
        else:
            flops_per_pair = fl["flops"] / max(fl["pairs"], 1)
        per_launch_pairs = sum(st[f] for f in kern[3]) / max(st["launches"], 1)
        per_launch_ms = st[kern[1]] / max(st["launches"], 1)
        achieved = flops_per_pair * per_launch_pairs / (per_launch_ms * 1e-3) / 1e12
        roofline = {"kernel": kern[0], "bound": "valu", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP64_PEAK_TFLOPS, 6), "traffic": hbm_traffic(kern[0]), "flops_per_pair": round(flops_per_pair, 1),
                    "pairs_per_launch": int(per_launch_pairs), "launch_ms": round(per_launch_ms, 4),
                    "flops_counted": {"edge": "edge contributions of the pairs that need them", "screen": "contact and containment",
                                      "all": "the whole refinement"}[kern[4]],
                    "flop_mix": {k: (fl["edge"] if kern[4] == "edge" else fl)[k] for k in ("add", "mul", "fma", "div", "sqrt")}}
        issued = issued_f64(kern[0])   # lanes the FP64 instructions occupied, masked or not (counter summary)
        if issued:
            roofline["issued_f64_lane_flops"] = issued
            roofline["lane_efficiency"] = round(flops_per_pair * per_launch_pairs / issued, 4)
        # cpu baseline: the oracle's relate self-join of the same 2.4M-parcel layer on the host's share of cores,
        # repeated until about --cpu-seconds of work
        threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))
        O.set_threads(threads)
        Of = O.Layer.polys(S)
        reps, t_cpu, n_cpu = 0, 0.0, 0
        while reps == 0 or t_cpu < args.cpu_seconds:
            t1 = time.perf_counter()
            r = O.join(Of, Of, want_area=False, want_relate=True, upper=True)
            t_cpu += time.perf_counter() - t1
            n_cpu += r["candidates"]
            reps += 1
            del r
        del Of
        cpu = {"value": round(n_cpu / t_cpu, 1), "unit": "refined pairs/s", "cores": threads, "kind": "port",
               "sample": f"the same {S.n}-piece relate self-join ({n_cpu // reps} candidates), run {reps} times in {t_cpu:.1f} s"}
        exp = O.join(Os, Os, want_area=False, want_relate=True, upper=True)
        # parity of the engine on the same sample
        Ls = ctx.polys(Ss)
        got = ctx.join(Ls, Ls, **opts)
        ge = {tuple(p): c for p, c in zip(got.pairs.tolist(), got.code.tolist())}
        ee = {tuple(p): c for p, c in zip(exp["pairs"].tolist(), exp["code"].tolist())}
        parity = {"rows_gpu": len(ge), "rows_oracle": len(ee), "missing": len(ee.keys() - ge.keys()), "extra": len(ge.keys() - ee.keys()),
                  "de9im_mismatch": sum(1 for k in ee.keys() & ge.keys() if ee[k] != ge[k]),
                  "candidates_equal": got.stats["candidates"] == exp["candidates"]}
        log(f"cpu baseline {cpu['value']:.0f} pairs/s on {threads} threads; parity {parity}; roofline {roofline}")

    configs = {}
    if not args.no_configs:   # the parcels x regions and the points / distance configurations, each with its own roofline
        configs["config1"] = config1_leg(gsj, ctx, args)
        log(f"config 1: {configs['config1']}")
        configs["config3"] = config3_leg(gsj, ctx, args, L, S)
        log(f"config 3: {configs['config3']}")
    serving = None if args.no_serve else serving_leg(gsj, ctx, args)
    if serving:
        log(f"serving: serial {serving['serial']}; served {serving['served']}; speedup {serving['speedup_pairs_per_s']}")
    national = None
    if not (args.no_serve or args.no_national):   # the serving configuration: resident national layers, windowed stream
        L.free()
        ctx.trim()
        args.national = args.national_scale
        national = national_leg(gsj, ctx, args)
    lat = np.array(times) * 1e3
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "refined pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (deterministic config-2 generator)",
        "config": {"workload": f"config 2: {M.n}-parcel self-join, full DE-9IM relate, subdivide 256, seed 1", "parcels": M.n,
                   "pieces": S.n, "vertices": S.nv, "candidates_per_step": cand, "rows_per_step": rows, "parallelism": "1 GPU"},
        "rows_per_s": round(rows * args.steps / t_all, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3), "path": "serial"},
        "breakdown_ms": {k: round(st[k], 3) for k in ("traverse_ms", "refine_ms", "compact_ms", "copy_ms", "k_screen_ms", "k_tiny_ms", "k_small_ms",
                                                      "k_medium_ms", "k_mixed_ms", "k_large_ms", "k_over_ms")},
        "load_ms": {"upload": round(t_up * 1e3, 1), "bvh_build": round(t_bvh * 1e3, 1), "bvh_build_device": round(bvh_dev_ms, 3), "prep": prep},
        "layer_parse": parse,
        "roofline_other": other_rooflines(S.n, st, bvh_dev_ms),
        "roofline": roofline, "cpu_baseline": cpu, "sample_parity": parity, **configs, "serving": serving, "national": national,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
