"""Launch-shape tuner: measures the pair kernels' launch shapes on this device, per size bucket of a launch, and writes
the table that contexts read at creation (postgis-cuda_amd/tuning/<arch>.txt; format in include/gsj.h,
gsj_ctx_set_tuning).

    python -m gsj.tune [--out FILE] [--side N] [--reps R]      (with postgis-cuda_amd on sys.path)

Workloads, all on one parcel layer of the mesh generator (N x N cells of 1/1549 degree, lognormal vertex counts,
seed 1, subdivided at 256 vertices; N = 1549 is the 2.4M-parcel relate configuration):
  large  (launches of >= 2^21 pairs): the whole relate self-join (about 10M candidate pairs) and a self dwithin of 5 m;
  medium (>= 2^15 pairs): relate and dwithin self-joins in 8 windows of 0.1-0.3 degree (about 1e5-1e6 pairs each);
  small  (< 2^15 pairs): the same in 64 windows of 0.005-0.03 degree (about 1e2-1e4 pairs each).
Each bucket starts from the built-in shape and takes one key at a time (coordinate descent, two passes): every value
of the key is timed (best of R runs of the bucket's workload, wall clock around the synchronous calls, rows handed
to a callback that keeps none) and kept when it is more than 2% faster than the value in place. Every shape tried
must give rows identical to the built-in shape's (a digest of the sorted rows of every request, from one more run that
copies them out) or the tuner stops with an error.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

BUCKETS = (("small", 0), ("medium", 1 << 15), ("large", 1 << 21))
# the values tried per key; the first is the built-in shape's
KEYS = (("tcap", (16, 0, 8, 12)), ("mcap", (64, 0, 48)), ("screen_occ", (5, 4)), ("screen_bpc", (16, 4, 8, 32, 64)),
        ("edge_bpc", (16, 4, 8, 32, 64)), ("mixed_bpc", (32, 8, 16, 64)), ("large_bpc", (16, 4, 8, 32)), ("trav_block", (256, 64, 128)),
        ("dw_bpc", (16, 4, 8, 32, 64)), ("ovf_lds", (1, 0, 2)), ("fork", (1, 0)), ("mixed", (1, 0)),
        ("large_cap", (96, 0, 64)), ("size_sort", (1, 0, 2, 3)), ("dw_size_sort", (0, 1, 3)))
DEFAULT = {k: v[0] for k, v in KEYS}


def table_text(shapes, header=()):
    lines = [f"# {h}" for h in header]
    for (name, np_min) in BUCKETS:
        s = shapes[name]
        lines.append(f"bucket np>={np_min} " + " ".join(f"{k}={s[k]}" for k, _ in KEYS) + f"   # {name}")
    return "\n".join(lines) + "\n"


def digest(rows):
    """order-free digest of a request's rows (pairs, areas, codes, distances as present)"""
    p = rows.pairs
    order = np.lexsort((p[:, 1], p[:, 0])) if len(p) else np.zeros(0, np.int64)
    h = hashlib.sha256(np.ascontiguousarray(p[order]).tobytes())
    for a in (rows.area, rows.code, rows.dist):
        if a is not None and len(a):
            h.update(np.ascontiguousarray(a[order]).tobytes())
    return h.hexdigest()


def windows(rng, n, lo, hi, x0, y0, ext):
    side = rng.uniform(lo, hi, n)
    cx = x0 + side / 2 + rng.uniform(0, 1, n) * (ext - side)
    cy = y0 + side / 2 + rng.uniform(0, 1, n) * (ext - side)
    return [(float(a - s / 2), float(b - s / 2), float(a + s / 2), float(b + s / 2)) for a, b, s in zip(cx, cy, side)]


def workload(ctx, L, bucket, x0, y0, ext, rng_seed=7):
    """the bucket's requests as callables f(collect) returning rows"""
    rel = dict(want_relate=True, upper=True)
    if bucket == "large":
        return [lambda c: ctx.join(L, L, collect=c, **rel), lambda c: ctx.dwithin(L, L, 5.0, same=True, upper=True, collect=c)]
    rng = np.random.default_rng(rng_seed)
    lo, hi, n = (0.1, 0.3, 8) if bucket == "medium" else (0.005, 0.03, 64)
    ws = windows(rng, n, lo * ext, hi * ext, x0, y0, ext)
    out = []
    for w in ws:
        out.append(lambda c, w=w: ctx.join(L, L, window=w, collect=c, **rel))
        out.append(lambda c, w=w: ctx.dwithin(L, L, 5.0, same=True, upper=True, window=w, collect=c))
    return out


def measure(ctx, reqs, reps):
    """(best wall seconds over reps with the rows left on the callback, digests of a run that copies them out)"""
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        for f in reqs:
            f(False)
        best = min(best, time.perf_counter() - t0)
    return best, [digest(f(True)) for f in reqs]


def shape_text(shapes, bucket, trial):
    s = dict(shapes)
    s[bucket] = trial
    return table_text(s)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=None, help="table file to write (default: tuning/<arch>.txt of the package)")
    ap.add_argument("--side", type=int, default=1549, help="mesh side in cells (1549: the 2.4M-parcel layer)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--buckets", default="small,medium,large")
    args = ap.parse_args()
    import gsj

    def log(msg):
        print(f"[tune] {msg}", file=sys.stderr, flush=True)

    ctx = gsj.Context(0)
    status = json.loads(ctx.status())
    arch = status.get("arch", "gfx950").split(":")[0]
    out = args.out or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuning", f"{arch}.txt")
    t0 = time.time()
    M = gsj.generate("parcels_mesh", nx=args.side, ny=args.side, w=1 / 1549, h=1 / 1549, seed=1)
    S = gsj.subdivide(M, 256)
    L = ctx.polys(S)
    x0, y0 = float(S.xy[:, 0].min()), float(S.xy[:, 1].min())   # the mesh's extent: windows lie inside it
    ext = min(float(S.xy[:, 0].max()) - x0, float(S.xy[:, 1].max()) - y0)
    log(f"layer: {M.n} parcels, {S.n} pieces, resident in {time.time() - t0:.1f}s; device {status.get('device')} ({arch})")
    shapes = {name: dict(DEFAULT) for name, _ in BUCKETS}
    report = {}
    for name, _ in BUCKETS:
        if name not in args.buckets.split(","):
            continue
        reqs = workload(ctx, L, name, x0, y0, ext)
        ctx.set_tuning(table_text(shapes))
        measure(ctx, reqs, 1)   # warm-up: scratch grows to its working size
        base_t, base_dg = measure(ctx, reqs, args.reps)
        cur, cur_t = dict(DEFAULT), base_t
        tried, trials = 0, []
        log(f"{name}: built-in shape {base_t * 1e3:.2f} ms for {len(reqs)} requests")
        for _ in range(args.passes):
            for key, values in KEYS:
                for v in values:
                    if v == cur[key]:
                        continue
                    trial = dict(cur)
                    trial[key] = v
                    ctx.set_tuning(shape_text(shapes, name, trial))
                    t, dg = measure(ctx, reqs, args.reps)
                    tried += 1
                    trials.append({"shape": {key: v}, "ms": round(t * 1e3, 3)})
                    if dg != base_dg:
                        bad = [i for i, (a, b) in enumerate(zip(dg, base_dg)) if a != b]
                        raise SystemExit(f"tune: rows differ under {name} {trial} (requests {bad[:8]})")
                    if t < 0.98 * cur_t:
                        log(f"{name}: {key}={v} {t * 1e3:.2f} ms (was {cur_t * 1e3:.2f})")
                        cur, cur_t = trial, t
                log(f"{name}: {key} -> {cur[key]} ({tried} shapes so far, {cur_t * 1e3:.2f} ms; tried "
                    + ", ".join(f"{list(x['shape'].values())[0]}: {x['ms']}" for x in trials if key in x["shape"]) + ")")
        shapes[name] = cur
        report[name] = {"requests": len(reqs), "built_in_ms": round(base_t * 1e3, 3), "tuned_ms": round(cur_t * 1e3, 3), "shapes_tried": tried,
                        "shape": cur, "trials": trials}
        log(f"{name}: tuned {cur_t * 1e3:.2f} ms vs built-in {base_t * 1e3:.2f} ms ({tried} shapes, rows identical under all)")
    dev = status.get("device") or arch
    header = [f"launch shapes for {arch} ({dev}, {status.get('cu')} CUs), written by gsj/tune.py on "
              f"{time.strftime('%Y-%m-%d')}",
              f"workloads: mesh {args.side}x{args.side} cells, seed 1, subdivide 256 ({S.n} pieces); best of {args.reps} runs"]
    for name, r in report.items():
        header.append(f"{name}: {r['requests']} requests, built-in {r['built_in_ms']} ms -> tuned {r['tuned_ms']} ms "
                      f"({r['shapes_tried']} shapes tried, rows identical)")
    text = table_text(shapes, header)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        f.write(text)
    print(json.dumps({"table": out, "report": report}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
