"""gpujoind end to end on the GPU: the daemon binary (on libgsj's C ABI and scheduler) speaking the reference's wire
protocol. Every response is checked against the library's direct calls on the same distinct, subdivided layer, with
the reference's row semantics restated here: dedup member expansion, self-join mirroring ("all") or ordering
("distinct"), transposed DE-9IM for mirrored rows, point/polygon swap, dense cluster labels."""
import json
import math
import os
import subprocess
import tempfile
import threading
import time

import numpy as np
import pytest

import gsj
from gsj import client as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DAEMON = os.path.join(ROOT, "postgis-cuda_amd", "bin", "gpujoind")
TOKEN = "s3cret"
CELL = 0.6 / 1095


def _dup_set(M, seed):
    """M with some geometries repeated (identical rows) in shuffled order, and text ids"""
    idx = np.concatenate([np.arange(M.n), np.arange(0, M.n, 7), np.arange(3, M.n, 13)])
    np.random.default_rng(seed).shuffle(idx)
    gro, rvo, xy = [0], [0], []
    for g in idx:
        for r in range(int(M.gro[g]), int(M.gro[g + 1])):
            xy.append(M.xy[int(M.rvo[r]):int(M.rvo[r + 1])])
            rvo.append(rvo[-1] + len(xy[-1]))
        gro.append(len(rvo) - 1)
    return gsj.HostSet(np.array(gro), np.array(rvo), np.concatenate(xy))


@pytest.fixture(scope="module")
def daemon():
    d = tempfile.mkdtemp()
    A = _dup_set(gsj.generate("parcels_mesh", nx=30, ny=30, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=21), 1)
    B = _dup_set(gsj.generate("parcels_mesh", nx=24, ny=24, x0=-95.2, y0=38.6, w=CELL, h=CELL, seed=22), 2)
    pts = gsj.generate_points(n=5000, x0=-95.2, y0=38.6, w=30 * CELL, h=30 * CELL, clusters=5, sigma_m=40.0, seed=3)
    for name, hs in (("a", A), ("b", B)):
        gsj.write_gset(os.path.join(d, name + ".gset"), hs)
        with open(os.path.join(d, name + ".gset.ids"), "w") as f:
            f.write("".join(f"{name}{i}\n" for i in range(hs.n)))
    gsj.write_pset(os.path.join(d, "p.pset"), pts)
    with open(os.path.join(d, "p.pset.ids"), "w") as f:
        f.write("".join(f"pt{i}\n" for i in range(len(pts))))
    with open(os.path.join(d, "token"), "w") as f:
        f.write(TOKEN + "\n")
    sock = os.path.join(d, "gpujoind.sock")
    log = open(os.path.join(d, "daemon.log"), "w")
    proc = subprocess.Popen([DAEMON, "--socket", sock, "--data-dir", d, "--token-file", os.path.join(d, "token"), "--lanes", "3", "--threads", "4"],
                            stdout=log, stderr=log)
    for _ in range(600):
        if os.path.exists(sock) or proc.poll() is not None:
            break
        time.sleep(0.1)
    assert proc.poll() is None, open(os.path.join(d, "daemon.log")).read()
    c = G.Client(sock, token=TOKEN)
    info = {"a": c.load("a", "gset:a.gset", subdivide=8), "b": c.load("b", "gset:b.gset", subdivide=0), "p": c.load("p", "pset:p.pset")}
    yield {"client": c, "dir": d, "sets": {"a": A, "b": B}, "pts": pts, "info": info, "sock": sock}
    proc.terminate()
    proc.wait(timeout=60)
    log.close()


class Side:
    """the daemon's view of a loaded polygon layer, rebuilt through the library: distinct set, members, pieces"""

    def __init__(self, ctx, hs, name, subdiv):
        U, moff, midx = gsj.dedup(hs)
        self.moff, self.midx = moff, midx
        self.ids = [f"{name}{i}" for i in range(hs.n)]
        self.layer = ctx.polys(gsj.subdivide(U, subdiv) if subdiv else U)

    def members(self, g):
        return [int(m) for m in self.midx[int(self.moff[g]):int(self.moff[g + 1])]]


def expand(A, B, pairs, vals, same, distinct, transpose=None):
    """reference row semantics (daemon/gpujoind.cu:231-283, 365-379)"""
    out = []
    for (pa, pb), v in zip(pairs.tolist(), vals):
        diag = same and pa == pb
        for ra in A.members(pa):
            for rb in B.members(pb):
                if not same:
                    out.append((A.ids[ra], B.ids[rb], v))
                elif distinct:
                    if ra == rb or (diag and ra > rb):
                        continue
                    if ra < rb:
                        out.append((A.ids[ra], A.ids[rb], v))
                    else:
                        out.append((A.ids[rb], A.ids[ra], v if diag or transpose is None else transpose(v)))
                else:
                    out.append((A.ids[ra], A.ids[rb], v))
                    if not diag:
                        out.append((A.ids[rb], A.ids[ra], v if transpose is None else transpose(v)))
    return out


def _key(rows):
    return sorted(rows, key=lambda r: (r[0], r[1], str(r[2])))


def _same_rows(got, exp):
    g, e = _key(got), _key(exp)
    assert len(g) == len(e)
    for x, y in zip(g, e):
        assert x[:2] == y[:2]
        if isinstance(y[2], float) and math.isnan(y[2]):
            assert math.isnan(x[2])
        else:
            assert x[2] == y[2]


def test_load_list_status_and_auth(daemon):
    c = daemon["client"]
    assert daemon["info"]["a"]["dedup"] and daemon["info"]["a"]["distinct"] < daemon["info"]["a"]["rows"] == daemon["sets"]["a"].n
    assert daemon["info"]["p"]["kind"] == "point"
    names = {L["layer"] for L in c.call(G.OP_LIST)}
    assert names == {"a", "b", "p"}
    st = c.call(G.OP_STATUS)
    assert set(st["layers"]) == names and st["mem_total_mb"] > 0
    with pytest.raises(G.DaemonError, match="authentication"):
        G.Client(daemon["sock"], token="wrong").call(G.OP_STATUS)
    with pytest.raises(G.DaemonError, match="no such layer"):
        c.join("a", "zz")
    with pytest.raises(G.DaemonError, match="self_pairs"):
        c.join("a", "a", self_pairs="some")


@pytest.mark.parametrize("self_pairs", ["all", "distinct"])
def test_join_self_and_two_layers(ctx, daemon, self_pairs):
    c = daemon["client"]
    A = Side(ctx, daemon["sets"]["a"], "a", 8)
    B = Side(ctx, daemon["sets"]["b"], "b", 0)
    r = ctx.join(A.layer, A.layer, want_area=True, upper=True)
    rows, st, _ = c.join("a", "a", want_area=True, self_pairs=self_pairs)
    _same_rows(rows, expand(A, A, r.pairs, r.area.tolist(), True, self_pairs == "distinct"))
    _, st2, _ = c.join("a", "a", want_area=True, self_pairs=self_pairs, stats=True)
    # stats JSON numbers carry 6 significant digits, as the reference prints them
    assert st2["rows"] == st["rows"] == len(rows) and abs(st2["area_sum"] - sum(x[2] for x in rows)) <= 1e-5 * abs(st2["area_sum"])
    r = ctx.join(A.layer, B.layer)
    rows, _, _ = c.join("a", "b", self_pairs=self_pairs)
    _same_rows(rows, expand(A, B, r.pairs, [math.nan] * len(r.pairs), False, False))


def test_relate_with_pattern_and_transposes(ctx, daemon):
    c = daemon["client"]
    A = Side(ctx, daemon["sets"]["a"], "a", 8)
    r = ctx.join(A.layer, A.layer, want_relate=True, upper=True)
    codes = [gsj.de9im_str(x) for x in r.code]
    tr = {s: gsj.de9im_str(gsj.de9im_transpose(x)) for s, x in zip(codes, r.code)}
    for sp in ("all", "distinct"):
        rows, _, _ = c.relate("a", "a", "", sp)
        _same_rows(rows, expand(A, A, r.pairs, codes, True, sp == "distinct", transpose=lambda s: tr[s]))
    rows, _, _ = c.relate("a", "a", "T********", "all")
    assert rows and all(x[2][0] != "F" for x in rows)


def test_points_and_cluster(ctx, daemon):
    c = daemon["client"]
    B = Side(ctx, daemon["sets"]["b"], "b", 0)
    pts = ctx.points(daemon["pts"])
    pip = ctx.points_in_polygons(pts, B.layer)
    exp = [(f"pt{p}", B.ids[m], math.nan) for (p, g) in pip.pairs.tolist() for m in B.members(g)]
    rows, st, _ = c.join("p", "b")
    _same_rows(rows, exp)
    rows, _, _ = c.join("b", "p")   # polygon x point runs as point x polygon, rows in request order
    _same_rows(rows, [(b, a, v) for a, b, v in exp])
    rel, _, _ = c.relate("p", "b", "", "all")
    assert {(a, b) for a, b, _ in rel} == {(a, b) for a, b, _ in exp} and {m for _, _, m in rel} <= {"0FFFFF212", "F0FFFF212"}
    root, _ = ctx.cluster(B.layer)
    rows, st, _ = c.cluster("b")
    dense, lab = {}, {}
    for g in range(len(B.moff) - 1):
        for m in B.members(g):
            lab[m] = int(root[g])
    exp = []
    for rr in range(len(B.ids)):
        dense.setdefault(lab[rr], len(dense))
        exp.append((B.ids[rr], dense[lab[rr]]))
    assert rows == exp and st["clusters"] == len(set(dense.values()))


def test_windowed_cluster(ctx, daemon):
    """CLUSTER with the trailing window: the source rows of the geometries meeting the window, numbered densely in row
    order from the engine's windowed components"""
    c = daemon["client"]
    B = Side(ctx, daemon["sets"]["b"], "b", 0)
    bb = B.layer.export()["bb"].astype(np.float64)
    x0, y0, x1, y1 = bb[:, 0].min(), bb[:, 1].min(), bb[:, 2].max(), bb[:, 3].max()
    win = (x0 + 0.3 * (x1 - x0), y0 + 0.2 * (y1 - y0), x0 + 0.7 * (x1 - x0), y0 + 0.6 * (y1 - y0))
    lab = {}
    for g, rt in ctx.cluster_rows(B.layer, window=win).pairs.tolist():
        for m in B.members(g):
            lab[m] = rt
    dense, exp = {}, []
    for rr in sorted(lab):
        dense.setdefault(lab[rr], len(dense))
        exp.append((B.ids[rr], dense[lab[rr]]))
    rows, st, _ = c.cluster("b", window=win)
    assert rows == exp and 0 < len(rows) < len(B.ids) and st["clusters"] == len(dense)


GSP_CAT = os.path.join(ROOT, "oracle", "_ref", "gsp_cat")


@pytest.mark.skipif(not os.path.exists(GSP_CAT), reason="the reference's client is built only where the reference tree is present")
def test_reference_client_reads_the_streams(daemon):
    """tools/gsp_cat.c of the reference, compiled unchanged from its source, drains this daemon's JOIN / CLUSTER /
    DWITHIN streams (frames, u16-length ids, f64 / u32 slots, end marker, stats JSON): its row count equals the rows
    this repository's client parses from the same request, and its errors come back as the reference's error reply."""
    c = daemon["client"]
    env = dict(os.environ, GSP_TOKEN=TOKEN)
    bb = daemon["sets"]["a"].xy
    w = (bb[:, 0].min() + 0.002, bb[:, 1].min() + 0.002, bb[:, 0].min() + 0.009, bb[:, 1].min() + 0.008)
    wa = "window=" + ",".join(repr(float(v)) for v in w)
    cases = [
        (G.OP_JOIN, ["a", "a", "1", "all"], lambda: c.join("a", "a", want_area=True)),
        (G.OP_JOIN, ["a", "b", "0", "distinct"], lambda: c.join("a", "b", self_pairs="distinct")),
        (G.OP_JOIN, ["p", "b", "0", "all"], lambda: c.join("p", "b")),
        (G.OP_JOIN, ["a", "a", "1", "all", wa], lambda: c.join("a", "a", want_area=True, window=w)),
        (G.OP_CLUSTER, ["b"], lambda: c.cluster("b")),
        (G.OP_CLUSTER, ["b", wa], lambda: c.cluster("b", window=w)),
        (G.OP_DWITHIN, ["a", "a", "dwithin", "4", "0", "all", "distance"], lambda: c.dwithin("a", "a", "dwithin", 4, 0, "all", extra="distance")),
        (G.OP_DWITHIN, ["a", "a", "buffer", "-1", "-1.5", "distinct"], lambda: c.dwithin("a", "a", "buffer", -1, -1.5, "distinct")),
    ]
    for op, args, ours in cases:
        r = subprocess.run([GSP_CAT, daemon["sock"], str(op), *args], env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (args, r.stderr)
        n = int(r.stdout.split("rows=")[1].split()[0])
        rows, st, _ = ours()
        assert n == len(rows) and n == st["rows"], (args, n, len(rows))
        assert json.loads(r.stdout.splitlines()[1])["rows"] == n
    r = subprocess.run([GSP_CAT, daemon["sock"], str(G.OP_JOIN), "a", "zz", "0", "all"], env=env, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no such layer" in r.stderr


def test_dwithin_modes(ctx, daemon):
    c = daemon["client"]
    A = Side(ctx, daemon["sets"]["a"], "a", 8)
    r = ctx.dwithin(A.layer, A.layer, 4.0, same=True, report_distance=True, upper=True)
    rows, st, _ = c.dwithin("a", "a", "dwithin", 4, 0, "all", extra="distance")
    _same_rows(rows, expand(A, A, r.pairs, r.dist.tolist(), True, False))
    rows, _, _ = c.dwithin("a", "a", "buffer", 1.5, 2.5, "distinct")   # two positive buffers: dwithin 4 m
    _same_rows(rows, expand(A, A, r.pairs, [math.nan] * len(r.pairs), True, True))
    need = (np.diff(A.moff) > 1).astype(np.uint8)
    e = ctx.intersects_buffered(A.layer, A.layer, 1.0, 1.0, same=True, need_self=need, upper=True)
    rows, st, _ = c.dwithin("a", "a", "buffer", -1, -1, "distinct")
    _same_rows(rows, expand(A, A, e.pairs, [math.nan] * len(e.pairs), True, True))
    assert st["kind"] == "buffer_intersects"
    B = Side(ctx, daemon["sets"]["b"], "b", 0)
    pts = ctx.points(daemon["pts"])
    pe = ctx.points_intersects_buffered(pts, B.layer, 2.0)
    rows, _, _ = c.dwithin("p", "b", "buffer", 0, -2, "all")
    _same_rows(rows, [(f"pt{p}", B.ids[m], math.nan) for (p, g) in pe.pairs.tolist() for m in B.members(g)])
    with pytest.raises(G.DaemonError, match="opposite sign"):
        c.dwithin("a", "b", "buffer", 1, -1, "all")
    with pytest.raises(G.DaemonError, match="negative buffer"):
        c.dwithin("p", "b", "buffer", -1, -1, "all")


def test_concurrent_clients_get_their_own_rows(daemon):
    c = daemon["client"]
    reqs = [("join", ("a", "a", True, "all")), ("join", ("a", "b", False, "all")), ("relate", ("a", "a", "", "distinct")),
            ("dwithin", ("a", "a", "dwithin", 3, 0, "all")), ("join", ("p", "b", False, "all"))]
    alone = [getattr(c, k)(*a)[0] for k, a in reqs]
    out = [None] * (3 * len(reqs))

    def run(i):
        k, a = reqs[i % len(reqs)]
        out[i] = getattr(c, k)(*a)[0]

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(out))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, rows in enumerate(out):
        _same_rows(rows, alone[i % len(reqs)])


def test_concurrent_windowed_requests(daemon):
    """windowed requests of one unsubdivided layer from many clients at once (the daemon's lanes batch them): each
    client gets the rows of the same request sent alone"""
    c = daemon["client"]
    xy = daemon["sets"]["b"].xy
    x0, y0, x1, y1 = xy[:, 0].min(), xy[:, 1].min(), xy[:, 0].max(), xy[:, 1].max()
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(24):
        f = rng.uniform(0, 0.7, 2)
        w = (x0 + f[0] * (x1 - x0), y0 + f[1] * (y1 - y0), x0 + (f[0] + 0.3) * (x1 - x0), y0 + (f[1] + 0.3) * (y1 - y0))
        k = i % 4
        if k == 0:
            reqs.append(("join", ("b", "b"), dict(want_area=True, self_pairs="all", window=w)))
        elif k == 1:
            reqs.append(("relate", ("b", "b", "", "distinct"), dict(window=w)))
        elif k == 2:
            reqs.append(("dwithin", ("b", "b", "dwithin", 3, 0, "all"), dict(window=w)))
        else:
            reqs.append(("dwithin", ("b", "b", "buffer", -1, -1.5, "distinct"), dict(window=w)))
    alone = [getattr(c, k)(*a, **kw)[0] for k, a, kw in reqs]
    out = [None] * len(reqs)

    def run(i):
        k, a, kw = reqs[i]
        out[i] = G.Client(daemon["sock"], token=TOKEN).__getattribute__(k)(*a, **kw)[0]

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(reqs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, rows in enumerate(out):
        _same_rows(rows, alone[i])
    assert any(len(r) for r in alone)


def test_load_from_a_binary_copy_stream(daemon):
    """LOAD copy:FILE parses a binary COPY dump on the device (rows of a, its ids, plus NULL, empty and wrong-kind rows
    the loader skips, and a few rows with the other byte order and Z): the layer answers joins with the rows of the
    gset-loaded a"""
    import ewkb_rows as E

    c, A = daemon["client"], daemon["sets"]["a"]
    rows = []
    for g in range(A.n):
        rings = [[tuple(p) for p in A.xy[int(A.rvo[r]):int(A.rvo[r + 1])]] for r in range(int(A.gro[g]), int(A.gro[g + 1]))]
        rows.append((f"a{g}", ("polygon", rings)))
        if g % 97 == 5:
            rows += [(f"null{g}", None), (f"empty{g}", ("polygon", [])), (f"pt{g}", ("point", 1.0, 2.0))]
    opts = {i: dict(le=False, z=True) for i in range(0, len(rows), 11)}
    with open(os.path.join(daemon["dir"], "a.copy"), "wb") as f:
        f.write(E.stream(rows, opts))
    info = c.load("acopy", "copy:a.copy", subdivide=8)
    extra = sum(1 for g in range(A.n) if g % 97 == 5)
    assert info["rows"] == A.n and info["stream_rows"] == A.n + 3 * extra
    assert (info["skipped_null"], info["skipped_empty"], info["skipped_type"]) == (extra, extra, extra)
    assert info["distinct"] == daemon["info"]["a"]["distinct"] and info["vertices"] == daemon["info"]["a"]["vertices"]
    for op in (lambda x: c.join(x, x, True, "distinct"), lambda x: c.relate(x, "b", "", "all"), lambda x: c.dwithin(x, x, "dwithin", 3, 0, "distinct")):
        _same_rows(op("acopy")[0], op("a")[0])
    assert c.call(G.OP_DROP, "acopy") is True


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    """tests/srf_replay.c: the extension's way of reading streams (1 MiB buffer, u16 ids, row layouts per SRF)"""
    exe = tmp_path_factory.mktemp("srf") / "srf_replay"
    subprocess.run(["gcc", "-O2", "-o", str(exe), os.path.join(ROOT, "tests", "srf_replay.c"), "-lm"], check=True)
    return str(exe)


def _replay(exe, sock, kind, op, *args):
    r = subprocess.run([exe, sock, TOKEN, kind, str(op), *map(str, args)], capture_output=True, timeout=300)
    lines = r.stdout.decode().splitlines()
    stats = json.loads(lines[-1][len("#stats "):]) if r.returncode == 0 else None
    return r.returncode, (lines[:-1] if r.returncode == 0 else lines), stats


def _as_lines(kind, rows):
    if kind == "join":
        return [f"{a}\t{b}\t" + ("NULL" if math.isnan(v) else "%.17g" % v) for a, b, v in rows]
    if kind == "relate":
        return [f"{a}\t{b}\t{code}" for a, b, code in rows]
    if kind == "cluster":
        return [f"{a}\t{k}" for a, k in rows]
    return [f"{a}\t{b}" for a, b, _ in rows]


def test_the_extension_reader_reads_every_stream(daemon, replay):
    """every streaming op as the extension's SRFs request and read it (gpu_join, gpu_relate, gpu_cluster_intersecting,
    gpu_dwithin / gpu_intersects_buffered pair rows, the stats modes): the rows this daemon streams are the rows its
    Python client reads, each row once, and the stats JSON counts them"""
    c, sock = daemon["client"], daemon["sock"]
    cases = [("join", G.OP_JOIN, ("a", "a", "1", "distinct"), lambda: c.join("a", "a", True, "distinct")),
             ("join", G.OP_JOIN, ("a", "b", "0", "all"), lambda: c.join("a", "b", False, "all")),
             ("relate", G.OP_RELATE, ("a", "b", "", "all"), lambda: c.relate("a", "b", "", "all")),
             ("relate", G.OP_RELATE, ("a", "a", "T*T***T**", "distinct"), lambda: c.relate("a", "a", "T*T***T**", "distinct")),
             ("cluster", G.OP_CLUSTER, ("a",), lambda: c.cluster("a")),
             ("pair", G.OP_DWITHIN, ("a", "a", "dwithin", "3", "0", "distinct"), lambda: c.dwithin("a", "a", "dwithin", 3, 0, "distinct")),
             ("pair", G.OP_DWITHIN, ("a", "b", "buffer", "-1", "-0.5", "all"), lambda: c.dwithin("a", "b", "buffer", -1, -0.5, "all")),
             ("pair", G.OP_DWITHIN, ("p", "b", "dwithin", "10", "0", "all"), lambda: c.dwithin("p", "b", "dwithin", 10, 0, "all"))]
    for kind, op, args, ref in cases:
        rc, lines, stats = _replay(replay, sock, kind, op, *args)
        assert rc == 0, (args, lines[-3:])
        rows, _, _ = ref()
        assert sorted(lines) == sorted(_as_lines(kind, rows)), args
        assert len(lines) == len(set(lines)) or kind == "cluster"
        assert stats["rows"] == len(lines), (args, stats)
    for op, args in ((G.OP_JOIN, ("a", "a", "1", "all", "stats")), (G.OP_DWITHIN, ("a", "a", "dwithin", "3", "0", "all", "stats"))):
        rc, lines, stats = _replay(replay, sock, "stats", op, *args)
        assert rc == 0 and lines == [] and stats["rows"] > 0, (args, lines)
    rc, lines, _ = _replay(replay, sock, "join", G.OP_JOIN, "nope", "a", "1", "all")
    assert rc == 1 and lines[0].startswith("ERROR: gpujoind: ") and "nope" in lines[0]
    rc, lines, _ = _replay(replay, sock, "join", G.OP_JOIN, "a", "a", "1", "some")
    assert rc == 1 and "self_pairs" in lines[0]


def test_ids_at_the_u16_limit(daemon, replay):
    """ids travel as u16 length + bytes: one of 65535 bytes arrives whole, a longer one is cut to 65535 bytes (as the
    reference daemon cuts it); rows of two such ids (131 KB each) pass the extension's 1 MiB buffer"""
    d = daemon["dir"]
    sq = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0), (0.0, 0.0)]
    sq2 = [(0.5, 0.5), (1.5, 0.5), (1.5, 1.5), (0.5, 1.5), (0.5, 0.5)]
    hs = gsj.HostSet(np.array([0, 1, 2], np.uint64), np.array([0, 5, 10], np.uint64), np.array(sq + sq2, float))
    gsj.write_gset(os.path.join(d, "long.gset"), hs)
    with open(os.path.join(d, "long.gset.ids"), "w") as f:
        f.write("x" * 65535 + "\n" + "y" * 70000 + "\n")
    c = daemon["client"]
    assert c.load("long", "gset:long.gset")["rows"] == 2
    rc, lines, stats = _replay(replay, daemon["sock"], "join", G.OP_JOIN, "long", "long", "1", "all")
    assert rc == 0 and stats["rows"] == 4
    ids = {tuple(l.split("\t")[:2]) for l in lines}
    assert ("x" * 65535, "y" * 65535) in ids and ("y" * 65535, "y" * 65535) in ids
    assert c.call(G.OP_DROP, "long") is True


def test_drop(daemon):
    c = daemon["client"]
    assert c.load("tmp", "gen:parcels_grid:nx=10;ny=10;seed=1")["rows"] == 100
    assert c.call(G.OP_DROP, "tmp") is True and c.call(G.OP_DROP, "tmp") is False


def test_row_limit_ends_the_stream_with_a_frame_error(daemon):
    """A second daemon with --max-rows: a join whose rows would exceed the limit opens its stream, then ends it with a
    GSP_FRAME_ERR frame (daemon/protocol.h's mid-stream error), read as an error by this repository's client and by
    the reference's own client; a request under the limit still completes on the same daemon."""
    d = daemon["dir"]
    sock = os.path.join(d, "limited.sock")
    log = open(os.path.join(d, "limited.log"), "w")
    proc = subprocess.Popen([DAEMON, "--socket", sock, "--data-dir", d, "--token-file", os.path.join(d, "token"), "--max-rows", "200",
                             "--lanes", "2", "--threads", "2"], stdout=log, stderr=log)
    try:
        for _ in range(600):
            if os.path.exists(sock) or proc.poll() is not None:
                break
            time.sleep(0.1)
        assert proc.poll() is None
        c = G.Client(sock, token=TOKEN)
        c.load("a", "gset:a.gset", subdivide=8)
        c.load("tiny", "gen:parcels_grid:nx=3;ny=3;seed=1")
        rows, st, _ = c.join("tiny", "tiny", want_area=True)
        assert len(rows) == st["rows"] == 9
        with pytest.raises(G.DaemonError, match="row limit of 200"):
            c.join("a", "a", want_area=True)
        exe = os.path.join(tempfile.mkdtemp(), "srf_replay")   # the extension's reader raises it as its ERROR
        subprocess.run(["gcc", "-O2", "-o", exe, os.path.join(ROOT, "tests", "srf_replay.c"), "-lm"], check=True)
        rc, lines, _ = _replay(exe, sock, "join", G.OP_JOIN, "a", "a", "1", "all")
        assert rc == 1 and lines[-1].startswith("ERROR: gpujoind: row limit of 200"), lines[-3:]
        if os.path.exists(GSP_CAT):
            r = subprocess.run([GSP_CAT, sock, str(G.OP_JOIN), "a", "a", "1", "all"], env=dict(os.environ, GSP_TOKEN=TOKEN),
                               capture_output=True, text=True, timeout=60)
            assert r.returncode == 1 and "stream error: row limit of 200" in r.stderr, r.stderr
    finally:
        proc.terminate()
        proc.wait(timeout=60)
        log.close()


def test_layer_files_stay_inside_the_data_directory(daemon):
    c = daemon["client"]
    outside = tempfile.mkdtemp()
    A = daemon["sets"]["a"]
    gsj.write_gset(os.path.join(outside, "x.gset"), A)
    os.symlink(os.path.join(outside, "x.gset"), os.path.join(daemon["dir"], "link.gset"))
    for src in (f"gset:{outside}/x.gset", "gset:../" + os.path.basename(outside) + "/x.gset", "gset:link.gset", "gset:/etc/passwd"):
        with pytest.raises(G.DaemonError, match="outside the data directory|cannot open"):
            c.load("evil", src)
    with pytest.raises(G.DaemonError, match="cannot open"):
        c.load("missing", "gset:nope.gset")
    assert c.load("abs", "gset:" + os.path.join(daemon["dir"], "a.gset"))["rows"] == A.n   # absolute, but inside


GSJLOAD = os.path.join(ROOT, "postgis-cuda_amd", "bin", "gsjload")


def test_load_client_drains_streams_and_times_them(daemon):
    """gsjload (the open-loop client of the serving benchmark) sends each request at its arrival time on a connection
    of its own and drains every row: its row counts equal what this repository's client parses from the same
    requests, errors are reported per request, and every latency is positive."""
    c = daemon["client"]
    xy = daemon["sets"]["b"].xy
    w = "window=" + ",".join(repr(float(v)) for v in (xy[:, 0].min(), xy[:, 1].min(), xy[:, 0].mean(), xy[:, 1].mean()))
    reqs = [(G.OP_JOIN, ["a", "a", "1", "all"]), (G.OP_RELATE, ["a", "b", "", "distinct"]), (G.OP_CLUSTER, ["b", w]),
            (G.OP_DWITHIN, ["b", "b", "dwithin", "3", "0", "distinct", w]), (G.OP_JOIN, ["p", "b", "0", "all"]), (G.OP_JOIN, ["a", "zz", "0", "all"])]
    text = "".join("\t".join([f"{0.01 * i:.3f}", str(op)] + args) + "\n" for i, (op, args) in enumerate(reqs))
    r = subprocess.run([GSJLOAD, "--socket", daemon["sock"], "--token", TOKEN], input=text, capture_output=True, text=True, timeout=120)
    recs = [json.loads(x) for x in r.stdout.splitlines()]
    per, summary = recs[:len(reqs)], recs[-1]
    assert len(recs) == len(reqs) + 1 and summary["summary"] and summary["failed"] == 1 and r.returncode == 2
    for (op, args), rec in zip(reqs[:-1], per[:-1]):
        rows, st, _ = c.call(op, *args)
        assert rec["ok"] and rec["rows"] == len(rows) == rec["stats"]["rows"] and rec["latency_ms"] > 0, (args, rec)
    assert not per[-1]["ok"] and "no such layer" in per[-1]["error"]   # the unknown layer
    assert summary["rows"] == sum(x["rows"] for x in per) and summary["p99_ms"] >= summary["p50_ms"] > 0
