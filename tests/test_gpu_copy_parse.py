"""GPU: the device parse of binary COPY streams (gsj_copy_parse_device) against the oracle's restatement of the
layer loader (pinned by tests/test_copy_parse.py): every rule's rows, sliver rings, point layers, randomly encoded
rows, a 250k-parcel stream written by gsj_copy_write, streams read in pieces, and the argument errors. Arrays are
compared bit for bit, ids and the loader's counters exactly."""
import numpy as np
import pytest

import ewkb_rows as E
import gsj
from oracle import oracle as O
from test_copy_parse import cases, sliver, sq

pytestmark = pytest.mark.gpu


def same(dev, orc):
    st = dev.stats
    assert (st["rows"], st["skipped_null"], st["skipped_type"], st["skipped_empty"], st["polys"], st["points"]) == \
        (orc["rows"], orc["nulls"], orc["bad"], orc["empty"], orc["polys"], orc["points"]), (st, orc["stats"])
    assert (st["consumed"], bool(st["done"])) == (orc["consumed"], orc["done"])
    assert dev.ids == orc["ids"]
    if orc["points"]:
        assert dev.polys is None and np.array_equal(dev.pts.view(np.uint64), orc["pxy"].view(np.uint64))
    else:
        assert dev.pts is None
        assert np.array_equal(dev.polys.gro, orc["gro"]) and np.array_equal(dev.polys.rvo, orc["rvo"])
        assert np.array_equal(dev.polys.xy.view(np.uint64), orc["xy"].view(np.uint64))


def test_every_rule(ctx):
    rows, opts = cases()
    buf = E.stream(rows, opts)
    same(gsj.copy_parse(ctx, buf), O.copy_parse(buf))


def test_sliver_rings_and_point_layers(ctx):
    rng = np.random.default_rng(11)
    rows = []
    for i in range(2000):
        x0, y0 = -118.5 + rng.uniform(0, 0.5), 33.8 + rng.uniform(0, 0.5)
        w, h = rng.uniform(1e-7, 3e-7), rng.uniform(0.5e-7, 1e-7)
        rows.append((f"s{i}", ("polygon", [sliver(x0, y0, w, h, ccw=bool(rng.integers(2))),
                                           sliver(x0 + 1e-8, y0 + 1e-8, w / 4, h / 4, ccw=bool(rng.integers(2)))])))
    buf = E.stream(rows)
    same(gsj.copy_parse(ctx, buf), O.copy_parse(buf))
    pts = [("e", ("polygon", [])), ("p0", ("point", 1.5, 2.5)), ("x", ("polygon", [sq(0, 0, 1)])), ("e1", ("polygon", [])),
           ("p1", ("point", -3.0, 4.0)), ("m", ("point", float("nan"), 1.0)), ("n", None)]
    buf = E.stream(pts, {1: dict(srid=4326, z=True), 4: dict(le=False, m=True, iso=True)})
    d = gsj.copy_parse(ctx, buf)
    same(d, O.copy_parse(buf))
    assert d.stats["kind"] == 2


def random_rows(rng, n):
    rows, opts = [], {}
    for i in range(n):
        x0, y0 = rng.uniform(-120, -70), rng.uniform(25, 48)
        k = rng.integers(8)
        if k < 3:
            g = ("polygon", [sq(x0, y0, 0.01, bool(rng.integers(2))), sq(x0 + 0.002, y0 + 0.002, 0.003, bool(rng.integers(2)))][: 1 + k % 2])
        elif k == 3:
            g = ("multi", [("polygon", [sq(x0, y0, 0.01)]), ("polygon", [sq(x0 + 0.02, y0, 0.01, False)])])
        elif k == 4:
            g = ("coll", [("polygon", [sq(x0, y0, 0.01)]), ("coll", [("multi", [("polygon", [[(0, 0), (1, 0), (0, 0)], sq(x0, y0, 0.1)])])])])
        elif k == 5:
            g = None
        elif k == 6:
            g = ("point", x0, y0)
        else:
            g = ("raw", E.polygon([sq(x0, y0, 0.01)])[: int(rng.integers(1, 60))], ("bad",))
        rows.append((f"r{i}" if rng.integers(20) else None, g))
        opts[i] = dict(le=bool(rng.integers(2)), srid=4326 if rng.integers(2) else None, z=bool(rng.integers(2)), m=bool(rng.integers(4) == 0),
                       iso=bool(rng.integers(2)))
    return rows, opts


def test_randomly_encoded_rows_and_pieces(ctx):
    rng = np.random.default_rng(3)
    rows, opts = random_rows(rng, 4000)
    buf = E.stream(rows, opts)
    whole = gsj.copy_parse(ctx, buf)
    same(whole, O.copy_parse(buf))
    # the same stream read in pieces: the rest of each piece and the layer kind carry over
    cuts = sorted(rng.integers(1, len(buf), 9).tolist()) + [len(buf)]
    carry, pos, kind, first, parts = b"", 0, 0, True, []
    for c in cuts:
        chunk = carry + buf[pos:c]
        pos = c
        d = gsj.copy_parse(ctx, chunk, header=first, kind=kind)
        o = O.copy_parse(chunk, header=first, kind=kind)
        same(d, o)
        if first and d.stats["consumed"] == 0:
            carry = chunk
            continue
        first, kind = False, d.stats["kind"]
        parts.append(d)
        carry = chunk[d.stats["consumed"]:]
    assert parts[-1].stats["done"] and carry == b""
    assert [i for p in parts for i in p.ids] == whole.ids
    assert np.array_equal(np.concatenate([p.polys.xy for p in parts if p.polys is not None]).view(np.uint64), whole.polys.xy.view(np.uint64))
    for k in ("rows", "skipped_null", "skipped_type", "skipped_empty", "polys", "points"):
        assert sum(p.stats[k] for p in parts) == whole.stats[k]


def test_a_written_mesh_parses_back(ctx):
    M = gsj.generate("parcels_mesh", nx=500, ny=500, w=1 / 1549, h=1 / 1549, seed=1)
    buf = gsj.copy_write(M, srid=4326)
    d = gsj.copy_parse(ctx, buf)
    same(d, O.copy_parse(buf))
    assert np.array_equal(d.polys.gro, M.gro) and np.array_equal(d.polys.rvo, M.rvo)
    assert np.array_equal(d.polys.xy.view(np.uint64), M.xy.view(np.uint64))
    assert d.ids == [str(i) for i in range(M.n)]


def test_argument_errors(ctx):
    buf = E.stream([("a", ("polygon", [sq(0, 0, 1)]))])
    with pytest.raises(gsj.GsjError, match="PGCOPY"):
        gsj.copy_parse(ctx, b"NOTACOPY\n\xff\r\n\x00" + buf[11:])
    with pytest.raises(gsj.GsjError, match="layer kind"):
        gsj.copy_parse(ctx, buf, kind=3)
    empty = gsj.copy_parse(ctx, buf[:15])   # an incomplete header: nothing read yet
    assert empty.stats["consumed"] == 0 and empty.stats["rows"] == 0 and empty.polys.n == 0
