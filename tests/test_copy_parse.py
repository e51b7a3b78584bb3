"""CPU tests pinning the oracle's binary COPY / EWKB parse (oracle/ewkb.cpp, restating daemon/loader.h:46-170)
against values written here independently (tests/ewkb_rows.py): every encoding the loader accepts (byte orders, SRID,
Z, M, ISO type codes, collections), the rows it skips (NULL, wrong kind, unsupported, empty, malformed), the ring
index rule of shells and holes, the layer kind of the first accepted row, streams cut into pieces, and ring
orientation decided by the exact rational sign of each ring's area, down to the sliver rings (1e-14 square degrees at
Los Angeles) whose naive shoelace sign is wrong. Two planted mistakes in the oracle (the naive shoelace, the first
kept ring taken as the shell) must each fail."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import ewkb_rows as E
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAN = float("nan")


def sq(x0, y0, s, ccw=True):
    r = [(x0, y0), (x0 + s, y0), (x0 + s, y0 + s), (x0, y0 + s), (x0, y0)]
    return r if ccw else list(reversed(r))


def sliver(x0, y0, w, h, ccw=True):
    """a thin quadrilateral of twice-area about w h at (x0, y0): exact sign by construction"""
    r = [(x0, y0), (x0 + w, y0 + h * 0.25), (x0 + w * 0.5, y0 + h), (x0 - w * 0.5, y0 + h * 0.75), (x0, y0)]
    return r if ccw else list(reversed(r))


def check(got, exp):
    assert got["rows"] == exp["rows"] and got["nulls"] == exp["nulls"], (got["stats"], exp)
    assert (got["bad"], got["empty"], got["polys"], got["points"]) == (exp["bad"], exp["empty"], exp["polys"], exp["points"]), got["stats"]
    assert got["ids"] == exp["ids"] and got["row"].tolist() == exp["row"]
    assert got["gro"].tolist() == exp["gro"] and got["rvo"].tolist() == exp["rvo"]
    assert np.array_equal(got["xy"].view(np.uint64), np.array(exp["xy"], float).reshape(-1, 2).view(np.uint64))
    assert np.array_equal(got["pxy"].view(np.uint64), np.array(exp["pxy"], float).reshape(-1, 2).view(np.uint64))


def cases():
    """(rows, per-row encoding options) covering every rule"""
    A = sq(-118.25, 34.05, 0.001)
    Acw = sq(-118.25, 34.05, 0.001, ccw=False)
    H = sq(-118.2498, 34.0502, 0.0002)          # a hole of A, counter-clockwise as written
    B = sq(-118.248, 34.05, 0.0005, ccw=False)
    rows = [
        ("a", ("polygon", [Acw, H])),                                   # shell and hole both the wrong way round
        ("b", ("polygon", [A])),
        ("c", ("multi", [("polygon", [Acw, H]), ("polygon", [B])])),    # every polygon's first ring is a shell
        ("d", ("coll", [("polygon", [B]), ("point", NAN, NAN), ("coll", [("polygon", [A])])])),
        ("e", ("polygon", [[(0, 0), (1, 1), (0, 0)], Acw])),           # ring 0 has 3 points: skipped; ring 1 stays a hole
        ("f", None),                                                    # NULL geometry
        (None, ("polygon", [A])),                                       # NULL id
        ("g", ("point", -118.0, 34.0)),                                 # a point in a polygon layer
        ("h", ("coll", [("polygon", [A]), ("point", 1.0, 2.0)])),      # a point inside a collection
        ("i", ("line", [(0, 0), (1, 1)])),
        ("j", ("polygon", [])),                                         # POLYGON EMPTY
        ("k", ("polygon", [[(0, 0), (1, 0), (0, 0)]])),                # no ring of 4 points
        ("l", ("multi", [])),
        ("m", ("point", NAN, NAN)),                                     # POINT EMPTY
        ("n", ("raw", E.polygon([A])[:-5], ("bad",))),                  # truncated value
        ("o", ("raw", b"", ("bad",))),                                  # zero-length value
        ("p", ("raw", E.header(4) + E.u32(0, True), ("bad",))),         # MULTIPOINT EMPTY: unsupported type
        ("", ("polygon", [[(0, 0), (2, 0), (4, 0), (0, 0)]])),          # zero-area shell: reversed (area not > 0)
        ("q", ("polygon", [A, [(0, 0), (0, 2), (0, 4), (0, 0)]])),      # zero-area hole: kept as written
        ("r", ("polygon", [sliver(-118.2437, 34.0522, 2e-7, 1e-7), sliver(-118.2437, 34.0522, 2e-7, 1e-7, ccw=False)])),
        ("s", ("polygon", [sliver(-118.2437, 34.0522, 2e-7, 1e-7, ccw=False), sliver(-118.2437, 34.0521, 1e-7, 2e-7)])),
    ]
    opts = {0: dict(le=False), 1: dict(srid=4326), 2: dict(z=True), 3: dict(m=True, le=False), 4: dict(z=True, m=True, iso=True),
            7: dict(iso=True, z=True), 17: dict(srid=4269, le=False), 19: dict(z=True, srid=4326), 20: dict(le=False, m=True, iso=True)}
    return rows, opts


def test_every_rule_against_independent_values():
    rows, opts = cases()
    got = O.copy_parse(E.stream(rows, opts))
    exp = E.expect(rows)
    check(got, exp)
    assert got["done"] and got["consumed"] == len(E.stream(rows, opts))
    assert exp["polys"] == 9 and exp["bad"] == 6 and exp["empty"] == 4 and exp["nulls"] == 2


def test_sliver_rings_take_the_exact_sign():
    """rings of 2e-14 square degrees at Los Angeles: the plain shoelace on absolute coordinates rounds products of
    ~4000 square degrees (ulp ~1e-12), far above the area; the compensated relative sum keeps the exact sign"""
    rng = np.random.default_rng(5)
    rows = []
    for i in range(400):
        x0, y0 = -118.5 + rng.uniform(0, 0.5), 33.8 + rng.uniform(0, 0.5)
        w, h = rng.uniform(1e-7, 3e-7), rng.uniform(0.5e-7, 1e-7)
        shell = sliver(x0, y0, w, h, ccw=bool(rng.integers(2)))
        hole = sliver(x0 + 1e-8, y0 + 1e-8, w / 4, h / 4, ccw=bool(rng.integers(2)))
        rows.append((f"s{i}", ("polygon", [shell, hole])))
    check(O.copy_parse(E.stream(rows)), E.expect(rows))
    flips = sum((E.exact_area2(r[1][1][0]) > 0) != (O.ring_twice_area(np.array(r[1][1][0])) > 0) for r in rows)
    assert flips == 0


def test_point_layers_take_the_first_accepted_kind():
    rows = [("e0", ("polygon", [[(0, 0), (1, 0), (0, 0)]])),   # no rings, before any point: empty
            ("p0", ("point", 1.5, 2.5)),
            ("x", ("polygon", [sq(0, 0, 1)])),                 # a polygon after the first point: wrong kind
            ("e1", ("polygon", [])),                           # no rings after the first point: wrong kind
            ("p1", ("point", -3.0, 4.0)),
            ("m", ("point", NAN, 1.0))]
    got = O.copy_parse(E.stream(rows, {1: dict(srid=4326, z=True), 4: dict(le=False)}))
    exp = E.expect(rows)
    check(got, exp)
    assert exp["points"] == 2 and exp["bad"] == 2 and exp["empty"] == 2


@pytest.mark.parametrize("seed", [0, 1])
def test_a_stream_parsed_in_pieces_equals_the_whole(seed):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(300):
        x0, y0 = rng.uniform(-120, -70), rng.uniform(25, 48)
        k = rng.integers(4)
        if k == 0:
            rows.append((f"r{i}", ("polygon", [sq(x0, y0, 0.01, bool(rng.integers(2)))])))
        elif k == 1:
            rows.append((f"r{i}", ("multi", [("polygon", [sq(x0, y0, 0.01)]), ("polygon", [sq(x0 + 0.02, y0, 0.01, False)])])))
        elif k == 2:
            rows.append((f"r{i}", None))
        else:
            rows.append((f"r{i}", ("point", x0, y0)))
    opts = {i: dict(le=bool(rng.integers(2)), srid=4326 if rng.integers(2) else None, z=bool(rng.integers(2))) for i in range(len(rows))}
    buf = E.stream(rows, opts)
    whole = O.copy_parse(buf)
    check(whole, E.expect(rows))
    cuts = sorted(rng.integers(1, len(buf), 6).tolist()) + [len(buf)]
    parts, carry, pos, kind, first = [], b"", 0, 0, True
    for c in cuts:
        chunk = carry + buf[pos:c]
        pos = c
        r = O.copy_parse(chunk, header=first, kind=kind)
        if first and r["consumed"] == 0:
            carry = chunk
            continue
        first = False
        kind = kind or (1 if r["polys"] else 2 if r["points"] else 0)
        parts.append(r)
        carry = chunk[r["consumed"]:]
    assert parts[-1]["done"] and carry == b""
    assert sum(p["rows"] for p in parts) == whole["rows"]
    assert [i for p in parts for i in p["ids"]] == whole["ids"]
    assert np.array_equal(np.concatenate([p["xy"] for p in parts]).view(np.uint64), whole["xy"].view(np.uint64))
    for k in ("bad", "empty", "nulls", "polys", "points"):
        assert sum(p[k] for p in parts) == whole[k]


def test_incomplete_rows_and_headers():
    rows = [("a", ("polygon", [sq(0, 0, 1)])), ("b", ("polygon", [sq(2, 0, 1)]))]
    buf = E.stream(rows)
    r = O.copy_parse(buf[:-5])                     # the second row cut short, no trailer
    assert r["rows"] == 1 and not r["done"] and r["consumed"] == len(E.stream(rows[:1], trailer=False))
    assert O.copy_parse(buf[:10])["consumed"] == 0  # an incomplete header waits for more
    with pytest.raises(ValueError, match="malformed COPY"):
        O.copy_parse(b"NOTACOPY\n\xff\r\n\x00" + buf[11:])
    ext = E.copy_stream([("a", E.polygon([sq(0, 0, 1)]))], ext=b"12345")   # a header extension is skipped
    assert O.copy_parse(ext)["polys"] == 1


@pytest.mark.parametrize("mutant,test", [(6, "test_sliver_rings_take_the_exact_sign"), (7, "test_every_rule_against_independent_values")])
def test_planted_mistakes_fail(mutant, test):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mutants"], check=True, capture_output=True)
    env = dict(os.environ, GSJ_ORACLE_LIB=os.path.join(ROOT, "oracle", "build", f"liboracle_m{mutant}.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", f"{os.path.abspath(__file__)}::{test}", "-p", "no:cacheprovider"],
                       env=env, capture_output=True, text=True, cwd=ROOT)
    assert r.returncode != 0, r.stdout[-2000:]


def test_written_streams_parse_back_to_the_set():
    """gsj_copy_write (host) -> the oracle's parse: a set whose rings are already oriented comes back unchanged
    (shells counter-clockwise start polygons, clockwise rings are their holes: POLYGON or MULTIPOLYGON values)"""
    import gsj

    M = gsj.generate("parcels_mesh", nx=30, ny=30, w=1 / 1549, h=1 / 1549, seed=3)
    shell, hole, other = sq(0, 0, 4), sq(1, 1, 1, ccw=False), sq(10, 0, 2)
    rings = [shell, hole, other, sq(20, 0, 1)]
    gro = np.concatenate([M.gro, [M.gro[-1] + 3, M.gro[-1] + 4]]).astype(np.uint64)
    lens = np.concatenate([np.diff(M.rvo.astype(np.int64)), [len(r) for r in rings]])
    rvo = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    xy = np.concatenate([M.xy, np.array([p for r in rings for p in r], float)])
    hs = gsj.HostSet(gro, rvo, xy)
    ids = [f"parcel-{i}" for i in range(hs.n)]
    for srid in (0, 4326):
        r = O.copy_parse(gsj.copy_write(hs, ids=ids, srid=srid))
        assert r["polys"] == hs.n and r["bad"] == 0 and r["ids"] == ids
        assert np.array_equal(r["gro"], hs.gro) and np.array_equal(r["rvo"], hs.rvo)
        assert np.array_equal(r["xy"].view(np.uint64), hs.xy.view(np.uint64))
    assert O.copy_parse(gsj.copy_write(hs))["ids"] == [str(i) for i in range(hs.n)]
