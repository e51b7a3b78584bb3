"""Test-side EWKB / PGCOPY writer (pure Python, independent of the library and the oracle): geometries described as
Python values are encoded in every form the loader accepts (either byte order, EWKB SRID / Z / M flags, ISO type
codes, collections) and framed as a binary COPY stream of (id text, geom) rows. `expect` states what the layer
loader's rules make of the same rows (daemon/loader.h:72-166), from the geometry values, with ring orientation decided
by the exact rational signed area; the tests compare the oracle (and through it the engine) with it."""
import math
import struct
from fractions import Fraction

SIG = b"PGCOPY\n\xff\r\n\x00"


def u32(v, le):
    return struct.pack("<I" if le else ">I", v)


def f64(v, le):
    return struct.pack("<d" if le else ">d", v)


def header(base, le=True, srid=None, z=False, m=False, iso=False):
    if iso:
        t = base + (3000 if z and m else 1000 if z else 2000 if m else 0)
    else:
        t = base | (0x80000000 if z else 0) | (0x40000000 if m else 0)
    if srid is not None:
        t |= 0x20000000
    out = bytes([1 if le else 0]) + u32(t, le)
    return out + (u32(srid, le) if srid is not None else b"")


def coords(pts, le, z=False, m=False):
    out = b""
    for p in pts:
        out += f64(p[0], le) + f64(p[1], le)
        if z:
            out += f64(1.5, le)
        if m:
            out += f64(-2.5, le)
    return out


def point(x, y, **kw):
    le, z, m = kw.get("le", True), kw.get("z", False), kw.get("m", False)
    return header(1, **kw) + coords([(x, y)], le, z, m)


def polygon(rings, **kw):
    le, z, m = kw.get("le", True), kw.get("z", False), kw.get("m", False)
    out = header(3, **kw) + u32(len(rings), le)
    for r in rings:
        out += u32(len(r), le) + coords(r, le, z, m)
    return out


def collection(members, base=6, **kw):
    le = kw.get("le", True)
    return header(base, **kw) + u32(len(members), le) + b"".join(members)


def linestring(pts, **kw):
    le = kw.get("le", True)
    return header(2, **kw) + u32(len(pts), le) + coords(pts, le)


def copy_stream(rows, trailer=True, ext=b""):
    """rows: (id str | None, geom bytes | None); binary COPY with two fields per row"""
    out = SIG + struct.pack(">I", 0) + struct.pack(">I", len(ext)) + ext
    for rid, g in rows:
        out += struct.pack(">h", 2)
        out += struct.pack(">i", -1) if rid is None else struct.pack(">i", len(rid.encode())) + rid.encode()
        out += struct.pack(">i", -1) if g is None else struct.pack(">i", len(g)) + g
    if trailer:
        out += struct.pack(">h", -1)
    return out


def exact_area2(ring):
    """twice the signed area of the closed ring's fan from its first point over points 1 .. n-2, exactly"""
    n = len(ring)
    ox, oy = Fraction(ring[0][0]), Fraction(ring[0][1])
    s = Fraction(0)
    for i in range(1, n - 2):
        ax, ay = Fraction(ring[i][0]) - ox, Fraction(ring[i][1]) - oy
        bx, by = Fraction(ring[i + 1][0]) - ox, Fraction(ring[i + 1][1]) - oy
        s += ax * by - ay * bx
    return s


def oriented(ring, shell):
    a = exact_area2(ring)
    return list(ring) if (a > 0) == shell else list(reversed(ring))


# ---------------------------------------------------------------------------------------------- abstract geometries
# ("point", x, y) | ("polygon", [ring, ...]) | ("multi", [geom, ...]) | ("coll", [geom, ...]) | ("line", [pt, ...])
# rings are lists of (x, y); encode() writes one as EWKB with the options of the row (byte order, SRID, Z, M, ISO)

def encode(g, le=True, srid=None, z=False, m=False, iso=False):
    kw = dict(le=le, z=z, m=m, iso=iso)
    top = dict(kw, srid=srid)
    if g[0] == "point":
        return point(g[1], g[2], **top)
    if g[0] == "polygon":
        return polygon(g[1], **top)
    if g[0] == "line":
        return header(2, **top) + u32(len(g[1]), le) + coords(g[1], le, z, m)
    members = [encode(x, **kw) for x in g[1]]
    return collection(members, base=6 if g[0] == "multi" else 7, **top)


def classify(g, depth=0):
    """what the loader's rules make of one value: ("bad",) | ("point", x, y) | ("poly", [(ring, shell), ...]) |
    ("empty",); a polygonal value without kept rings is ("poly", [])"""
    if g[0] == "point":
        if math.isnan(g[1]) or math.isnan(g[2]):
            return ("empty",)
        return ("bad",) if depth > 0 else ("point", g[1], g[2])
    if g[0] == "polygon":
        return ("poly", [(r, k == 0) for k, r in enumerate(g[1]) if len(r) >= 4])
    if g[0] in ("multi", "coll"):
        kept, poly = [], False
        for x in g[1]:
            c = classify(x, depth + 1)
            if c[0] == "bad" or c[0] == "point":
                return ("bad",)
            if c[0] == "poly":
                poly = True
                kept += c[1]
        return ("poly", kept) if poly else ("empty",)
    return ("bad",)


def expect(rows, kind_in=0):
    """the layer the rules build from rows of (id, geometry value | raw bytes marked ("raw", bytes, cls) | None)"""
    o = dict(gro=[0], rvo=[0], xy=[], pxy=[], ids=[], row=[], rows=0, nulls=0, bad=0, empty=0, polys=0, points=0)
    have_poly, have_pt = kind_in == 1, kind_in == 2
    for i, (rid, g) in enumerate(rows):
        o["rows"] += 1
        if rid is None or g is None:
            o["nulls"] += 1
            continue
        c = g[2] if g[0] == "raw" else classify(g)
        if c[0] == "bad" or (have_poly and c[0] == "point") or (have_pt and c[0] == "poly"):
            o["bad"] += 1
        elif c[0] == "empty" or (c[0] == "poly" and not c[1]):
            o["empty"] += 1
        elif c[0] == "point":
            have_pt = True
            o["points"] += 1
            o["pxy"].append((c[1], c[2]))
            o["ids"].append(rid)
            o["row"].append(i)
        else:
            have_poly = True
            o["polys"] += 1
            for r, shell in c[1]:
                o["xy"] += oriented(r, shell)
                o["rvo"].append(len(o["xy"]))
            o["gro"].append(len(o["rvo"]) - 1)
            o["ids"].append(rid)
            o["row"].append(i)
    return o


def stream(rows, opts=None, trailer=True):
    """rows as a binary COPY stream; opts[i]: encode() options of row i"""
    enc = []
    for i, (rid, g) in enumerate(rows):
        if g is None:
            enc.append((rid, None))
        elif g[0] == "raw":
            enc.append((rid, g[1]))
        else:
            enc.append((rid, encode(g, **((opts or {}).get(i, {})))))
    return copy_stream(enc, trailer=trailer)
