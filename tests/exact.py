"""Exact rational geometry (fractions.Fraction) used to pin the oracle: every routine here is the textbook
definition, written independently of the engine's floating-point formulation."""
from fractions import Fraction as F


def fr(x):
    return x if isinstance(x, F) else F(float(x))


def orient(a, b, c):
    """exact (b - a) x (c - a)"""
    ax, ay, bx, by, cx, cy = map(fr, (*a, *b, *c))
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)


def sign(v):
    return (v > 0) - (v < 0)


def left_of_sos(a, b, p):
    """p + (eps, eps^2) strictly left of a->b, eps -> 0+ (symbolic perturbation)"""
    o = orient(a, b, p)
    if o != 0:
        return o > 0
    ax, ay, bx, by = map(fr, (*a, *b))
    # d/dpx orient = -(by - ay), d/dpy orient = (bx - ax)
    t1 = -(by - ay)
    if t1 != 0:
        return t1 > 0
    return (bx - ax) > 0


def on_segment(p, a, b):
    if orient(a, b, p) != 0:
        return False
    return min(fr(a[0]), fr(b[0])) <= fr(p[0]) <= max(fr(a[0]), fr(b[0])) and min(fr(a[1]), fr(b[1])) <= fr(p[1]) <= max(fr(a[1]), fr(b[1]))


def segments_intersect(p0, p1, q0, q1):
    """closed segments share a point"""
    d1, d2 = sign(orient(q0, q1, p0)), sign(orient(q0, q1, p1))
    d3, d4 = sign(orient(p0, p1, q0)), sign(orient(p0, p1, q1))
    if d1 * d2 < 0 and d3 * d4 < 0:
        return True
    return on_segment(p0, q0, q1) or on_segment(p1, q0, q1) or on_segment(q0, p0, p1) or on_segment(q1, p0, p1)


def point_in_rings(p, rings):
    """even-odd point-in-polygon, exact: 2 on the boundary, 1 interior, 0 outside"""
    px, py = fr(p[0]), fr(p[1])
    inside = False
    for ring in rings:
        for i in range(len(ring) - 1):
            a, b = ring[i], ring[i + 1]
            if on_segment(p, a, b):
                return 2
            ax, ay, bx, by = fr(a[0]), fr(a[1]), fr(b[0]), fr(b[1])
            if (ay > py) != (by > py):
                x = ax + (py - ay) * (bx - ax) / (by - ay)
                if x > px:
                    inside = not inside
    return 1 if inside else 0


def shoelace(ring):
    s = F(0)
    for i in range(len(ring) - 1):
        s += fr(ring[i][0]) * fr(ring[i + 1][1]) - fr(ring[i + 1][0]) * fr(ring[i][1])
    return s / 2


def clip_convex(subject, clipper):
    """Sutherland-Hodgman of a (possibly non-convex) subject ring against a convex CCW clipper; exact"""
    def inside(p, a, b):
        return orient(a, b, p) >= 0

    def cross_pt(p, q, a, b):
        px, py, qx, qy = map(fr, (*p, *q))
        ax, ay, bx, by = map(fr, (*a, *b))
        d1 = (bx - ax) * (py - ay) - (by - ay) * (px - ax)
        d2 = (bx - ax) * (qy - ay) - (by - ay) * (qx - ax)
        t = d1 / (d1 - d2)
        return (px + t * (qx - px), py + t * (qy - py))

    out = [(fr(x), fr(y)) for x, y in subject[:-1]]
    clip = [(fr(x), fr(y)) for x, y in clipper]
    for i in range(len(clip) - 1):
        a, b = clip[i], clip[i + 1]
        inp, out = out, []
        if not inp:
            break
        for j in range(len(inp)):
            p, q = inp[j], inp[(j + 1) % len(inp)]
            pin, qin = inside(p, a, b), inside(q, a, b)
            if pin:
                out.append(p)
                if not qin:
                    out.append(cross_pt(p, q, a, b))
            elif qin:
                out.append(cross_pt(p, q, a, b))
    if len(out) < 3:
        return F(0)
    s = F(0)
    for j in range(len(out)):
        p, q = out[j], out[(j + 1) % len(out)]
        s += p[0] * q[1] - q[0] * p[1]
    return s / 2
