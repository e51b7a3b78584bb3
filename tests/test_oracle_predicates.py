"""Pin the oracle's exact predicates (orient, left_of with simulation of simplicity, seg_isect, outward float
rounding) against exact rational arithmetic. Inputs are origin-relative coordinates as the engine sees them: small
multiples of the ulp of a lon/lat coordinate, plus adversarial near-collinear and shared-vertex configurations."""
import math
import random

import numpy as np
import pytest

from oracle import oracle as O
import exact as X

U = math.ulp(95.5)   # spacing of lon/lat coordinates near 95 degrees


def rnd_pt(rng, span=10 ** 9):
    return (rng.randint(-span, span) * U, rng.randint(-span, span) * U)


def test_orient_sign_random():
    rng = random.Random(1)
    for _ in range(20000):
        a, b, c = rnd_pt(rng), rnd_pt(rng), rnd_pt(rng)
        assert np.sign(O.orient(*a, *b, *c)) == X.sign(X.orient(a, b, c))


def test_orient_sign_near_collinear():
    rng = random.Random(2)
    bad = 0
    for _ in range(20000):
        a, b = rnd_pt(rng), rnd_pt(rng)
        t = rng.random()
        # a point on (or one unit beside) the line a-b
        cx = a[0] + t * (b[0] - a[0]) + rng.choice((-1, 0, 1)) * U
        cy = a[1] + t * (b[1] - a[1]) + rng.choice((-1, 0, 1)) * U
        c = (round(cx / U) * U, round(cy / U) * U)
        if np.sign(O.orient(*a, *b, *c)) != X.sign(X.orient(a, b, c)):
            bad += 1
    assert bad == 0


def test_orient_exactly_collinear_is_zero():
    rng = random.Random(3)
    for _ in range(5000):
        a = rnd_pt(rng)
        d = (rng.randint(-1000, 1000) * U, rng.randint(-1000, 1000) * U)
        k1, k2 = rng.randint(-50, 50), rng.randint(-50, 50)
        b = (a[0] + k1 * d[0], a[1] + k1 * d[1])
        c = (a[0] + k2 * d[0], a[1] + k2 * d[1])
        assert X.orient(a, b, c) == 0
        assert O.orient(*a, *b, *c) == 0.0


def test_left_of_simulation_of_simplicity():
    rng = random.Random(4)
    for _ in range(20000):
        a, b = rnd_pt(rng, 1000), rnd_pt(rng, 1000)
        if rng.random() < 0.5:   # p on the line a-b (exact), or anywhere
            k = rng.randint(-3, 3)
            p = (a[0] + k * (b[0] - a[0]), a[1] + k * (b[1] - a[1]))
        else:
            p = rnd_pt(rng, 1000)
        assert O.left_of(*a, *b, *p) == X.left_of_sos(a, b, p)
        if X.orient(a, b, p) == 0 and a != b:
            # consistent for an edge and its reverse: exactly one side claims the point
            assert O.left_of(*a, *b, *p) != O.left_of(*b, *a, *p)


def test_seg_isect_closed_segments():
    rng = random.Random(5)
    for _ in range(20000):
        mode = rng.randrange(4)
        p0, p1 = rnd_pt(rng, 100), rnd_pt(rng, 100)
        if mode == 0:
            q0, q1 = rnd_pt(rng, 100), rnd_pt(rng, 100)
        elif mode == 1:   # shared endpoint
            q0, q1 = p1, rnd_pt(rng, 100)
        elif mode == 2:   # collinear, possibly overlapping
            k0, k1 = rng.randint(-2, 3), rng.randint(-2, 3)
            q0 = (p0[0] + k0 * (p1[0] - p0[0]), p0[1] + k0 * (p1[1] - p0[1]))
            q1 = (p0[0] + k1 * (p1[0] - p0[0]), p0[1] + k1 * (p1[1] - p0[1]))
        else:   # T junction: q0 on p's line
            q0 = (2 * p0[0] - p1[0], 2 * p0[1] - p1[1]) if rng.random() < 0.5 else p0
            q1 = rnd_pt(rng, 100)
        assert O.seg_isect(*p0, *p1, *q0, *q1) == X.segments_intersect(p0, p1, q0, q1)


def test_frame_shift_is_exact():
    """(o_b - o_a) + (v - o_b) == v - o_a for lon/lat-scale coordinates: shared vertices are bit-identical in
    every pair frame (what makes touch tests exact)."""
    rng = random.Random(6)
    for _ in range(20000):
        base = rng.uniform(-120, -60), rng.uniform(20, 50)
        oa = (base[0] + rng.uniform(0, 0.01), base[1] + rng.uniform(0, 0.01))
        ob = (base[0] + rng.uniform(0, 0.01), base[1] + rng.uniform(0, 0.01))
        v = (max(oa[0], ob[0]) + rng.uniform(0, 0.01), max(oa[1], ob[1]) + rng.uniform(0, 0.01))
        for k in (0, 1):
            assert (ob[k] - oa[k]) + (v[k] - ob[k]) == v[k] - oa[k]


@pytest.mark.parametrize("v", [0.0, 1.0, -1.0, 1e-30, -118.5, 34.0000001, 95.123456789, 3.4e38, -7.25e-39, 2.0 ** -149])
def test_outward_float_rounding(v):
    lo, hi = O.f_down(v), O.f_up(v)
    assert lo <= v <= hi                                     # encloses
    assert np.float32(lo) == lo and np.float32(hi) == hi     # floats
    if lo != hi:   # tightest: one float step apart and v strictly between
        assert np.nextafter(np.float32(lo), np.float32(np.inf)) == np.float32(hi)


def test_outward_float_rounding_random():
    rng = np.random.default_rng(7)
    for v in rng.uniform(-180, 180, 20000):
        lo, hi = np.float32(O.f_down(v)), np.float32(O.f_up(v))
        assert float(lo) <= v <= float(hi)
        assert lo == hi or np.nextafter(lo, np.float32(np.inf)) == hi
