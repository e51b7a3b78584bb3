"""CPU: the oracle's piece -> parent aggregation of subdivided layers (engine/engine.cuh:464-470 cut-edge weights,
:963-1042 fold and codes) pinned by subdivision invariance: parent pairs, DE-9IM codes and areas are the same with and
without subdivision, outside the two cases the reference's own rules leave open (see tests/invariance.py), and a
build of the oracle with any one of five planted mistakes in those rules fails the same checks."""
import os
import subprocess
import sys

import numpy as np
import pytest

import invariance as I

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CASES = [(ci, name, maxv) for ci, (name, _, _, mvs) in enumerate(I.cases()) for maxv in mvs]


@pytest.fixture(scope="module")
def cases():
    return I.cases()


@pytest.mark.parametrize("ci,name,maxv", CASES, ids=[f"{n.split()[0]}-{n.split()[1]}-maxv{m}" for _, n, m in CASES])
def test_oracle_is_invariant_under_subdivision(cases, ci, name, maxv):
    _, M, R, _ = cases[ci]
    rep = I.report(I.oracle_join, M, maxv, R=R)
    assert I.failures(rep["one_sided"]) == 0, rep
    assert rep["one_sided"]["rows"] > 0
    if "self" in rep:
        assert I.failures(rep["self"]) == 0, rep
        if maxv >= 256:   # the config-2 setting: neither exception occurs in the sample
            assert rep["self"]["exempt"] == 0 and rep["one_sided"]["exempt"] == 0, rep


def test_exceptions_are_rare_and_shrink_with_piece_size(cases):
    _, M, _, _ = cases[0]
    reps = {m: I.report(I.oracle_join, M, m) for m in (8, 16, 64, 256)}
    exempt = [reps[m]["self"]["exempt"] for m in (8, 16, 64, 256)]
    assert exempt == sorted(exempt, reverse=True) and exempt[-1] == 0
    assert reps[8]["lossy_parents"] > 0 and reps[16]["lossy_parents"] == 0   # slivers only at the deepest cuts
    assert reps[8]["self"]["exempt"] < 0.06 * reps[8]["self"]["rows"]


def test_rectangle_overlaps_equal_closed_form():
    """the rectangles' overlap areas with the square are products of interval lengths, subdivided or not"""
    from oracle import oracle as O
    import gsj

    SQ, RC, exact = I.square_and_rectangles()
    for maxv in (None, 8, 16):
        P = SQ if maxv is None else gsj.subdivide(SQ, maxv)
        r = O.join(O.Layer.polys(P), O.Layer.polys(RC), want_area=True)
        assert len(r["pairs"]) == np.count_nonzero(exact > 0)
        np.testing.assert_allclose(r["area"], exact[r["pairs"][:, 1]], rtol=1e-13, atol=0)


def test_planted_mistakes_fail_the_invariance_checks():
    subprocess.run(["make", "-j5", "-C", os.path.join(ROOT, "oracle"), "mutants"], check=True, capture_output=True)
    procs = {}
    for k in (1, 2, 3, 4, 5):
        env = dict(os.environ, GSJ_ORACLE_LIB=os.path.join(ROOT, "oracle", "build", f"liboracle_m{k}.so"),
                   PYTHONPATH=os.pathsep.join([HERE, os.path.join(ROOT, "postgis-cuda_amd"), ROOT]), OMP_NUM_THREADS="2")
        procs[k] = subprocess.Popen([sys.executable, os.path.join(HERE, "invariance.py")], env=env, stdout=subprocess.PIPE,
                                    stderr=subprocess.PIPE, text=True)
    for k, p in procs.items():
        out, err = p.communicate(timeout=600)
        assert p.returncode == 0, err
        n = int(out.split()[-1])
        assert n > 0, f"planted mistake {k} passed the invariance checks"
