"""Pins the oracle's general gap cost solver (Waterman-Smith-Beyer,
pyalign/algorithm/solver.h:3047-3184) before the GPU is compared against it.

1. The reference's known-answer test (pyalign/tests/test_wsb.py:10-35).
2. Brute-force enumeration of every jump path on tiny integer inputs, with
   cost vectors that are not subadditive (so split gaps must win).
3. Closed forms: linear and affine costs given as general vectors reproduce
   the linear and Gotoh solvers' scores and optimal alignments exactly.
"""

import zlib

import numpy as np
import pytest

import oracle as O
from pyalign_amd import gaps
from tests import brute


def eqsim(a, b, eq, ne):
    return np.array([[eq if x == y else ne for y in b] for x in a], dtype=np.float32)


def general(ws, wt=None):
    ws = np.asarray(ws, dtype=np.float32)
    return ("general", ws, ws if wt is None else np.asarray(wt, dtype=np.float32))


def edge_sets(alignments):
    return sorted(tuple(map(tuple, sorted(e.tolist()))) for _, _, e in alignments)


def test_kat_wsb_cg_ccga():  # pyalign/tests/test_wsb.py:10-35
    S = eqsim("CG", "CCGA", 1, -1)
    lim_s, lim_t = O.rounded_limit(2), O.rounded_limit(4)
    g = gaps.LogarithmicGapCost(3, 1)
    gap = ("general", g.costs(lim_s + 1), g.costs(lim_t + 1))
    al = O.all_alignments(S, locality=O.GLOBAL, gap=gap, lim_s=lim_s, lim_t=lim_t)
    assert all(round(float(s), 1) == -3.7 for s, _, _ in al)
    assert edge_sets(al) == [((0, 0), (1, 1)), ((0, 0), (1, 3))]


def rand_costs(rng, n, allow_split=True):
    """w(0) = 0 and integers w(k) >= 0; with allow_split some long gaps cost
    more than two shorter ones, so the recurrence must chain jumps."""
    w = np.zeros(n, dtype=np.float32)
    w[1:] = rng.integers(0, 6, size=n - 1)
    if not allow_split:
        w[1:] = np.sort(w[1:])
    return w


@pytest.mark.parametrize("maximize", [True, False])
def test_brute_global(maximize):
    rng = np.random.default_rng(zlib.crc32(repr(("wsb-global", maximize)).encode()))
    d = O.MAXIMIZE if maximize else O.MINIMIZE
    for _ in range(14):
        n, m = (int(x) for x in rng.integers(1, 5, size=2))
        S = rng.integers(-3, 4, size=(n, m)).astype(np.float32)
        lim_s, lim_t = O.rounded_limit(n), O.rounded_limit(m)
        ws, wt = rand_costs(rng, lim_s + 1), rand_costs(rng, lim_t + 1)
        opt, sets = brute.brute_wsb_global(S.tolist(), ws.tolist(), wt.tolist(), maximize)
        gap = general(ws, wt)
        r = O.solve(S, locality=O.GLOBAL, direction=d, gap=gap)
        assert r.score == opt
        al = O.all_alignments(S, locality=O.GLOBAL, direction=d, gap=gap)
        assert al and all(s == opt for s, _, _ in al)
        assert edge_sets(al) == sorted(sets)
        assert tuple(map(tuple, sorted(r.edges().tolist()))) in set(sets)


def test_brute_local():
    rng = np.random.default_rng(zlib.crc32(b"wsb-local"))
    for _ in range(10):
        n, m = (int(x) for x in rng.integers(1, 5, size=2))
        S = rng.integers(-3, 4, size=(n, m)).astype(np.float32)
        ws = rand_costs(rng, O.rounded_limit(n) + 1)
        wt = rand_costs(rng, O.rounded_limit(m) + 1)
        assert O.solve(S, locality=O.LOCAL, gap=general(ws, wt)).score == brute.brute_wsb_local(
            S.tolist(), ws.tolist(), wt.tolist())


def test_brute_semiglobal():
    rng = np.random.default_rng(zlib.crc32(b"wsb-semiglobal"))
    for _ in range(10):
        n, m = (int(x) for x in rng.integers(1, 5, size=2))
        S = rng.integers(-3, 4, size=(n, m)).astype(np.float32)
        ws = rand_costs(rng, O.rounded_limit(n) + 1)
        wt = rand_costs(rng, O.rounded_limit(m) + 1)
        assert O.solve(S, locality=O.SEMIGLOBAL, gap=general(ws, wt)).score == brute.brute_wsb_semiglobal(
            S.tolist(), ws.tolist(), wt.tolist())


def test_split_gap_beats_one_long_gap():
    """w(k) grows fast: s = 3 symbols against t = 1 symbol with bad matches.
    The best alignment opens t's gap on the border (w_t(1) = 1) and then
    crosses s in three one-long jumps down the column (3 x w_s(1)), not one
    three-long gap (w_s(3) = 20), and matches nothing: -4."""
    S = np.array([[-9.0], [-9.0], [5.0]], dtype=np.float32)
    w = np.array([0, 1, 10, 20, 30], dtype=np.float32)
    r = O.solve(S, locality=O.GLOBAL, gap=("general", w, w), lim_s=4, lim_t=4)
    assert r.score == -4.0
    assert r.path().tolist() == [[2, 0], [1, 0], [0, 0], [-1, 0]]
    opt, _ = brute.brute_wsb_global(S.tolist(), w.tolist(), w.tolist())
    assert opt == -4.0


@pytest.mark.parametrize("locality", [O.LOCAL, O.GLOBAL, O.SEMIGLOBAL])
@pytest.mark.parametrize("maximize", [True, False])
def test_linear_costs_match_linear_solver(locality, maximize):
    """w(k) = g k: the same optimum, and the same optimal alignments."""
    rng = np.random.default_rng(zlib.crc32(repr(("wsb-lin", locality, maximize)).encode()))
    d = O.MAXIMIZE if maximize else O.MINIMIZE
    for _ in range(10):
        n, m = (int(x) for x in rng.integers(1, 9, size=2))
        S = rng.integers(-4, 5, size=(n, m)).astype(np.float32)
        if not maximize:
            S = -S
        lim_s, lim_t = O.rounded_limit(n), O.rounded_limit(m)
        g = float(rng.integers(0, 3))
        ws = (np.arange(lim_s + 1) * g).astype(np.float32)
        wt = (np.arange(lim_t + 1) * g).astype(np.float32)
        a = O.solve(S, locality=locality, direction=d, gap=general(ws, wt))
        b = O.solve(S, locality=locality, direction=d, gap=("linear", g))
        assert a.score == b.score
        ea = edge_sets(O.all_alignments(S, locality=locality, direction=d, gap=general(ws, wt)))
        eb = edge_sets(O.all_alignments(S, locality=locality, direction=d, gap=("linear", g)))
        assert ea == eb


@pytest.mark.parametrize("locality", [O.LOCAL, O.GLOBAL, O.SEMIGLOBAL])
def test_affine_costs_match_gotoh(locality):
    """w(k) = open + extend (k - 1) (pyalign/gaps.py AffineGapCost.costs):
    the Gotoh solver's optimum."""
    rng = np.random.default_rng(zlib.crc32(repr(("wsb-aff", locality)).encode()))
    for _ in range(10):
        n, m = (int(x) for x in rng.integers(1, 10, size=2))
        S = rng.integers(-4, 6, size=(n, m)).astype(np.float32)
        lim_s, lim_t = O.rounded_limit(n), O.rounded_limit(m)
        open_, ext = float(rng.integers(2, 6)), float(rng.integers(0, 3))
        g = gaps.AffineGapCost(open_, ext)
        a = O.solve(S, locality=locality, gap=("general", g.costs(lim_s + 1), g.costs(lim_t + 1)))
        b = O.solve(S, locality=locality, gap=("affine", open_, ext))
        assert a.score == b.score


def test_solution_matrices_borders():
    """Global borders are w(k) * sgn (the corner from w_t, solver.h:3114-3123);
    local and semiglobal borders are zero."""
    S = eqsim("ACGT", "AGT", 2, -1)
    ws = np.array([0, 3, 4.5, 5, 5.5], dtype=np.float32)
    wt = np.array([0, 2, 2.5, 3, 3.5], dtype=np.float32)
    cfg = O.make_config(locality=O.GLOBAL, gap=("general", ws, wt), lim_s=4, lim_t=4)
    v = O.OracleResult(cfg, S).values()[0]
    assert np.array_equal(v[:, 0], -ws)
    assert np.array_equal(v[0, :], -wt[:4])
    for loc in (O.LOCAL, O.SEMIGLOBAL):
        cfg = O.make_config(locality=loc, gap=("general", ws, wt), lim_s=4, lim_t=4)
        v = O.OracleResult(cfg, S).values()[0]
        assert not v[:, 0].any() and not v[0, :].any()
