"""Pins the oracle's Solution matrices (values, traceback matrix and edge
lists, pyalign/algorithm/solver.h:2355-2471) on tiny integer inputs.

* D values: every cell's value is the brute-force optimum of the sub-problem
  that ends at that cell (tests/brute.py, no dynamic programming).
* P / Q values (affine): the closed form "best gap run ending here" over the
  brute-force D values, P(i, j) = opt_k D(k-1, j) - w(i-k+1), which the
  Gotoh recursion computes incrementally.
* Tracebacks: each cell's predecessors are exactly the candidates that
  attain its value, in the reference's push order ([local zero], diagonal,
  P / up, Q / left; solver.h:2841-2852, 2993-3033), gap origins nearest
  first; one-optimal tracebacks keep the first of them.
Integer-valued inputs keep every sum exact in any order.
"""

import zlib

import numpy as np
import pytest

import oracle as O
from tests import brute

NO_TB = -32768  # no_traceback<int16_t>() (solver.h:182-185)


def rand_case(rng, n, m):
    return rng.integers(-3, 4, size=(n, m)).astype(np.float32)


def run_w(gap, k):
    return brute.run_cost(gap, k)


def brute_D(S, gap, maximize, locality):
    """D(i, j) for every cell from enumeration of the sub-problems."""
    n, m = S.shape
    D = np.zeros((n, m))
    for i in range(n):
        for j in range(m):
            cands = []
            # a start (a, b) leaves s[:a] and t[:b] before the path; a == i + 1
            # or b == j + 1 leaves one side empty (the path is one gap run)
            if locality == O.GLOBAL:
                starts = [(0, 0)]
            elif locality == O.SEMIGLOBAL:
                starts = [(a, 0) for a in range(i + 2)] + [(0, b) for b in range(1, j + 2)]
            else:
                starts = [(a, b) for a in range(i + 2) for b in range(j + 2)]
            for a, b in starts:
                rows, cols = i + 1 - a, j + 1 - b
                if rows == 0 and cols == 0:
                    continue
                if rows == 0 or cols == 0:
                    cands.append((-1.0 if maximize else 1.0) * run_w(gap, rows + cols))
                    continue
                sub = S[a:i + 1, b:j + 1].tolist()
                cands.append(brute.brute_global(sub, gap, maximize)[0])
            if locality == O.LOCAL:
                cands.append(0.0)
            D[i, j] = brute.best(cands, maximize)
    return D


def border(locality, gap, maximize, k):
    """D(k-1, -1) = D(-1, k-1): w(k) * sgn for global, 0 otherwise."""
    if locality != O.GLOBAL or k == 0:
        return 0.0
    return (-1.0 if maximize else 1.0) * run_w(gap, k)


def full_D(Dint, locality, gap, maximize):
    """D with the border row/column, indexed [i + 1][j + 1]."""
    n, m = Dint.shape
    F = np.zeros((n + 1, m + 1))
    F[1:, 1:] = Dint
    for i in range(n + 1):
        F[i, 0] = border(locality, gap, maximize, i)
    for j in range(m + 1):
        F[0, j] = border(locality, gap, maximize, j)
    return F


def gap_closed_form(F, gap, maximize, vertical):
    """P (vertical) or Q: best over gap origins of D(origin) - w(run); also
    the attaining origins, nearest first."""
    n, m = F.shape[0] - 1, F.shape[1] - 1
    sgn = -1.0 if maximize else 1.0
    V = np.zeros((n, m))
    origins = {}
    for i in range(n):
        for j in range(m):
            c = []
            if vertical:
                for k in range(i, -1, -1):  # gap covers rows k..i, origin (k-1, j)
                    c.append((F[k, j + 1] + sgn * run_w(gap, i - k + 1), (k - 1, j)))
            else:
                for k in range(j, -1, -1):
                    c.append((F[i + 1, k] + sgn * run_w(gap, j - k + 1), (i, k - 1)))
            v = brute.best([x for x, _ in c], maximize)
            V[i, j] = v
            origins[(i, j)] = [o for x, o in c if x == v]
    return V, origins


def expected_lists(S, F, gap, maximize, locality):
    """Per layer and interior cell: the traceback_n entries in push order."""
    n, m = S.shape
    sgn = -1.0 if maximize else 1.0
    lists = [dict()]
    if gap[0] == "affine":
        P, po = gap_closed_form(F, gap, maximize, True)
        Q, qo = gap_closed_form(F, gap, maximize, False)
        lists += [po, qo]
    for i in range(n):
        for j in range(m):
            d = F[i + 1, j + 1]
            c = []
            if locality == O.LOCAL:
                c.append((0.0, [(NO_TB, NO_TB)]))
            c.append((F[i, j] + S[i, j], [(i - 1, j - 1)]))
            if gap[0] == "affine":
                c.append((P[i, j], po[(i, j)]))
                c.append((Q[i, j], qo[(i, j)]))
            else:
                c.append((F[i, j + 1] + sgn * gap[1], [(i - 1, j)]))
                c.append((F[i + 1, j] + sgn * gap[1], [(i, j - 1)]))
            lists[0][(i, j)] = [p for x, pts in c if x == d for p in pts]
    return lists


CASES = [(loc, gap, mx) for loc in (O.GLOBAL, O.LOCAL, O.SEMIGLOBAL)
         for gap in (("linear", 1), ("linear", 0), ("affine", 3, 1), ("affine", 2, 0))
         for mx in (True, False)]


@pytest.mark.parametrize("locality,gap,maximize", CASES)
def test_values_and_tracebacks_against_enumeration(locality, gap, maximize):
    rng = np.random.default_rng(zlib.crc32(repr((locality, gap, maximize)).encode()))
    d = O.MAXIMIZE if maximize else O.MINIMIZE
    for _ in range(6):
        n, m = (int(x) for x in rng.integers(1, 5, size=2))
        S = rand_case(rng, n, m)
        F = full_D(brute_D(S, gap, maximize, locality), locality, gap, maximize)
        lists = expected_lists(S, F, gap, maximize, locality)
        for goal in (O.GOAL_ONE, O.GOAL_ALL):
            r = O.OracleResult(O.make_config(locality=locality, direction=d, gap=gap, goal=goal,
                                             lim_s=8, lim_t=8), S)
            vals = r.values()
            assert vals.shape == (3 if gap[0] == "affine" else 1, n + 1, m + 1)
            assert np.array_equal(vals[0], F)
            tbm = r.traceback_matrix()
            edges = r.traceback_edges()
            if gap[0] == "affine":
                P, _ = gap_closed_form(F, gap, maximize, True)
                Q, _ = gap_closed_form(F, gap, maximize, False)
                assert np.array_equal(vals[1, 1:, 1:], P)
                assert np.array_equal(vals[2, 1:, 1:], Q)
                worst = -np.inf if maximize else np.inf
                assert np.all(vals[1, 0, 1:] == worst) and np.all(vals[2, 1:, 0] == worst)
                assert np.array_equal(vals[1, :, 0], F[:, 0]) and np.array_equal(vals[2, 0, :], F[0, :])
            for layer, cells in enumerate(lists):
                src = [tuple(e[0]) for e in edges[layer].tolist()]
                assert src == sorted(src)  # cells row by row
                got = {}
                for (fu, fv), (u, v) in edges[layer].tolist():
                    got.setdefault((fu, fv), []).append((u, v))
                for (i, j), want in cells.items():
                    if goal == O.GOAL_ALL:
                        assert got.get((i, j), []) == want, (layer, i, j)
                        first = want[0] if want else (NO_TB, NO_TB)
                    else:
                        first = want[0]
                        assert got[(i, j)] == [first], (layer, i, j)
                    assert tuple(tbm[layer, i + 1, j + 1]) == first
                # border cells: cleared tracebacks
                assert np.all(tbm[layer, 0, :] == NO_TB) and np.all(tbm[layer, :, 0] == NO_TB)
                n_border = (n + 1) * (m + 1) - n * m
                if goal == O.GOAL_ONE:
                    assert len(edges[layer]) == (n + 1) * (m + 1)
                    b = [e for e in edges[layer].tolist() if e[0][0] < 0 or e[0][1] < 0]
                    assert len(b) == n_border and all(e[1] == [NO_TB, NO_TB] for e in b)
                else:
                    assert all(e[0][0] >= 0 and e[0][1] >= 0 for e in edges[layer].tolist())


@pytest.mark.parametrize("maximize", [False, True])
def test_dtw_values_and_tracebacks(maximize):
    rng = np.random.default_rng(91)
    d = O.MAXIMIZE if maximize else O.MINIMIZE
    worst = -np.inf if maximize else np.inf
    for _ in range(8):
        n, m = (int(x) for x in rng.integers(1, 5, size=2))
        C = rng.integers(0, 6, size=(n, m)).astype(np.float32)
        r = O.OracleResult(O.make_config(solver_type=O.DTW, direction=d, goal=O.GOAL_ONE, lim_s=8, lim_t=8), C)
        vals = r.values()[0]
        F = np.full((n + 1, m + 1), worst)
        F[0, 0] = 0.0
        for i in range(n):
            for j in range(m):
                F[i + 1, j + 1] = brute.brute_dtw(C[:i + 1, :j + 1].tolist(), maximize)
        assert np.array_equal(vals, F)
        tbm = r.traceback_matrix()[0]
        for i in range(n):
            for j in range(m):
                c = [(F[i, j], (i - 1, j - 1)), (F[i, j + 1], (i - 1, j)), (F[i + 1, j], (i, j - 1))]
                opt = brute.best([x for x, _ in c], maximize)
                assert tuple(tbm[i + 1, j + 1]) == next(p for x, p in c if x == opt)
