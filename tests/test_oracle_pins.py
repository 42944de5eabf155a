"""Pins for the oracle parts that the reference's KATs leave open (CPU only).

1. Global affine border vectors at solver limits >= 1024 that are not powers
   of two.  AffineCost::vector is xt::linspace<float>(0, (n-1)*u, n) + v with
   w(0) = 0 (pyalign/algorithm/solver.h:2876-2880).  When the extend u is
   dyadic (1, 1/2, 1/4, 1/8) every k*u up to the limits here is exact in
   float32 and in double, so step = (n-1)*u / (n-1) = u exactly and every
   way of evaluating linspace (start + k*step, a lerp, the exact stop at
   k = n-1) gives k*u: w(k) = fl32(k*u + v), whatever xtensor's generator
   order is.  The mathematics fixes these values independently of n.
2. The emission order of all optimal alignments (solver.h:1648-1716: a LIFO
   stack, a cell's traceback candidates pushed in candidate order so the
   last one is explored first; local seeds pushed i, j descending,
   solver.h:2067-2091), traced by hand on the reference KATs that have two
   optimal alignments; the expected lists below are written out from that
   trace, not produced by running anything.
3. The border cells of Solution.values (solver.h:2918-2951 for Gotoh,
   2806-2814 linear, 2021-2029 / 2253-2261 local and semiglobal zero
   borders): the constructor fills Q's column 0 and P's row 0 with the worst
   value, then D's and P's column 0 with w_s(i)*sgn, then D's and Q's row 0
   with w_t(j)*sgn -- so P(0, 0) = Q(0, 0) = D(0, 0) = w(0)*sgn, which is
   -0.0 when maximising (sgn = -1) and +0.0 when minimising; local and
   semiglobal fill +0 where global fills w*sgn.
"""

import numpy as np
import pytest

import oracle as O


def _w(k, open_, extend):
    """w(k) for k >= 1 of AffineGapCost(open, extend) after the factory's
    flip: u = fl32(extend), v = fl32(open - extend) (pyalign/gaps.py:164-176,
    factory.h:48-52); k*u is exact for dyadic u."""
    u = np.float32(extend)
    v = np.float32(float(open_) - float(extend))
    ku = np.float32(k) * u
    assert float(ku) == k * float(u)  # exact
    return np.float32(ku + v)


@pytest.mark.parametrize("lim", [1100, 1537, 5045])
@pytest.mark.parametrize("extend", [1.0, 0.5, 0.25, 0.125])
@pytest.mark.parametrize("direction", [O.MAXIMIZE, O.MINIMIZE])
def test_affine_border_vector_dyadic_extend_non_power_of_two_limit(lim, extend, direction):
    open_ = 2.0 + 3 * extend
    rng = np.random.default_rng(lim)
    # len_s = lim: the D / P column shows w_s(k) for every k = 0 .. lim,
    # including the linspace endpoint k = n - 1 = lim
    S = rng.integers(-2, 3, size=(lim, 2)).astype(np.float32)
    r = O.OracleResult(O.make_config(locality=O.GLOBAL, direction=direction, gap=("affine", open_, extend),
                                     goal=O.GOAL_ONE, lim_s=lim, lim_t=1030), S)
    vals = r.values()
    sgn = np.float32(-1.0 if direction == O.MAXIMIZE else 1.0)
    want = np.array([np.float32(0.0) * sgn] + [_w(k, open_, extend) * sgn for k in range(1, lim + 1)], np.float32)
    for layer in (0, 1):  # D and P: column 0 holds w_s(i) * sgn
        assert np.array_equal(vals[layer][:, 0].view(np.uint32), want.view(np.uint32)), layer
    wt = np.array([np.float32(0.0) * sgn] + [_w(k, open_, extend) * sgn for k in (1, 2)], np.float32)
    for layer in (0, 2):  # D and Q: row 0 holds w_t(j) * sgn (t limit 1030)
        assert np.array_equal(vals[layer][0, :].view(np.uint32), wt.view(np.uint32)), layer


def _edge_lists(r):
    out = []
    while True:
        x = r.all_next()
        if x is None:
            return out
        out.append(tuple(sorted(tuple(map(int, e)) for e in x[2].tolist())))


def _dense(a, b, eq, ne):
    return np.array([[eq if x == y else ne for y in b] for x in a], np.float32)


def test_all_optimal_order_sw_kat():
    """test_sw.py:10-31: AATCG / AACG, local, linear gap 2, eq 1 ne -1.

    Hand trace: the maximum 2 sits at (1, 1) and (4, 3).  Seeds are pushed
    scanning i, j descending, so (4, 3) goes in first and (1, 1) is on top.
    (1, 1) <- (0, 0) [diag, 1 + 1] <- (-1, -1) [diag, 0 + 1]: path cells
    (1,1), (0,0) are left diagonally: edges (0,0), (1,1).  Then (4, 3) <-
    (3, 2) <- (2, 1), where D(2, 1) = 0 stops the local walk: edges (3,2),
    (4,3)."""
    S = _dense("AATCG", "AACG", 1, -1)
    r = O.solve(S, locality=O.LOCAL, gap=("linear", 2), goal=O.GOAL_ALL)
    assert _edge_lists(r) == [((0, 0), (1, 1)), ((3, 2), (4, 3))]


def test_all_optimal_order_batch_kat():
    """test_batch.py:10-41, third problem: AATCG / AGTT, global, linear 2.

    Hand trace: D(4, 3) = -2 holds [diag (3, 2), up (3, 3)]; both are pushed,
    (3, 3) last, so it is explored first: (3,3) <- (2,2) <- (1,1) <- (0,0)
    <- (-1,-1) (the move (4,3) -> (3,3) changes only u, so (4, 3) is no
    edge): edges (0,0) (1,1) (2,2) (3,3).  Then (3, 2): its only candidate
    is up (2, 2), then the same diagonal run: edges (0,0) (1,1) (2,2) (4,3)."""
    S = _dense("AATCG", "AGTT", 1, -1)
    r = O.solve(S, locality=O.GLOBAL, gap=("linear", 2), goal=O.GOAL_ALL)
    assert _edge_lists(r) == [((0, 0), (1, 1), (2, 2), (3, 3)), ((0, 0), (1, 1), (2, 2), (4, 3))]


def test_all_optimal_order_gotoh_kat():
    """test_gotoh.py:13-32: CC / ACCT, minimize, eq 0 ne 1, affine (5, 1).

    Hand trace (w1 = 5, extension 1): D(1, 3) = 7 ties diag D(0, 2) + 1 and
    Q(1, 3) = Q(1, 2) + 1, whose traceback is copied from the run's origin
    (1, 1) (Q(1, 2) = D(1, 1) + 5).  Pushed [(0, 2), (1, 1)], so (1, 1) is
    explored first: (1,1) <- (0,0) <- (-1,-1): edges (0,0) (1,1).  Then
    (0, 2) <- (-1, 1): edges (0,2) (1,3)."""
    S = _dense("CC", "ACCT", 0, 1)
    r = O.solve(S, locality=O.GLOBAL, direction=O.MINIMIZE, gap=("affine", 5, 1), goal=O.GOAL_ALL)
    got = []
    while True:
        x = r.all_next()
        if x is None:
            break
        assert x[0] == np.float32(7)
        got.append(tuple(sorted(tuple(map(int, e)) for e in x[2].tolist())))
    assert got == [((0, 0), (1, 1)), ((0, 2), (1, 3))]


def _lin_border(k, g, sgn):
    # xt::arange<index>(0, n) * gap * sgn (solver.h:2806-2814): multiplied, not accumulated
    return np.float32(np.float32(np.float32(k) * np.float32(g)) * sgn)


@pytest.mark.parametrize("direction", [O.MAXIMIZE, O.MINIMIZE])
@pytest.mark.parametrize("loc", [O.GLOBAL, O.LOCAL, O.SEMIGLOBAL])
def test_solution_border_cells_affine(direction, loc):
    rng = np.random.default_rng(7 + loc + 3 * direction)
    ls, lt, lim_s, lim_t = 5, 7, 8, 16
    S = rng.integers(-3, 4, size=(ls, lt)).astype(np.float32)
    open_, extend = 3.0, 0.5
    r = O.OracleResult(O.make_config(locality=loc, direction=direction, gap=("affine", open_, extend),
                                     goal=O.GOAL_ONE, lim_s=lim_s, lim_t=lim_t), S)
    D, P, Q = r.values()
    sgn = np.float32(-1.0 if direction == O.MAXIMIZE else 1.0)
    worst = np.float32(-np.inf if direction == O.MAXIMIZE else np.inf)
    if loc == O.GLOBAL:
        col = np.array([np.float32(0.0) * sgn] + [_w(k, open_, extend) * sgn for k in range(1, ls + 1)], np.float32)
        row = np.array([np.float32(0.0) * sgn] + [_w(k, open_, extend) * sgn for k in range(1, lt + 1)], np.float32)
    else:
        col = np.zeros(ls + 1, np.float32)  # +0.0 everywhere
        row = np.zeros(lt + 1, np.float32)
    bits = lambda x: np.asarray(x, np.float32).view(np.uint32)  # noqa: E731  (keeps the sign of zero)
    assert np.array_equal(bits(D[:, 0]), bits(col))
    assert np.array_equal(bits(D[0, :]), bits(row))
    assert np.array_equal(bits(P[:, 0]), bits(col))          # P's column 0, P(0, 0) included
    assert np.array_equal(bits(P[0, 1:]), bits(np.full(lt, worst)))  # P's row 0 beyond it: worst
    assert np.array_equal(bits(Q[0, :]), bits(row))          # Q's row 0, Q(0, 0) included
    assert np.array_equal(bits(Q[1:, 0]), bits(np.full(ls, worst)))  # Q's column 0 below it: worst
    if loc == O.GLOBAL:
        assert np.signbit(D[0, 0]) == (direction == O.MAXIMIZE)


@pytest.mark.parametrize("direction", [O.MAXIMIZE, O.MINIMIZE])
@pytest.mark.parametrize("loc", [O.GLOBAL, O.LOCAL, O.SEMIGLOBAL])
def test_solution_border_cells_linear_and_dtw(direction, loc):
    rng = np.random.default_rng(17 + loc + 3 * direction)
    ls, lt = 6, 4
    S = rng.integers(-3, 4, size=(ls, lt)).astype(np.float32)
    g = 0.7
    r = O.OracleResult(O.make_config(locality=loc, direction=direction, gap=("linear", g), goal=O.GOAL_ONE,
                                     lim_s=8, lim_t=4), S)
    (D,) = r.values()
    sgn = np.float32(-1.0 if direction == O.MAXIMIZE else 1.0)
    bits = lambda x: np.asarray(x, np.float32).view(np.uint32)  # noqa: E731
    if loc == O.GLOBAL:
        assert np.array_equal(bits(D[:, 0]), bits([_lin_border(k, g, sgn) for k in range(ls + 1)]))
        assert np.array_equal(bits(D[0, :]), bits([_lin_border(k, g, sgn) for k in range(lt + 1)]))
    else:
        assert np.array_equal(bits(D[:, 0]), bits(np.zeros(ls + 1)))
        assert np.array_equal(bits(D[0, :]), bits(np.zeros(lt + 1)))
    # DTW (solver.h:3205-3210): every border cell worst, D(0, 0) = +0
    r = O.OracleResult(O.make_config(solver_type=O.DTW, direction=direction, goal=O.GOAL_ONE, lim_s=8, lim_t=4),
                       np.abs(S))
    (D,) = r.values()
    worst = np.float32(-np.inf if direction == O.MAXIMIZE else np.inf)
    assert np.array_equal(bits(D[1:, 0]), bits(np.full(ls, worst)))
    assert np.array_equal(bits(D[0, 1:]), bits(np.full(lt, worst)))
    assert bits(D[0, 0]) == bits(np.float32(0.0))
