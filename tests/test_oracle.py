"""Pins the CPU oracle before anything is compared against it.

1. The reference's own known-answer tests (pyalign/tests/*.py and the
   README example, README.md:94-110), re-run through the oracle.
2. Brute-force enumeration of every alignment on tiny integer inputs
   (tests/brute.py), for all localities, both gap models and directions.
3. Closed forms that a dropped term, a wrong sign or a transposed operand
   would break.
"""

import zlib

import numpy as np
import pytest

import oracle as O
from tests import brute


def eqsim(a, b, eq, ne):
    return np.array([[eq if x == y else ne for y in b] for x in a], dtype=np.float32)


def dict_sim(pairs, default=0):
    def f(x, y):  # pyalign/problems/function.py:29-44 (symmetric lookup)
        a = pairs.get((x, y))
        b = pairs.get((y, x))
        if a is None and b is not None:
            return b
        if a is not None:
            return a
        return default
    return f


def edge_sets(alignments):
    return sorted(tuple(map(tuple, sorted(e.tolist()))) for _, _, e in alignments)


# ---------------------------------------------------------------- KATs ---

def test_kat_nw_aatcg_aacg():  # pyalign/tests/test_nw.py:9-29
    r = O.solve(eqsim("AATCG", "AACG", 1, -1), locality=O.GLOBAL, gap=("linear", 2))
    assert r.score == 2.0
    assert sorted(r.edges().tolist()) == [[0, 0], [1, 1], [3, 2], [4, 3]]


def test_kat_sw_aatcg_aacg():  # pyalign/tests/test_sw.py:10-31
    al = O.all_alignments(eqsim("AATCG", "AACG", 1, -1), locality=O.LOCAL, gap=("linear", 2))
    assert all(s == 2 for s, _, _ in al)
    assert edge_sets(al) == [((0, 0), (1, 1)), ((3, 2), (4, 3))]


@pytest.mark.parametrize("a,b,score,expected", [  # pyalign/tests/test_batch.py:10-41
    ("AATCG", "AACG", 2, [((0, 0), (1, 1), (3, 2), (4, 3))]),
    ("AATGC", "AACG", 0, [((0, 0), (1, 1), (2, 2), (3, 3))]),
    ("AATCG", "AGTT", -2, [((0, 0), (1, 1), (2, 2), (3, 3)), ((0, 0), (1, 1), (2, 2), (4, 3))]),
])
def test_kat_batch(a, b, score, expected):
    # the reference batches all three into one SIMD call with master shape (5, 4)
    al = O.all_alignments(eqsim(a, b, 1, -1), locality=O.GLOBAL, gap=("linear", 2), lim_s=8, lim_t=4)
    assert all(s == score for s, _, _ in al)
    assert edge_sets(al) == sorted(expected)


def test_kat_gotoh_cc_acct_minimize():  # pyalign/tests/test_gotoh.py:13-32
    al = O.all_alignments(eqsim("CC", "ACCT", 0, 1), locality=O.GLOBAL,
                          direction=O.MINIMIZE, gap=("affine", 5, 1))
    assert all(s == 7.0 for s, _, _ in al)
    assert edge_sets(al) == [((0, 0), (1, 1)), ((0, 2), (1, 3))]


def test_kat_gotoh_ggtgtga_tcgcgt():  # pyalign/tests/test_gotoh.py:38-84
    f = dict_sim({("T", "C"): -1, ("A", "T"): 0, ("G", "T"): -2,
                  ("A", "A"): 6, ("T", "T"): 6, ("G", "G"): 6})
    a, b = "GGTGTGA", "TCGCGT"
    S = np.array([[f(x, y) for y in b] for x in a], dtype=np.float32)
    al = O.all_alignments(S, locality=O.GLOBAL, gap=("affine", 11, 1))
    assert all(s == -2 for s, _, _ in al)
    assert edge_sets(al) == [((0, 0), (1, 1), (3, 2), (4, 3), (5, 4), (6, 5))]


def test_kat_gotoh_aaaggg():  # pyalign/tests/test_gotoh.py:86-119
    al = O.all_alignments(eqsim("AAAGGG", "TTAAAAGGGGTT", 0, -1), locality=O.GLOBAL,
                          gap=("affine", 5, 1))
    assert all(s == -14 for s, _, _ in al)
    assert edge_sets(al) == sorted([
        ((0, 3), (1, 4), (2, 5), (3, 6), (4, 7), (5, 8)),
        ((0, 0), (1, 1), (2, 2), (3, 9), (4, 10), (5, 11))])


def test_kat_gotoh_taaatttgc():  # pyalign/tests/test_gotoh.py:121-150
    al = O.all_alignments(eqsim("TAAATTTGC", "TCGCCTTAC", 10, -30), locality=O.GLOBAL,
                          gap=("affine", 40, 1))
    assert all(s == -60 for s, _, _ in al)
    assert edge_sets(al) == sorted([((0, 0), (1, 7), (8, 8)), ((0, 6), (1, 7), (8, 8))])


@pytest.mark.parametrize("open_cost,score", [(25, -44), (30, -54)])
def test_kat_gotoh_agat_ctct(open_cost, score):  # pyalign/tests/test_gotoh.py:152-176
    al = O.all_alignments(eqsim("AGAT", "CTCT", 10, -30), locality=O.GLOBAL,
                          gap=("affine", open_cost, 1))
    assert all(s == score for s, _, _ in al)
    assert edge_sets(al) == [((3, 3),)]


def test_kat_readme_industry_interest():  # README.md:94-110
    r = O.solve(eqsim("INDUSTRY", "INTEREST", 1, -1), locality=O.GLOBAL, gap=("linear", 0))
    assert r.score == 4.0
    assert sorted(r.edges().tolist()) == [[0, 0], [1, 1], [4, 6], [5, 7]]


# -------------------------------------------------------- brute force ---

def rand_case(rng, n, m, lo=-3, hi=4):
    return rng.integers(lo, hi, size=(n, m)).astype(np.float32)


GAPS = [("linear", 0), ("linear", 1), ("linear", 2), ("affine", 3, 1), ("affine", 2, 2), ("affine", 5, 0)]


@pytest.mark.parametrize("gap", GAPS)
@pytest.mark.parametrize("maximize", [True, False])
def test_brute_global(gap, maximize):
    rng = np.random.default_rng(zlib.crc32(repr((gap, maximize)).encode()))
    for _ in range(12):
        n, m = rng.integers(1, 5, size=2)
        S = rand_case(rng, n, m)
        opt, sets = brute.brute_global(S.tolist(), gap, maximize)
        d = O.MAXIMIZE if maximize else O.MINIMIZE
        r = O.solve(S, locality=O.GLOBAL, direction=d, gap=gap)
        assert r.score == opt
        al = O.all_alignments(S, locality=O.GLOBAL, direction=d, gap=gap)
        assert all(s == opt for s, _, _ in al)
        assert edge_sets(al) == sorted(sets)
        # the one-optimal path is one of the optimal alignments
        assert tuple(map(tuple, sorted(r.edges().tolist()))) in set(sets)


@pytest.mark.parametrize("gap", GAPS)
def test_brute_local(gap):
    rng = np.random.default_rng(7 + len(str(gap)))
    for _ in range(8):
        n, m = rng.integers(1, 5, size=2)
        S = rand_case(rng, n, m)
        opt = brute.brute_local(S.tolist(), gap, True)
        assert O.solve(S, locality=O.LOCAL, gap=gap).score == opt


@pytest.mark.parametrize("gap", GAPS)
def test_brute_semiglobal(gap):
    rng = np.random.default_rng(11 + len(str(gap)))
    for _ in range(8):
        n, m = rng.integers(1, 5, size=2)
        S = rand_case(rng, n, m)
        opt = brute.brute_semiglobal(S.tolist(), gap, True)
        assert O.solve(S, locality=O.SEMIGLOBAL, gap=gap).score == opt


@pytest.mark.parametrize("maximize", [False, True])
def test_brute_dtw(maximize):
    rng = np.random.default_rng(5)
    for _ in range(10):
        n, m = rng.integers(1, 6, size=2)
        C = rng.integers(0, 9, size=(n, m)).astype(np.float32)
        d = O.MAXIMIZE if maximize else O.MINIMIZE
        r = O.solve(C, solver_type=O.DTW, direction=d)
        assert r.score == brute.brute_dtw(C.tolist(), maximize)


def path_score(path, S, gap, maximize, locality):
    """Re-score a reference-format path (seed first) from the gap model."""
    sgn = -1.0 if maximize else 1.0
    total = 0.0
    for (u, v), (pu, pv) in zip(path[:-1], path[1:]):
        if u - pu == 1 and v - pv == 1:
            total += S[u][v]
        elif v == pv:
            total += sgn * brute.run_cost(gap, u - pu)
        else:
            total += sgn * brute.run_cost(gap, v - pv)
    if locality == O.GLOBAL:
        lu, lv = path[-1]
        if lu >= 0 and lv < 0:
            total += sgn * brute.run_cost(gap, lu + 1)
        elif lv >= 0 and lu < 0:
            total += sgn * brute.run_cost(gap, lv + 1)
    return total


@pytest.mark.parametrize("locality", [O.LOCAL, O.GLOBAL, O.SEMIGLOBAL])
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 4, 1)])
def test_path_rescoring(locality, gap):
    rng = np.random.default_rng(3)
    for _ in range(20):
        n, m = rng.integers(1, 12, size=2)
        S = rand_case(rng, n, m, -4, 6)
        r = O.solve(S, locality=locality, gap=gap)
        assert path_score(r.path().tolist(), S.tolist(), gap, True, locality) == r.score
        for s, p, e in O.all_alignments(S, locality=locality, gap=gap):
            assert path_score(p.tolist(), S.tolist(), gap, True, locality) == s


# -------------------------------------------------------- closed forms ---

def test_identical_global_is_n_eq():
    s = "ACGTTGCAAGT"
    for gap in [("linear", 1), ("affine", 3, 1)]:
        assert O.solve(eqsim(s, s, 2, -1), locality=O.GLOBAL, gap=gap).score == 2 * len(s)


def test_local_of_substring():
    rng = np.random.default_rng(9)
    t = "".join(rng.choice(list("ACGT"), 40))
    s = t[11:23]
    r = O.solve(eqsim(s, t, 1, -3), locality=O.LOCAL, gap=("affine", 5, 2))
    assert r.score == len(s)
    assert sorted(r.edges().tolist()) == [[i, 11 + i] for i in range(len(s))]


def test_semiglobal_embedded():
    rng = np.random.default_rng(10)
    t = "".join(rng.choice(list("ACGT"), 50))
    s = t[17:35]
    r = O.solve(eqsim(s, t, 1, -1), locality=O.SEMIGLOBAL, gap=("affine", 2, 0.5))
    assert r.score == len(s)


def test_affine_open_equal_extend_is_linear():
    rng = np.random.default_rng(12)
    for _ in range(10):
        n, m = rng.integers(1, 30, size=2)
        S = rand_case(rng, n, m)
        for loc in (O.LOCAL, O.GLOBAL, O.SEMIGLOBAL):
            a = O.solve(S, locality=loc, gap=("affine", 2, 2)).score
            b = O.solve(S, locality=loc, gap=("linear", 2)).score
            assert a == b


def test_minimize_negated_is_negated_maximize():
    rng = np.random.default_rng(13)
    for _ in range(10):
        n, m = rng.integers(1, 30, size=2)
        S = rand_case(rng, n, m)
        for gap in (("linear", 1), ("affine", 3, 1)):
            a = O.solve(S, locality=O.GLOBAL, gap=gap).score
            # minimizing -S with the same (positive) gap costs mirrors maximizing S
            c = O.solve(-S, locality=O.GLOBAL, direction=O.MINIMIZE, gap=gap).score
            assert c == -a


def test_dtw_identical_is_zero_on_diagonal():
    x = np.cumsum(np.random.default_rng(14).normal(size=(20, 3)), axis=0).astype(np.float32)
    C = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1).astype(np.float32)
    r = O.solve(C, solver_type=O.DTW, direction=O.MINIMIZE)
    assert r.score == 0
    assert r.path().tolist() == [[i, i] for i in range(19, -2, -1)]


def test_global_affine_border_power_of_two():
    # w(k) = fl32(fl32(k*u) + v) for power-of-two limits (solver.h:2876-2880)
    S = np.zeros((3, 5), np.float32)
    r = O.solve(S, locality=O.GLOBAL, gap=("affine", 1.0, 0.2), lim_s=4, lim_t=8)
    vals = r.values()
    u, v = np.float32(0.2), np.float32(np.float64(1.0) - np.float64(0.2))
    for k in range(1, 4):
        assert vals[0, k, 0] == -np.float32(np.float32(k * u) + v)
    for k in range(1, 6):
        assert vals[0, 0, k] == -np.float32(np.float32(k * u) + v)
    assert np.signbit(vals[0, 0, 0])  # -0.0 when maximizing
    assert np.isneginf(vals[1, 0, 1]) and np.isneginf(vals[2, 1, 0])


def test_empty_inputs():
    # global: the seed is a border cell (solver.h:2179-2187)
    r = O.solve(np.zeros((0, 3), np.float32), locality=O.GLOBAL, gap=("linear", 2))
    assert r.score == -6.0 and r.path().tolist() == [[-1, 2]]
    # local with an empty side has no seed: score is the worst value
    r = O.solve(np.zeros((0, 3), np.float32), locality=O.LOCAL, gap=("linear", 2))
    assert not r.has_seed and np.isneginf(r.score)
    r = O.solve(np.zeros((2, 0), np.float32), locality=O.SEMIGLOBAL, gap=("linear", 2))
    assert r.score == 0.0


def test_batched_lanes_match_single():
    """16-lane baseline (reference SIMD layout) equals one-lane solves (H9)."""
    rng = np.random.default_rng(15)
    K = 5
    sim = rng.integers(-3, 5, size=(K, K)).astype(np.float32)
    sim = np.triu(sim) + np.triu(sim, 1).T
    n = 40
    ls = rng.integers(0, 30, n).astype(np.int64)
    lt = rng.integers(0, 30, n).astype(np.int64)
    cs = [rng.integers(0, K, x).astype(np.uint8) for x in ls]
    ct = [rng.integers(0, K, x).astype(np.uint8) for x in lt]
    off_s = np.concatenate([[0], np.cumsum(ls)[:-1]]).astype(np.int64)
    off_t = np.concatenate([[0], np.cumsum(lt)[:-1]]).astype(np.int64)
    codes_s = np.concatenate(cs + [np.zeros(1, np.uint8)])
    codes_t = np.concatenate(ct + [np.zeros(1, np.uint8)])
    members = np.arange(n, dtype=np.int32)
    boff = np.arange(0, n + 16, 16).clip(max=n).astype(np.int64)
    boff = np.unique(boff)
    for loc, gap in [(O.LOCAL, ("affine", 4, 1)), (O.GLOBAL, ("linear", 1)), (O.SEMIGLOBAL, ("affine", 3, 2))]:
        cfg = O.make_config(locality=loc, goal=O.GOAL_SCORE, gap=gap, lim_s=32, lim_t=32)
        _, scores = O.baseline_indexed_score(cfg, codes_s, off_s, ls, codes_t, off_t, lt, sim,
                                             members, boff, 2)
        for k in range(n):
            S = sim[np.ix_(cs[k], ct[k])]
            r = O.OracleResult(O.make_config(locality=loc, goal=O.GOAL_SCORE, gap=gap,
                                             lim_s=32, lim_t=32), S)
            assert np.float32(scores[k]).tobytes() == r.score.tobytes(), (k, loc)
