"""Parity of the HIP solver with the reference semantics, on the GPU.

Expected values come from the reference's own known-answer tests
(pyalign/tests/*.py, README.md:94-110) and from the CPU oracle
(oracle/pa_oracle.cpp), which the CPU suite pins first.  Scores must be
bit-identical (float32 bytes), paths and alignments identical, all-optimal
enumerations identical in content and order.
"""

import typing

import zlib

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from pyalign_amd import _lib, algorithm

pytestmark = pytest.mark.gpu

SOLVERS = {"local": (pa.LocalSolver, O.LOCAL), "global": (pa.GlobalSolver, O.GLOBAL),
           "semiglobal": (pa.SemiglobalSolver, O.SEMIGLOBAL)}


def gap_obj(g):
    return pa.LinearGapCost(g[1]) if g[0] == "linear" else pa.AffineGapCost(g[1], g[2])


def rounded(x):
    return pa.solve.SizeRounder()(x)


def dense(problem):
    if isinstance(problem, pa.BinaryProblem):
        eq, ne = problem.binary_similarity_values()
        a = pa.IntegerEncoder().encode(problem.s)
        b = pa.IntegerEncoder().encode(problem.t)
        return np.where(np.asarray(a)[:, None] == np.asarray(b)[None, :], eq, ne).astype(np.float32)
    return problem.matrix.astype(np.float32)


def factories(alphabet, sim, direction="maximize"):
    yield pa.general(sim, direction=direction)
    yield pa.alphabetic(alphabet, sim, direction=direction)
    if sim.binary_similarity_values is not None:
        yield pa.binary(*sim.binary_similarity_values, direction=direction)


def edges_sorted(alignments):
    return sorted(tuple(map(tuple, a.edges.tolist())) for a in alignments)


# ---------------------------------------------------------- reference KATs ---

def test_kat_nw():  # pyalign/tests/test_nw.py:9-29
    for pf in factories("ACGT", pa.Equality(eq=1, ne=-1)):
        al = pa.GlobalSolver(gap_cost=pa.LinearGapCost(2), codomain=pa.Alignment).solve(pf.new_problem("AATCG", "AACG"))
        assert al.score == 2.0
        assert al.edges.tolist() == [[0, 0], [1, 1], [3, 2], [4, 3]]


def test_kat_sw():  # pyalign/tests/test_sw.py:10-31
    for pf in factories("ACGT", pa.Equality(eq=1, ne=-1)):
        solver = pa.LocalSolver(gap_cost=pa.LinearGapCost(2), codomain=typing.Iterator[pa.Alignment])
        al = list(solver.solve(pf.new_problem("AATCG", "AACG")))
        assert all(a.score == 2 for a in al)
        assert edges_sorted(al) == [((0, 0), (1, 1)), ((3, 2), (4, 3))]


def test_kat_batch():  # pyalign/tests/test_batch.py:10-41
    for pf in factories("ACGT", pa.Equality(eq=1, ne=-1)):
        solver = pa.GlobalSolver(gap_cost=pa.LinearGapCost(2), codomain=typing.Iterator[pa.Alignment],
                                 return_dup=False)
        res = solver.solve([pf.new_problem("AATCG", "AACG"), pf.new_problem("AATGC", "AACG"),
                            pf.new_problem("AATCG", "AGTT")])
        exp = [(2, [((0, 0), (1, 1), (3, 2), (4, 3))]), (0, [((0, 0), (1, 1), (2, 2), (3, 3))]),
               (-2, [((0, 0), (1, 1), (2, 2), (3, 3)), ((0, 0), (1, 1), (2, 2), (4, 3))])]
        for it, (score, edges) in zip(res, exp):
            al = list(it)
            assert all(a.score == score for a in al)
            assert edges_sorted(al) == sorted(edges)


GOTOH_KATS = [  # pyalign/tests/test_gotoh.py:13-176
    ("ACT", pa.Equality(eq=0, ne=1), "minimize", (5, 1), "CC", "ACCT", 7.0,
     [((0, 0), (1, 1)), ((0, 2), (1, 3))]),
    ("ACGT", pa.Dict({("T", "C"): -1, ("A", "T"): 0, ("G", "T"): -2, ("A", "A"): 6, ("T", "T"): 6, ("G", "G"): 6}),
     "maximize", (11, 1), "GGTGTGA", "TCGCGT", -2, [((0, 0), (1, 1), (3, 2), (4, 3), (5, 4), (6, 5))]),
    ("AGT", pa.Equality(eq=0, ne=-1), "maximize", (5, 1), "AAAGGG", "TTAAAAGGGGTT", -14,
     [((0, 3), (1, 4), (2, 5), (3, 6), (4, 7), (5, 8)), ((0, 0), (1, 1), (2, 2), (3, 9), (4, 10), (5, 11))]),
    ("ACGT", pa.Equality(eq=10, ne=-30), "maximize", (40, 1), "TAAATTTGC", "TCGCCTTAC", -60,
     [((0, 0), (1, 7), (8, 8)), ((0, 6), (1, 7), (8, 8))]),
    ("ACGT", pa.Equality(eq=10, ne=-30), "maximize", (25, 1), "AGAT", "CTCT", -44, [((3, 3),)]),
    ("ACGT", pa.Equality(eq=10, ne=-30), "maximize", (30, 1), "AGAT", "CTCT", -54, [((3, 3),)]),
]


@pytest.mark.parametrize("case", range(len(GOTOH_KATS)))
def test_kat_gotoh(case):
    alphabet, sim, direction, (o, e), a, b, score, edges = GOTOH_KATS[case]
    for pf in factories(alphabet, sim, direction):
        for codomain in (typing.List[pa.Alignment], typing.Iterator[pa.Alignment]):
            solver = pa.GlobalSolver(gap_cost=pa.AffineGapCost(o, e), codomain=codomain)
            al = list(solver.solve(pf.new_problem(a, b)))
            assert all(x.score == score for x in al)
            assert edges_sorted(al) == sorted(edges)


def test_kat_readme():  # README.md:94-110
    al = pa.global_alignment("INDUSTRY", "INTEREST", gap_cost=0, eq=1, ne=-1)
    assert al.score == 4.0
    assert al.edges.tolist() == [[0, 0], [1, 1], [4, 6], [5, 7]]
    pf = pa.general(pa.Equality(eq=1, ne=-1), direction="maximize")
    sol = pa.GlobalSolver(gap_cost=pa.LinearGapCost(0.2), codomain=pa.Solution).solve(pf.new_problem("INDUSTRY", "INTEREST"))
    assert sol.alignment.edges.tolist() == [[0, 0], [1, 1], [4, 6], [5, 7]]
    assert sol.algorithm.name if hasattr(sol, "algorithm") else True


# ------------------------------------------------------ randomized vs oracle ---

def random_bag(rng, pf, alphabet, n, lo, hi):
    seqs = []
    for _ in range(n):
        a = "".join(rng.choice(list(alphabet), int(rng.integers(lo, hi + 1))))
        b = "".join(rng.choice(list(alphabet), int(rng.integers(lo, hi + 1))))
        seqs.append((a, b))
    return seqs, [pf.new_problem(a, b) for a, b in seqs]


def make_table(rng, alphabet, half=False):
    K = len(alphabet)
    t = rng.integers(-3, 5, size=(K, K)).astype(np.float32)
    if half:
        t = t / 2
    t = np.triu(t) + np.triu(t, 1).T
    return t


def oracle_for(problem, loc, direction, gap, goal, lims):
    return O.OracleResult(O.make_config(locality=loc, direction=direction, gap=gap, goal=goal,
                                        lim_s=lims[0], lim_t=lims[1]), dense(problem))


FORMS = ["general", "alphabetic", "binary"]
GAPS = [("linear", 1), ("linear", 0.5), ("affine", 4, 1), ("affine", 2.5, 0.5)]


def build_factory(form, rng, direction):
    alphabet = "ACGT"
    if form == "binary":
        return alphabet, pa.binary(eq=2, ne=-1, direction=direction)
    table = make_table(rng, alphabet, half=True)
    lookup = {(x, y): table[i, j] for i, x in enumerate(alphabet) for j, y in enumerate(alphabet)}
    w = lambda x, y: lookup[(x, y)]  # noqa: E731
    if form == "alphabetic":
        return alphabet, pa.alphabetic(alphabet, w, direction=direction)
    return alphabet, pa.general(w, direction=direction)


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", GAPS, ids=str)
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
@pytest.mark.parametrize("form", FORMS)
def test_random_score_and_solution(loc, gap, direction, form):
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, direction, form)).encode()))
    alphabet, pf = build_factory(form, rng, direction)
    seqs, probs = random_bag(rng, pf, alphabet, 37, 0, 40)
    cls, oloc = SOLVERS[loc]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    # scores
    got = cls(gap_cost=gap_obj(gap), codomain=pa.Score).solve(probs)
    for p, g in zip(probs, got):
        ref = oracle_for(p, oloc, odir, gap, O.GOAL_SCORE, lims)
        assert np.float32(g).tobytes() == ref.score.tobytes(), (p.s, p.t, g, ref.score)
    # one optimal solution: score, path, alignment
    sols = cls(gap_cost=gap_obj(gap), codomain=pa.Solution).solve(probs)
    for p, s in zip(probs, sols):
        ref = oracle_for(p, oloc, odir, gap, O.GOAL_ONE, lims)
        if not ref.has_seed:
            assert s.score is None and s.path is None
            continue
        assert np.float32(s.score).tobytes() == ref.score.tobytes(), (p.s, p.t)
        assert np.array_equal(np.asarray(s.path, dtype=np.int32), ref.path()), (p.s, p.t, s.path, ref.path())
        assert sorted(s.alignment.edges.tolist()) == sorted(ref.edges().tolist())


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 3, 1), ("affine", 1, 1)], ids=str)
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
def test_random_all_optimal_in_order(loc, gap, direction):
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, direction, "all")).encode()))
    alphabet, pf = build_factory("alphabetic", rng, direction)
    seqs, probs = random_bag(rng, pf, alphabet, 25, 0, 9)
    cls, oloc = SOLVERS[loc]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    its = cls(gap_cost=gap_obj(gap), codomain=typing.Iterator[pa.Solution]).solve(probs)
    for p, it in zip(probs, its):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=odir, gap=gap, goal=O.GOAL_ALL,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        expect = []
        while True:
            x = ref.all_next()
            if x is None:
                break
            expect.append(x)
        got = list(it)
        assert len(got) == len(expect), (p.s, p.t)
        for g, (score, path, edges) in zip(got, expect):
            assert np.float32(g.score).tobytes() == np.float32(score).tobytes()
            assert np.array_equal(np.asarray(g.path, dtype=np.int32), path)
            assert sorted(g.alignment.edges.tolist()) == sorted(edges.tolist())


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 5, 1)], ids=str)
def test_long_problems_multi_strip(loc, gap):
    """Rows beyond one strip go through the row buffer between strips."""
    rng = np.random.default_rng(21)
    alphabet = "ACDEFGHIKLMNPQRSTVWY"
    table = make_table(rng, alphabet)
    np.fill_diagonal(table, 6)
    lookup = {(x, y): table[i, j] for i, x in enumerate(alphabet) for j, y in enumerate(alphabet)}
    pf = pa.alphabetic(list(alphabet), lambda x, y: lookup[(x, y)])
    seqs, probs = random_bag(rng, pf, alphabet, 6, 500, 1300)
    cls, oloc = SOLVERS[loc]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    for codomain, goal in ((pa.Score, O.GOAL_SCORE), (pa.Solution, O.GOAL_ONE)):
        res = cls(gap_cost=gap_obj(gap), codomain=codomain).solve(probs)
        for p, r in zip(probs, res):
            ref = oracle_for(p, oloc, O.MAXIMIZE, gap, goal, lims)
            score = r if codomain is pa.Score else r.score
            assert np.float32(score).tobytes() == ref.score.tobytes()
            if codomain is pa.Solution:
                assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path())


def test_c2_shaped_batch_scores_bit_exact():
    """Local Gotoh (10, 1), 20-letter integer table, log-uniform lengths."""
    rng = np.random.default_rng(2)
    K = 20
    table = rng.integers(-4, 4, size=(K, K))
    table = np.triu(table) + np.triu(table, 1).T
    np.fill_diagonal(table, rng.integers(4, 12, size=K))
    table = table.astype(np.float32)
    n = 1500
    ls = np.floor(np.exp(rng.uniform(np.log(50), np.log(700), n))).astype(np.int32)
    lt = np.floor(np.exp(rng.uniform(np.log(50), np.log(700), n))).astype(np.int32)
    cs = [rng.integers(0, K, x).astype(np.uint8) for x in ls]
    ct = [rng.integers(0, K, x).astype(np.uint8) for x in lt]
    off_s = np.concatenate([[0], np.cumsum(ls)[:-1]])
    off_t = np.concatenate([[0], np.cumsum(lt)[:-1]])
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL, gap_cost=pa.AffineGapCost(10, 1)))
    solver = algorithm.m16.create_solver(1024, 1024, opts)
    for schedule in (_lib.PA_SCHED_AUTO, _lib.PA_SCHED_PLAIN):
        got = solver.solve_ragged("score", "indexed", ls, lt, schedule=schedule, codes_s=np.concatenate(cs),
                                  codes_t=np.concatenate(ct), off_s=off_s, off_t=off_t, table=table)
        for k in range(0, n, 7):
            ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=("affine", 10, 1), goal=O.GOAL_SCORE,
                                               lim_s=int(ls[k]), lim_t=int(lt[k])), table[np.ix_(cs[k], ct[k])])
            assert np.float32(got[k]).tobytes() == ref.score.tobytes(), (k, schedule)


def test_shim_reference_shaped_calls_with_stale_lanes():
    """The reference-shaped entry points with unused lanes holding stale data."""
    rng = np.random.default_rng(4)
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Alignment), locality=algorithm.Locality.GLOBAL,
        gap_cost=pa.AffineGapCost(3, 1)))
    solver = algorithm.m16.create_solver(32, 32, opts)
    B = solver.batch_size
    K = 4
    sim = make_table(rng, "ACGT")
    a = rng.integers(0, K, size=(B, 20)).astype(np.uint32)
    b = rng.integers(0, K, size=(B, 25)).astype(np.uint32)
    length = np.stack([rng.integers(0, 21, B), rng.integers(0, 26, B)]).astype(np.uint16)
    res = solver.solve_indexed_for_alignment(a, b, sim, length)
    assert len(res) == B
    for k in range(B):
        S = sim[np.ix_(a[k, :length[0, k]], b[k, :length[1, k]])]
        ref = O.OracleResult(O.make_config(locality=O.GLOBAL, gap=("affine", 3, 1), goal=O.GOAL_ONE,
                                           lim_s=32, lim_t=32), S)
        assert np.float32(res[k].score).tobytes() == ref.score.tobytes()
        s_to_t = np.full(length[0, k], -1)
        for u, v in ref.edges():
            s_to_t[u] = v
        assert res[k].s_to_t.tolist() == s_to_t.tolist()
    with pytest.raises(ValueError, match="out of bounds index in a"):
        solver.solve_indexed_for_score(a + 10, b, sim, length)
    with pytest.raises(ValueError, match="batch size"):
        solver.solve_indexed_for_score(a[:3], b[:3], sim, length[:, :3])


def test_length_limits_raise():
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=False, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL, gap_cost=pa.LinearGapCost(1)))
    solver = algorithm.m16.create_solver(8, 8, opts)
    a = np.zeros((1, 9), np.uint32)
    with pytest.raises(RuntimeError, match="exceeds configured maximum length 8"):
        solver.solve_binary_for_score(a, a, 1.0, -1.0, np.array([[9], [9]]))


# ------------------------------------------------------------------ DTW ---

def _dtw_oracle(C, direction, goal, lims):
    return O.OracleResult(O.make_config(solver_type=O.DTW, direction=direction, goal=goal,
                                        lim_s=lims[0], lim_t=lims[1]), C)


@pytest.mark.parametrize("direction", ["minimize", "maximize"])
def test_dtw_dense_general_problems(direction):
    rng = np.random.default_rng(31)
    pf = pa.general(lambda x, y: float(np.float32(abs(x - y) * 0.5)), direction=direction)
    seqs = [(rng.integers(0, 9, int(rng.integers(0, 30))), rng.integers(0, 9, int(rng.integers(0, 30))))
            for _ in range(30)]
    probs = [pf.new_problem(list(a), list(b)) for a, b in seqs]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    for codomain, goal in ((pa.Score, O.GOAL_SCORE), (pa.Solution, O.GOAL_ONE)):
        res = pa.ElasticSolver(codomain=codomain).solve(probs)
        for p, r in zip(probs, res):
            ref = _dtw_oracle(p.matrix.astype(np.float32), odir, goal, lims)
            score = r if codomain is pa.Score else r.score
            assert np.float32(score).tobytes() == ref.score.tobytes(), (p.s, p.t)
            if codomain is pa.Solution:
                assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path())
                assert r.complexity.name == "DTW"


@pytest.mark.parametrize("dim", [1, 2, 3, 4])
def test_dtw_series_form(dim):
    rng = np.random.default_rng(40 + dim)
    pf = pa.series(direction="minimize")
    data = [(np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(1, 90)), dim)), 0),
             np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(1, 90)), dim)), 0)) for _ in range(25)]
    probs = [pf.new_problem(x, y) for x, y in data]
    lims = (rounded(max(len(x) for x, _ in data)), rounded(max(len(y) for _, y in data)))
    for codomain, goal in ((pa.Score, O.GOAL_SCORE), (pa.Solution, O.GOAL_ONE)):
        res = pa.ElasticSolver(codomain=codomain).solve(probs)
        for p, r in zip(probs, res):
            ref = _dtw_oracle(p.matrix, O.MINIMIZE, goal, lims)
            score = r if codomain is pa.Score else r.score
            assert np.float32(score).tobytes() == ref.score.tobytes()
            if codomain is pa.Solution:
                assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path())


def test_dtw_series_long_multi_strip():
    rng = np.random.default_rng(50)
    pf = pa.series(direction="minimize")
    data = [(np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(600, 1500)), 3)), 0),
             np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(600, 1500)), 3)), 0)) for _ in range(4)]
    probs = [pf.new_problem(x, y) for x, y in data]
    lims = (rounded(max(len(x) for x, _ in data)), rounded(max(len(y) for _, y in data)))
    res = pa.ElasticSolver(codomain=pa.Solution).solve(probs)
    for p, r in zip(probs, res):
        ref = _dtw_oracle(p.matrix, O.MINIMIZE, O.GOAL_ONE, lims)
        assert np.float32(r.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path())


@pytest.mark.parametrize("case", ["ties", "inf", "nan"])
def test_dtw_series_min_folds_and_non_finite(case):
    """Series of small integers (many equal candidates): over finite series
    the kernels fold with min and take each code from equality tests, which
    must give the oracle's choices; a bag holding an infinite or NaN value
    keeps the strict compare-and-select folds, also the oracle's."""
    rng = np.random.default_rng({"ties": 61, "inf": 62, "nan": 63}[case])
    pf = pa.series(direction="minimize")
    lens = [int(rng.integers(3, 160)) for _ in range(19)] + [700]
    data = [(rng.integers(-2, 3, size=(a, 2)).astype(np.float32),
             rng.integers(-2, 3, size=(int(rng.integers(3, 160)), 2)).astype(np.float32)) for a in lens]
    if case == "inf":
        data[3][0][2, 1] = np.inf
    elif case == "nan":
        data[5][1][1, 0] = np.nan
    probs = [pf.new_problem(x, y) for x, y in data]
    lims = (rounded(max(len(x) for x, _ in data)), rounded(max(len(y) for _, y in data)))
    res = pa.ElasticSolver(codomain=pa.Solution).solve(probs)
    scores = pa.ElasticSolver(codomain=pa.Score).solve(probs)
    for k, (p, r, sc) in enumerate(zip(probs, res, scores)):
        ref = _dtw_oracle(p.matrix, O.MINIMIZE, O.GOAL_ONE, lims)
        if np.isnan(ref.score):
            assert np.isnan(r.score) and np.isnan(sc), k
        else:
            assert np.float32(r.score).tobytes() == ref.score.tobytes(), k
            assert np.float32(sc).tobytes() == ref.score.tobytes(), k
        assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path()), k


def test_c2_full_length_range_against_oracle():
    """The batched local Gotoh workload's generator over its whole length
    range, [50, 2000]: 4,096 pairs solved in one bag on the fast kernel
    (every strip height it serves), every 8th pair against the oracle."""
    import bench
    table, ls, lt, cs, ct, off_s, off_t = bench.c2_inputs(n=4096, seed=12)
    assert ls.max() > 1800 and lt.max() > 1800
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL, gap_cost=pa.AffineGapCost(10, 1)))
    lim = int(max(ls.max(), lt.max()))
    solver = algorithm.m16.create_solver(lim, lim, opts)
    got = solver.solve_ragged("score", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=off_s, off_t=off_t,
                              table=table)
    for k in range(0, len(ls), 8):
        S = table[np.ix_(cs[off_s[k]:off_s[k] + ls[k]], ct[off_t[k]:off_t[k] + lt[k]])]
        ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=("affine", 10, 1), goal=O.GOAL_SCORE,
                                           lim_s=lim, lim_t=lim), S)
        assert np.float32(got[k]).tobytes() == ref.score.tobytes(), k


@pytest.mark.parametrize("loc", list(SOLVERS))
def test_return_dup_all_optimal_in_order(loc):
    """return_dup=True: the DFS emits every path, duplicates of an already
    emitted compressed alignment included (options.h:21,
    solver.h:1705-1711), in the oracle's order."""
    rng = np.random.default_rng(zlib.crc32(("dup" + loc).encode()))
    alphabet = "ACGT"
    pf = pa.alphabetic(alphabet, lambda x, y: 3.0 if x == y else -5.0)
    seqs, probs = random_bag(rng, pf, alphabet, 25, 0, 8)
    # A-A, then G / T bridged by one gap on each side in either order (one
    # compressed alignment reached by two paths), then C-C
    seqs.append(("AGC", "ATC"))
    probs.append(pf.new_problem("AGC", "ATC"))
    cls, oloc = SOLVERS[loc]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    dups = 0
    for gap in (("linear", 1), ("affine", 2, 1)):
        its = cls(gap_cost=gap_obj(gap), codomain=typing.Iterator[pa.Solution], return_dup=True).solve(probs)
        for p, it in zip(probs, its):
            ref = O.OracleResult(O.make_config(locality=oloc, direction=O.MAXIMIZE, gap=gap, goal=O.GOAL_ALL,
                                               remove_dup=False, lim_s=lims[0], lim_t=lims[1]), dense(p))
            expect = []
            while True:
                x = ref.all_next()
                if x is None:
                    break
                expect.append(x)
            got = list(it)
            assert len(got) == len(expect), (p.s, p.t)
            keys = [tuple(map(tuple, e.tolist())) for _, _, e in expect]
            dups += len(keys) - len(set(keys))
            for g, (score, path, edges) in zip(got, expect):
                assert np.float32(g.score).tobytes() == np.float32(score).tobytes()
                assert np.array_equal(np.asarray(g.path, dtype=np.int32), path)
    assert dups > 0  # duplicates were emitted, not deduplicated
