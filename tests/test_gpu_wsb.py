"""General gap costs (Waterman-Smith-Beyer, pyalign/algorithm/solver.h:3047-3184)
on the GPU against the oracle, whose general solver the CPU suite pins by
the reference's known-answer test, brute force and closed forms
(tests/test_oracle_wsb.py).  Scores and Solution values bit-identical,
paths, tracebacks and all-optimal enumerations identical and in order.
"""

import typing
import zlib

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from tests.test_gpu_matrices import assert_matrices
from tests.test_gpu_parity import SOLVERS, build_factory, dense, edges_sorted, factories, random_bag, rounded

pytestmark = pytest.mark.gpu

# non-subadditive integer costs: some gaps are cheaper as two jumps
STEPPY = pa.UserFuncGapCost(lambda k: [0, 2, 5, 5, 9, 9, 9, 14][min(k, 7)] + (k > 7) * (k - 7))
GAPS = {"log": pa.LogarithmicGapCost(3, 1), "exp": pa.ExponentialGapCost(2, 0.5), "steppy": STEPPY,
        "smooth": pa.smooth_gap_cost(4)}


def oracle_gap(g, lims):
    return ("general", g.costs(lims[0] + 1), g.costs(lims[1] + 1))


def test_kat_wsb():  # pyalign/tests/test_wsb.py:10-35
    for pf in factories("ACG", pa.Equality(eq=1, ne=-1)):
        solver = pa.GlobalSolver(gap_cost=pa.LogarithmicGapCost(3, 1), codomain=typing.Iterator[pa.Alignment])
        p = pf.new_problem("CG", "CCGA")
        al = list(solver.solve(p))
        assert all(round(x.score, 1) == -3.7 for x in al)
        assert edges_sorted(al) == [((0, 0), (1, 1)), ((0, 0), (1, 3))]
        ref = O.OracleResult(O.make_config(locality=O.GLOBAL, goal=O.GOAL_ALL,
                                           gap=oracle_gap(pa.LogarithmicGapCost(3, 1), (2, 4)), lim_s=2, lim_t=4),
                             dense(p))
        assert np.float32(al[0].score).tobytes() == np.float32(ref.all_next()[0]).tobytes()


def test_algorithm_metadata():
    pf = pa.general(pa.Equality(eq=1, ne=-1))
    sol = pa.GlobalSolver(gap_cost=pa.LogarithmicGapCost(3, 1), codomain=pa.Solution).solve(
        pf.new_problem("ACGT", "AGT"))
    assert sol.complexity.name == "Waterman-Smith-Beyer"  # solver.h:3104-3108
    assert sol.complexity.runtime == "n^3"


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", list(GAPS))
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
@pytest.mark.parametrize("form", ["general", "alphabetic", "binary"])
def test_random_score_and_solution(loc, gap, direction, form):
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, direction, form, "wsb")).encode()))
    alphabet, pf = build_factory(form, rng, direction)
    seqs, probs = random_bag(rng, pf, alphabet, 23, 0, 30)
    cls, oloc = SOLVERS[loc]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    og = oracle_gap(GAPS[gap], lims)
    got = cls(gap_cost=GAPS[gap], codomain=pa.Score).solve(probs)
    for p, g in zip(probs, got):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=odir, gap=og, goal=O.GOAL_SCORE,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        assert np.float32(g).tobytes() == ref.score.tobytes(), (p.s, p.t, g, ref.score)
    sols = cls(gap_cost=GAPS[gap], codomain=pa.Solution).solve(probs)
    for p, s in zip(probs, sols):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=odir, gap=og, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        if not ref.has_seed:
            assert s.score is None and s.path is None
            continue
        assert np.float32(s.score).tobytes() == ref.score.tobytes(), (p.s, p.t)
        assert np.array_equal(np.asarray(s.path, dtype=np.int32), ref.path()), (p.s, p.t, s.path, ref.path())
        assert sorted(s.alignment.edges.tolist()) == sorted(ref.edges().tolist())


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
def test_random_all_optimal_in_order(loc, direction):
    """Integer costs and an integer table make many ties."""
    rng = np.random.default_rng(zlib.crc32(repr((loc, direction, "wsb-all")).encode()))
    alphabet = "ACGT"
    t = rng.integers(-2, 3, size=(4, 4)).astype(np.float32)
    t = np.triu(t) + np.triu(t, 1).T
    if direction == "minimize":
        t = -t
    lookup = {(x, y): t[i, j] for i, x in enumerate(alphabet) for j, y in enumerate(alphabet)}
    pf = pa.alphabetic(alphabet, lambda x, y: lookup[(x, y)], direction=direction)
    seqs, probs = random_bag(rng, pf, alphabet, 20, 0, 8)
    cls, oloc = SOLVERS[loc]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    gap = pa.UserFuncGapCost(lambda k: [0, 1, 3, 3, 4, 6, 6, 7, 8][min(k, 8)])
    its = cls(gap_cost=gap, codomain=typing.Iterator[pa.Solution]).solve(probs)
    for p, it in zip(probs, its):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=odir, gap=oracle_gap(gap, lims),
                                           goal=O.GOAL_ALL, lim_s=lims[0], lim_t=lims[1]), dense(p))
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
@pytest.mark.parametrize("goal", [O.GOAL_ONE, O.GOAL_ALL])
def test_solution_matrices(loc, goal):
    rng = np.random.default_rng(zlib.crc32(repr((loc, goal, "wsb-mat")).encode()))
    alphabet, pf = build_factory("alphabetic", rng, "maximize")
    seqs, probs = random_bag(rng, pf, alphabet, 6, 1, 12)
    cls, oloc = SOLVERS[loc]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    codomain = pa.Solution if goal == O.GOAL_ONE else typing.List[pa.Solution]
    res = cls(gap_cost=STEPPY, codomain=codomain).solve(probs)
    for p, r in zip(probs, res):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=O.MAXIMIZE, gap=oracle_gap(STEPPY, lims),
                                           goal=goal, lim_s=lims[0], lim_t=lims[1]), dense(p))
        for sol in ([r] if goal == O.GOAL_ONE else r):
            assert_matrices(sol, ref, (p.s, p.t))


@pytest.mark.parametrize("loc", list(SOLVERS))
def test_longer_problems(loc):
    """Anti-diagonals longer than the 256-thread workgroup."""
    rng = np.random.default_rng(zlib.crc32(repr((loc, "wsb-long")).encode()))
    alphabet = "ACDEFGHIKLMNPQRSTVWY"
    t = rng.integers(-4, 4, size=(20, 20)).astype(np.float32)
    t = np.triu(t) + np.triu(t, 1).T
    np.fill_diagonal(t, 6)
    lookup = {(x, y): t[i, j] for i, x in enumerate(alphabet) for j, y in enumerate(alphabet)}
    pf = pa.alphabetic(list(alphabet), lambda x, y: lookup[(x, y)])
    seqs, probs = random_bag(rng, pf, alphabet, 3, 260, 420)
    cls, oloc = SOLVERS[loc]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    g = pa.LogarithmicGapCost(4, 1.5)
    res = cls(gap_cost=g, codomain=pa.Solution).solve(probs)
    for p, r in zip(probs, res):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=O.MAXIMIZE, gap=oracle_gap(g, lims),
                                           goal=O.GOAL_ONE, lim_s=lims[0], lim_t=lims[1]), dense(p))
        assert np.float32(r.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path())


def test_asymmetric_costs():
    """Different general costs for s and t (a dict gap_cost)."""
    rng = np.random.default_rng(5)
    alphabet, pf = build_factory("general", rng, "maximize")
    seqs, probs = random_bag(rng, pf, alphabet, 12, 1, 20)
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    gs, gt = pa.LogarithmicGapCost(2, 1), STEPPY
    sols = pa.SemiglobalSolver(gap_cost={"s": gs, "t": gt}, codomain=pa.Solution).solve(probs)
    for p, s in zip(probs, sols):
        ref = O.OracleResult(O.make_config(locality=O.SEMIGLOBAL, gap=("general", gs.costs(lims[0] + 1),
                                                                       gt.costs(lims[1] + 1)),
                                           goal=O.GOAL_ONE, lim_s=lims[0], lim_t=lims[1]), dense(p))
        assert np.float32(s.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(s.path, dtype=np.int32), ref.path())
