"""Solution matrices on the GPU: Solution.values, traceback("matrix") and
traceback("edges") (pyalign/solve.py:44-57, solver.h:2355-2471) against the
oracle, whose matrices the CPU suite pins by enumeration
(tests/test_oracle_matrices.py).  Values must be bit-identical (float32
bytes, so signed zeros and infinities count), tracebacks identical.
"""

import typing
import zlib

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from tests.test_gpu_parity import SOLVERS, build_factory, dense, gap_obj, random_bag, rounded

pytestmark = pytest.mark.gpu


def assert_matrices(sol, ref, what):
    vals = np.asarray(sol.values, dtype=np.float32)
    rv = ref.values()
    assert vals.shape == rv.shape, what
    assert np.array_equal(vals.view(np.uint32), rv.view(np.uint32)), what
    tbm = np.asarray(sol.traceback("matrix")).astype(np.int32)
    assert np.array_equal(tbm, ref.traceback_matrix()), what
    got, want = sol.traceback("edges"), ref.traceback_edges()
    assert len(got) == len(want), what
    for layer, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(np.asarray(g).astype(np.int32), w), (what, layer)
    assert sol.shape == rv.shape


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 4, 1), ("affine", 2.5, 0.5)], ids=str)
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
@pytest.mark.parametrize("form", ["general", "alphabetic", "binary"])
def test_solution_matrices(loc, gap, direction, form):
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, direction, form, "mat")).encode()))
    alphabet, pf = build_factory(form, rng, direction)
    seqs, probs = random_bag(rng, pf, alphabet, 9, 0, 24)
    cls, oloc = SOLVERS[loc]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    solver = cls(gap_cost=gap_obj(gap), codomain=pa.Solution)
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    # a bag goes to the GPU in one call; a single problem through the batch form
    for p, sol in zip(probs, solver.solve(probs)):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=odir, gap=gap, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        assert_matrices(sol, ref, (p.s, p.t))
    p = probs[0]
    sol = solver.solve(p)
    ref = O.OracleResult(O.make_config(locality=oloc, direction=odir, gap=gap, goal=O.GOAL_ONE,
                                       lim_s=rounded(len(p.s)), lim_t=rounded(len(p.t))), dense(p))
    assert_matrices(sol, ref, (p.s, p.t))


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 3, 1), ("affine", 1, 1)], ids=str)
def test_all_optimal_solution_matrices(loc, gap):
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, "all-mat")).encode()))
    alphabet, pf = build_factory("alphabetic", rng, "maximize")
    seqs, probs = random_bag(rng, pf, alphabet, 8, 1, 9)
    cls, oloc = SOLVERS[loc]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    lists = cls(gap_cost=gap_obj(gap), codomain=typing.List[pa.Solution]).solve(probs)
    for p, sols in zip(probs, lists):
        ref = O.OracleResult(O.make_config(locality=oloc, direction=O.MAXIMIZE, gap=gap, goal=O.GOAL_ALL,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        for sol in sols:  # every solution of a problem carries the same matrices
            assert_matrices(sol, ref, (p.s, p.t))


@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 5, 1)], ids=str)
def test_multi_strip_solution_matrices(gap):
    """Rows beyond one strip (row buffer between strips, several strips of
    traceback words)."""
    rng = np.random.default_rng(71)
    alphabet = "ACDEFGHIKLMNPQRSTVWY"
    table = rng.integers(-4, 5, size=(20, 20)).astype(np.float32)
    table = np.triu(table) + np.triu(table, 1).T
    lookup = {(x, y): table[i, j] for i, x in enumerate(alphabet) for j, y in enumerate(alphabet)}
    pf = pa.alphabetic(list(alphabet), lambda x, y: lookup[(x, y)])
    seqs, probs = random_bag(rng, pf, alphabet, 3, 520, 700)
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    for p, sol in zip(probs, pa.LocalSolver(gap_cost=gap_obj(gap), codomain=pa.Solution).solve(probs)):
        ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=gap, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        assert_matrices(sol, ref, (len(p.s), len(p.t)))


@pytest.mark.parametrize("direction", ["minimize", "maximize"])
def test_dtw_solution_matrices(direction):
    rng = np.random.default_rng(73)
    pf = pa.general(lambda x, y: float(np.float32(abs(x - y) * 0.5)), direction=direction)
    seqs = [(rng.integers(0, 9, int(rng.integers(0, 30))), rng.integers(0, 9, int(rng.integers(0, 30))))
            for _ in range(10)]
    probs = [pf.new_problem(list(a), list(b)) for a, b in seqs]
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    for p, sol in zip(probs, pa.ElasticSolver(codomain=pa.Solution).solve(probs)):
        ref = O.OracleResult(O.make_config(solver_type=O.DTW, direction=odir, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), p.matrix.astype(np.float32))
        assert_matrices(sol, ref, (len(p.s), len(p.t)))


def test_dtw_series_solution_matrices():
    rng = np.random.default_rng(74)
    pf = pa.series(direction="minimize")
    data = [(np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(1, 60)), 2)), 0),
             np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(1, 60)), 2)), 0)) for _ in range(6)]
    probs = [pf.new_problem(x, y) for x, y in data]
    lims = (rounded(max(len(x) for x, _ in data)), rounded(max(len(y) for _, y in data)))
    for p, sol in zip(probs, pa.ElasticSolver(codomain=pa.Solution).solve(probs)):
        ref = O.OracleResult(O.make_config(solver_type=O.DTW, direction=O.MINIMIZE, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), p.matrix)
        assert_matrices(sol, ref, (len(p.s), len(p.t)))
