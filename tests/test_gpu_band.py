"""Band mode on the GPU: the strips of a long problem computed concurrently
by several wavefronts, the bottom row of each strip handed to the next
through the row buffer and a published column count (and, for the forms it
serves, the long mode that replaces it: tests/test_gpu_long.py).  Results must equal
the oracle's exactly, as in every other schedule; each test also checks
that the schedule really ran problems in band mode.
"""

import zlib

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from pyalign_amd import algorithm

pytestmark = pytest.mark.gpu

LOCS = {"local": (algorithm.Locality.LOCAL, O.LOCAL), "global": (algorithm.Locality.GLOBAL, O.GLOBAL),
        "semiglobal": (algorithm.Locality.SEMIGLOBAL, O.SEMIGLOBAL)}


def make_solver(loc, gap, direction, codomain, lim_s, lim_t):
    d = algorithm.Direction.MAXIMIZE if direction == "maximize" else algorithm.Direction.MINIMIZE
    g = pa.LinearGapCost(gap[1]) if gap[0] == "linear" else pa.AffineGapCost(gap[1], gap[2])
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=d, return_dup=False,
        codomain=pa.solve.Codomain(codomain), locality=LOCS[loc][0], gap_cost=g))
    return algorithm.m16.create_solver(lim_s, lim_t, opts)


def lengths(rng, n, lo_s, hi_s, lo_t, hi_t):
    return (rng.integers(lo_s, hi_s + 1, n).astype(np.int32), rng.integers(lo_t, hi_t + 1, n).astype(np.int32))


def ragged_codes(rng, ls, lt, K):
    cs = rng.integers(0, K, int(ls.sum())).astype(np.uint8)
    ct = rng.integers(0, K, int(lt.sum())).astype(np.uint8)
    os_ = np.concatenate([[0], np.cumsum(ls.astype(np.int64))[:-1]])
    ot = np.concatenate([[0], np.cumsum(lt.astype(np.int64))[:-1]])
    return cs, ct, os_, ot


def sym_table(rng, K, lo=-4, hi=5):
    t = rng.integers(lo, hi, size=(K, K)).astype(np.float32)
    return np.triu(t) + np.triu(t, 1).T


def dense_of(form, k, cs, ct, os_, ot, ls, lt, table=None, eq=None, ne=None):
    a = cs[os_[k]:os_[k] + ls[k]]
    b = ct[ot[k]:ot[k] + lt[k]]
    if form == "indexed":
        return table[a][:, b]
    return np.where(a[:, None] == b[None, :], eq, ne).astype(np.float32)


@pytest.mark.parametrize("mode", ["long", "band"])
@pytest.mark.parametrize("loc", list(LOCS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 5, 1)], ids=str)
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
@pytest.mark.parametrize("form", ["indexed", "binary"])
def test_band_scores_and_paths(loc, gap, direction, form, mode, monkeypatch):
    """mode "long": critical-path problems in the long mode (workgroup
    sweeps, the default where the form allows it); "band": band mode
    (strips on separate wavefronts), which the long mode replaces for these
    forms."""
    if mode == "band":
        monkeypatch.setenv("PYALIGN_AMD_NO_LONG_MODE", "1")
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, direction, form, "band")).encode()))
    ls, lt = lengths(rng, 3, 1100, 1700, 300, 700)
    K = 4 if form == "binary" else 6
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, K)
    table = sym_table(rng, K)
    if direction == "minimize":
        table = -table
    eq, ne = (2.0, -1.0) if direction == "maximize" else (-2.0, 1.0)
    payload = dict(codes_s=cs, codes_t=ct, off_s=os_, off_t=ot)
    if form == "indexed":
        payload["table"] = table
    else:
        payload.update(codes_s=cs.astype(np.uint32), codes_t=ct.astype(np.uint32), eq=eq, ne=ne)
    lim_s, lim_t = 2048, 1024
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    for codomain, kind, goal in ((pa.Score, "score", O.GOAL_SCORE), (pa.Solution, "solution", O.GOAL_ONE)):
        solver = make_solver(loc, gap, direction, codomain, lim_s, lim_t)
        res = solver.solve_ragged(kind, form, ls, lt, **payload)
        # local maximisation over a table, score only, runs the fast kernel,
        # whose tall strips (32 rows a lane) beat band mode at these shapes
        # (not the long mode)
        fast = loc == "local" and direction == "maximize" and form == "indexed" and kind == "score"
        groups, band, _ = solver.last_schedule()
        if mode == "band":
            assert band == (0 if fast else 3), solver.last_schedule()
        elif not fast:  # the long mode, or tall strips where the planner's cost model prefers them
            assert band >= 1, solver.last_schedule()
        for k in range(3):
            S = dense_of(form, k, cs, ct, os_, ot, ls, lt, table, eq, ne)
            ref = O.OracleResult(O.make_config(locality=LOCS[loc][1], direction=odir, gap=gap, goal=goal,
                                               lim_s=lim_s, lim_t=lim_t), S)
            if kind == "score":
                assert np.float32(res[k]).tobytes() == ref.score.tobytes(), k
            else:
                assert np.float32(res[k].score).tobytes() == ref.score.tobytes(), k
                assert np.array_equal(np.asarray(res[k].path, dtype=np.int32), ref.path()), k


def test_band_mixed_with_short_problems_one_launch():
    """Long problems in band mode next to short ones in the same queue."""
    rng = np.random.default_rng(81)
    ls_long, lt_long = lengths(rng, 2, 2500, 3000, 800, 1000)
    ls_short, lt_short = lengths(rng, 300, 10, 60, 10, 60)
    ls = np.concatenate([ls_short[:150], ls_long, ls_short[150:]])
    lt = np.concatenate([lt_short[:150], lt_long, lt_short[150:]])
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, 5)
    table = sym_table(rng, 5)
    solver = make_solver("semiglobal", ("affine", 3, 1), "maximize", pa.Solution, 4096, 1024)
    res = solver.solve_ragged("solution", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot,
                              table=table)
    assert solver.last_schedule()[1] >= 1
    for k in list(range(140, 160)) + [0, len(ls) - 1]:
        S = dense_of("indexed", k, cs, ct, os_, ot, ls, lt, table)
        ref = O.OracleResult(O.make_config(locality=O.SEMIGLOBAL, gap=("affine", 3, 1), goal=O.GOAL_ONE,
                                           lim_s=4096, lim_t=1024), S)
        assert np.float32(res[k].score).tobytes() == ref.score.tobytes(), k
        assert np.array_equal(np.asarray(res[k].path, dtype=np.int32), ref.path()), k


def test_band_fast_job_local_score():
    """Local maximisation over an integer table (the score-only fast path):
    band strips of the long problems run in a general-kernel launch of their
    own beside the fast kernel."""
    rng = np.random.default_rng(82)
    ls_long, lt_long = lengths(rng, 2, 3000, 4000, 1500, 2000)
    ls_short, lt_short = lengths(rng, 100, 50, 300, 50, 300)
    ls = np.concatenate([ls_long, ls_short])
    lt = np.concatenate([lt_long, lt_short])
    K = 20
    table = sym_table(rng, K)
    np.fill_diagonal(table, rng.integers(4, 12, K))
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, K)
    solver = make_solver("local", ("affine", 10, 1), "maximize", pa.Score, 4096, 2048)
    res = solver.solve_ragged("score", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot, table=table)
    assert solver.last_schedule()[1] >= 1
    for k in [0, 1, 2, 50, 101]:
        S = dense_of("indexed", k, cs, ct, os_, ot, ls, lt, table)
        ref = O.solve(S, locality=O.LOCAL, gap=("affine", 10, 1), lim_s=4096, lim_t=2048)
        assert np.float32(res[k]).tobytes() == ref.score.tobytes(), k


def test_band_solution_matrices(monkeypatch):
    """Solution.values and tracebacks of a problem solved in band mode (the
    matrices always are: the long mode keeps no cell values)."""
    monkeypatch.setenv("PYALIGN_AMD_NO_LONG_MODE", "1")
    rng = np.random.default_rng(83)
    ls, lt = lengths(rng, 2, 1100, 1300, 200, 260)
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, 4)
    table = sym_table(rng, 4)
    solver = make_solver("global", ("affine", 4, 1), "maximize", pa.Solution, 2048, 256)
    res = solver.solve_ragged("solution", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot,
                              table=table)
    assert solver.last_schedule()[1] == 2
    k = 1
    S = dense_of("indexed", k, cs, ct, os_, ot, ls, lt, table)
    ref = O.OracleResult(O.make_config(locality=O.GLOBAL, gap=("affine", 4, 1), goal=O.GOAL_ONE,
                                       lim_s=2048, lim_t=256), S)
    vals = np.asarray(res[k].values, dtype=np.float32)
    assert np.array_equal(vals.view(np.uint32), ref.values().view(np.uint32))
    assert np.array_equal(np.asarray(res[k].traceback_as_matrix).astype(np.int32), ref.traceback_matrix())


def test_band_dtw_series():
    rng = np.random.default_rng(84)
    pf = pa.series(direction="minimize")
    data = [(np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(1500, 2000)), 3)), 0),
             np.cumsum(rng.normal(0, 0.1, size=(int(rng.integers(300, 500)), 3)), 0)) for _ in range(2)]
    probs = [pf.new_problem(x, y) for x, y in data]
    res = pa.ElasticSolver(codomain=pa.Solution).solve(probs)
    lims = (pa.solve.SizeRounder()(max(len(x) for x, _ in data)), pa.solve.SizeRounder()(max(len(y) for _, y in data)))
    for p, r in zip(probs, res):
        ref = O.OracleResult(O.make_config(solver_type=O.DTW, direction=O.MINIMIZE, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), p.matrix)
        assert np.float32(r.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(r.path, dtype=np.int32), ref.path())


def test_long_pair_full_size_score_and_solution():
    """The single long semiglobal pair: |B| = 10,000 uniform ACGT, A a
    5,000-long window with substitutions and indels; affine (2.0, 0.5),
    binary eq 1 ne -1, score and single Solution."""
    import bench
    a, b = bench.c3_inputs()
    pf = pa.binary(eq=1, ne=-1)
    problem = pf.new_problem(a, b)
    ca = np.frombuffer(a.encode(), np.uint8)
    cb = np.frombuffer(b.encode(), np.uint8)
    S = np.where(ca[:, None] == cb[None, :], 1.0, -1.0).astype(np.float32)
    lims = (pa.solve.SizeRounder()(len(a)), pa.solve.SizeRounder()(len(b)))
    score = pa.SemiglobalSolver(gap_cost=pa.AffineGapCost(2.0, 0.5), codomain=pa.Score).solve(problem)
    ref = O.solve(S, locality=O.SEMIGLOBAL, gap=("affine", 2.0, 0.5), lim_s=lims[0], lim_t=lims[1])
    assert np.float32(score).tobytes() == ref.score.tobytes()
    sol = pa.SemiglobalSolver(gap_cost=pa.AffineGapCost(2.0, 0.5), codomain=pa.Solution).solve(problem)
    ref1 = O.OracleResult(O.make_config(locality=O.SEMIGLOBAL, gap=("affine", 2.0, 0.5), goal=O.GOAL_ONE,
                                        lim_s=lims[0], lim_t=lims[1]), S)
    assert np.float32(sol.score).tobytes() == ref1.score.tobytes()
    assert np.array_equal(np.asarray(sol.path, dtype=np.int32), ref1.path())
