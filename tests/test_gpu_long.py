"""Long-problem mode on the GPU (pa_long_kernel): a problem on the critical
path is swept by the waves of workgroups, one strip of 64 rows per wave,
bottom rows handed from wave to wave through LDS and from workgroup to
workgroup through tagged device-scope words.  Every score and path must
equal the oracle's; shapes cover several bands, a last strip of a few rows,
idle waves in the last band and columns fewer than a wave's lanes.
"""

import zlib

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from tests.test_gpu_band import LOCS, dense_of, make_solver, ragged_codes, sym_table

pytestmark = pytest.mark.gpu

SHAPES = [(2049, 700), (1024, 1500), (300, 4000), (769, 1200), (4200, 90)]


@pytest.mark.parametrize("shape", SHAPES, ids=str)
@pytest.mark.parametrize("loc", list(LOCS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 5, 1)], ids=str)
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
@pytest.mark.parametrize("form", ["indexed", "binary"])
def test_long_mode_single_problem(shape, loc, gap, direction, form):
    rng = np.random.default_rng(zlib.crc32(repr((shape, loc, gap, direction, form, "long")).encode()))
    ls = np.array([shape[0]], np.int32)
    lt = np.array([shape[1]], np.int32)
    K = 4 if form == "binary" else 7
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
    lim_s, lim_t = max(4200, int(ls[0])), max(1024, int(lt[0]))
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    S = dense_of(form, 0, cs, ct, os_, ot, ls, lt, table, eq, ne)
    for codomain, kind, goal in ((pa.Score, "score", O.GOAL_SCORE), (pa.Solution, "solution", O.GOAL_ONE)):
        solver = make_solver(loc, gap, direction, codomain, lim_s, lim_t)
        res = solver.solve_ragged(kind, form, ls, lt, **payload)
        groups, band, _ = solver.last_schedule()
        fast = loc == "local" and direction == "maximize" and form == "indexed" and kind == "score"
        if shape != (4200, 90):  # the short wide case may well stay on the per-wave layouts
            assert band == 1 and groups == 0, (solver.last_schedule(), fast)
        ref = O.OracleResult(O.make_config(locality=LOCS[loc][1], direction=odir, gap=gap, goal=goal,
                                           lim_s=lim_s, lim_t=lim_t), S)
        if kind == "score":
            assert np.float32(res[0]).tobytes() == ref.score.tobytes()
        else:
            assert np.float32(res[0].score).tobytes() == ref.score.tobytes()
            assert np.array_equal(np.asarray(res[0].path, dtype=np.int32), ref.path())


def test_long_mode_several_problems_and_short_ones():
    """Three long problems (bands taken in one ticket order across them)
    beside 200 short ones in the same solve."""
    rng = np.random.default_rng(85)
    ls = np.concatenate([[3000, 2500, 1900], rng.integers(10, 80, 200)]).astype(np.int32)
    lt = np.concatenate([[1800, 2200, 2600], rng.integers(10, 80, 200)]).astype(np.int32)
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, 5)
    table = sym_table(rng, 5)
    for kind, codomain, goal in (("score", pa.Score, O.GOAL_SCORE), ("solution", pa.Solution, O.GOAL_ONE)):
        solver = make_solver("local", ("affine", 4, 1), "maximize", codomain, 4096, 4096)
        res = solver.solve_ragged(kind, "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot,
                                  table=table)
        assert solver.last_schedule()[1] == 3
        for k in [0, 1, 2, 3, 100, 202]:
            S = dense_of("indexed", k, cs, ct, os_, ot, ls, lt, table)
            ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=("affine", 4, 1), goal=goal,
                                               lim_s=4096, lim_t=4096), S)
            got = res[k] if kind == "score" else res[k].score
            assert np.float32(got).tobytes() == ref.score.tobytes(), k
            if kind == "solution":
                assert np.array_equal(np.asarray(res[k].path, dtype=np.int32), ref.path()), k


def test_long_mode_repeats_and_graph_replay():
    """The same long problem solved repeatedly (the tagged hand-off words are
    cleared before every launch, also in a replayed graph), with new codes in
    between."""
    rng = np.random.default_rng(86)
    ls = np.array([2300], np.int32)
    lt = np.array([1700], np.int32)
    solver = make_solver("global", ("affine", 3, 1), "maximize", pa.Solution, 4096, 2048)
    for rep in range(4):
        cs, ct, os_, ot = ragged_codes(rng, ls, lt, 4) if rep % 2 == 0 else (cs, ct, os_, ot)
        table = sym_table(np.random.default_rng(7), 4)
        res = solver.solve_ragged("solution", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot,
                                  table=table)
        S = dense_of("indexed", 0, cs, ct, os_, ot, ls, lt, table)
        ref = O.OracleResult(O.make_config(locality=O.GLOBAL, gap=("affine", 3, 1), goal=O.GOAL_ONE,
                                           lim_s=4096, lim_t=2048), S)
        assert np.float32(res[0].score).tobytes() == ref.score.tobytes(), rep
        assert np.array_equal(np.asarray(res[0].path, dtype=np.int32), ref.path()), rep


@pytest.mark.parametrize("loc", list(LOCS))
@pytest.mark.parametrize("neg_zero", [False, True])
def test_long_mode_fold_choice(loc, neg_zero):
    """Traceback codes from max folds and equality tests (finite similarities,
    no -0) and from strict compare-and-select (a -0 in the table) give the
    oracle's score and path; integer similarities make ties common."""
    rng = np.random.default_rng(zlib.crc32(repr((loc, neg_zero, "fold")).encode()))
    ls = np.array([1100], np.int32)
    lt = np.array([1600], np.int32)
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, 5)
    table = sym_table(rng, 5, -2, 3)
    if neg_zero:
        table[0, 1] = table[1, 0] = np.float32(-0.0)
    S = dense_of("indexed", 0, cs, ct, os_, ot, ls, lt, table)
    for gap in (("linear", 1), ("affine", 3, 1)):
        solver = make_solver(loc, gap, "maximize", pa.Solution, 2048, 2048)
        res = solver.solve_ragged("solution", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot,
                                  table=table)
        assert solver.last_schedule()[1] == 1
        ref = O.OracleResult(O.make_config(locality=LOCS[loc][1], gap=gap, goal=O.GOAL_ONE, lim_s=2048,
                                           lim_t=2048), S)
        assert np.float32(res[0].score).tobytes() == ref.score.tobytes(), gap
        assert np.array_equal(np.asarray(res[0].path, dtype=np.int32), ref.path()), gap


@pytest.mark.parametrize("loc", ["local", "global"])
def test_long_pair_tape_gap_gradients(loc):
    """A tape of a long binary pair (pa_forward_ragged, gap gradients only)
    runs outside long mode, whose step-packed codes only a solve's walk
    reads: its backward walks the per-lane layout, and the gap gradients
    count the gap runs of the oracle's path."""
    import ctypes

    import torch

    from pyalign_amd import algorithm
    from pyalign_amd._lib import check, lib
    from tests.test_gpu_autograd import gap_counts
    rng = np.random.default_rng(zlib.crc32(repr((loc, "tape")).encode()))
    ls = np.array([2300], np.int32)
    lt = np.array([1900], np.int32)
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, 4)
    solver = make_solver(loc, ("affine", 3, 1), "maximize", pa.Solution, 4096, 2048)
    rg, keep = algorithm.make_ragged("binary", ls, lt, codes_s=cs.astype(np.uint32), codes_t=ct.astype(np.uint32),
                                     off_s=os_, off_t=ot, eq=1.0, ne=-1.0)
    scores = torch.empty(1, device="cuda")
    h = ctypes.c_void_p()
    check(lib().pa_forward_ragged(solver._h, ctypes.byref(rg), ctypes.c_void_p(scores.data_ptr()), ctypes.byref(h)))
    check(lib().pa_solver_sync(solver._h))
    g = torch.ones(1, device="cuda")
    gg = torch.zeros(2, device="cuda")
    check(lib().pa_backward(h, ctypes.c_void_p(g.data_ptr()), None, None, None, ctypes.c_void_p(gg.data_ptr())))
    lib().pa_tape_free(h)
    S = dense_of("binary", 0, cs, ct, os_, ot, ls, lt, eq=1.0, ne=-1.0)
    ref = O.OracleResult(O.make_config(locality=LOCS[loc][1], gap=("affine", 3, 1), goal=O.GOAL_ONE,
                                       lim_s=4096, lim_t=2048), S)
    assert np.float32(scores[0].item()).tobytes() == ref.score.tobytes()
    runs, ext = gap_counts(ref.path().tolist(), LOCS[loc][1])
    assert gg.cpu().tolist() == [-float(runs), -float(ext)]
    del keep
