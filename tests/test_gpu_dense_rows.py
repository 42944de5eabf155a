"""Dense similarity input with rows staged through LDS (every non-empty
problem at least 8 columns wide, unit column stride): results must equal the
oracle bit for bit, and equal the per-cell load path of the same kernel
(PYALIGN_AMD_NO_DENSE_STAGING) in scores, paths and gradients."""

import typing
import zlib

import numpy as np
import pytest
import torch

import oracle as O
import pyalign_amd as pa
from pyalign_amd import algorithm, torch_ops

pytestmark = pytest.mark.gpu

LOCS = {"local": (algorithm.Locality.LOCAL, O.LOCAL), "global": (algorithm.Locality.GLOBAL, O.GLOBAL),
        "semiglobal": (algorithm.Locality.SEMIGLOBAL, O.SEMIGLOBAL)}


def make_solver(loc, gap, direction, codomain, lim):
    d = algorithm.Direction.MAXIMIZE if direction == "maximize" else algorithm.Direction.MINIMIZE
    g = pa.LinearGapCost(gap[1]) if gap[0] == "linear" else pa.AffineGapCost(gap[1], gap[2])
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=d, return_dup=False,
        codomain=pa.solve.Codomain(codomain), locality=LOCS[loc][0], gap_cost=g))
    return algorithm.m16.create_solver(lim, lim, opts)


def bag(rng, n, lo, hi):
    ls = rng.integers(lo, hi + 1, n).astype(np.int32)
    lt = rng.integers(lo, hi + 1, n).astype(np.int32)
    ls[::17] = 0  # empty problems stay allowed
    sizes = ls.astype(np.int64) * lt
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    dense = (rng.integers(-4, 5, int(sizes.sum())) / 2).astype(np.float32)
    return ls, lt, off, dense


def solve(loc, gap, direction, kind, ls, lt, off, dense, monkeypatch, staged):
    with monkeypatch.context() as m:
        if not staged:
            m.setenv("PYALIGN_AMD_NO_DENSE_STAGING", "1")
        solver = make_solver(loc, gap, direction, pa.Solution if kind == "solution" else pa.Score, 512)
    return solver.solve_ragged(kind, "dense", ls, lt, dense=dense, dense_off=off)


@pytest.mark.parametrize("loc", list(LOCS))
@pytest.mark.parametrize("gap", [("linear", 1), ("affine", 3, 1)], ids=str)
@pytest.mark.parametrize("direction", ["maximize", "minimize"])
def test_staged_rows_match_oracle_and_per_cell_loads(loc, gap, direction, monkeypatch):
    rng = np.random.default_rng(zlib.crc32(repr((loc, gap, direction, "drows")).encode()))
    ls, lt, off, dense = bag(rng, 60, 8, 300)
    odir = O.MAXIMIZE if direction == "maximize" else O.MINIMIZE
    for kind, goal in (("score", O.GOAL_SCORE), ("solution", O.GOAL_ONE)):
        got = solve(loc, gap, direction, kind, ls, lt, off, dense, monkeypatch, True)
        ref_path = solve(loc, gap, direction, kind, ls, lt, off, dense, monkeypatch, False)
        for k in range(len(ls)):
            a, b = int(ls[k]), int(lt[k])
            S = dense[off[k]:off[k] + a * b].reshape(a, b)
            ref = O.OracleResult(O.make_config(locality=LOCS[loc][1], direction=odir, gap=gap, goal=goal,
                                               lim_s=512, lim_t=512), S)
            if kind == "score":
                assert np.float32(got[k]).tobytes() == ref.score.tobytes(), k
                assert np.float32(got[k]).tobytes() == np.float32(ref_path[k]).tobytes(), k
            elif ref.has_seed:
                assert np.float32(got[k].score).tobytes() == ref.score.tobytes(), k
                assert np.array_equal(np.asarray(got[k].path, dtype=np.int32), ref.path()), k
                assert np.array_equal(np.asarray(got[k].path), np.asarray(ref_path[k].path)), k


def test_differentiable_workload_gradients_identical(monkeypatch):
    """Embedding similarities of the differentiable workload's shape (lengths
    64..512 in a padded [N, 512, 512] tensor): scores and every gradient equal
    with and without staging."""
    rng = np.random.default_rng(47)
    n = 64
    ls = rng.integers(64, 513, n).astype(np.int32)
    lt = rng.integers(64, 513, n).astype(np.int32)
    g = torch.Generator(device="cpu").manual_seed(47)
    A = (torch.randn(n, 512, 64, generator=g) / 8).cuda()
    B = (torch.randn(n, 512, 64, generator=g) / 8).cuda()
    S = torch.bmm(A, B.transpose(1, 2).contiguous()).contiguous()

    def run(staged):
        torch_ops._SOLVERS.clear()
        with monkeypatch.context() as m:
            if not staged:
                m.setenv("PYALIGN_AMD_NO_DENSE_STAGING", "1")
            Sg = S.clone().requires_grad_(True)
            go = torch.tensor(1.0, requires_grad=True)
            ge = torch.tensor(0.2, requires_grad=True)
            s = torch_ops.align_score(Sg, ls, lt, go, ge, locality="global")
        s.sum().backward()
        torch_ops._SOLVERS.clear()
        return [s.detach().cpu().numpy(), Sg.grad.cpu().numpy(), go.grad.numpy().reshape(1), ge.grad.numpy().reshape(1)]

    for x, y in zip(run(True), run(False)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("loc", list(LOCS))
def test_staged_rows_all_optimal_alignments(loc, monkeypatch):
    """All-optimal tracebacks (8 tie bits per cell) from staged rows equal the
    per-cell path's: half-integer similarities make ties common, and the first
    alignments of every problem's enumeration must agree in order and maps."""
    rng = np.random.default_rng(zlib.crc32(repr((loc, "drows-all")).encode()))
    ls, lt, off, dense = bag(rng, 24, 8, 40)
    runs = []
    for staged in (True, False):
        with monkeypatch.context() as m:
            if not staged:
                m.setenv("PYALIGN_AMD_NO_DENSE_STAGING", "1")
            solver = make_solver(loc, ("affine", 3, 1), "maximize", typing.Iterator[pa.Alignment], 64)
        its = solver.solve_ragged("alignment_iterator", "dense", ls, lt, dense=dense, dense_off=off)
        out = []
        for it in its:
            seq = []
            for _ in range(6):
                a = it.next()
                if a is None:
                    break
                seq.append((np.float32(a.score).tobytes(), np.asarray(a.s_to_t).tobytes()))
            out.append(seq)
        runs.append(out)
    assert runs[0] == runs[1]
    assert sum(len(x) > 1 for x in runs[0]) > 0  # some problems have tied optima


@pytest.mark.parametrize("loc", ["global", "semiglobal"])
def test_staged_rows_in_band_mode(loc, monkeypatch):
    """Two long dense problems (band mode: strips on concurrent wavefronts,
    rows staged through LDS) among short ones: scores and paths equal the
    oracle's and the per-cell path's."""
    rng = np.random.default_rng(zlib.crc32(repr((loc, "drows-band")).encode()))
    ls = np.concatenate([rng.integers(10, 60, 150), [2600, 2900], rng.integers(10, 60, 150)]).astype(np.int32)
    lt = np.concatenate([rng.integers(10, 60, 150), [900, 1000], rng.integers(10, 60, 150)]).astype(np.int32)
    sizes = ls.astype(np.int64) * lt
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    dense = (rng.integers(-4, 5, int(sizes.sum())) / 2).astype(np.float32)
    res = []
    for staged in (True, False):
        with monkeypatch.context() as m:
            if not staged:
                m.setenv("PYALIGN_AMD_NO_DENSE_STAGING", "1")
            solver = make_solver(loc, ("affine", 3, 1), "maximize", pa.Solution, 4096)
        res.append(solver.solve_ragged("solution", "dense", ls, lt, dense=dense, dense_off=off))
        assert solver.last_schedule()[1] >= 1, solver.last_schedule()
    for k in [0, 149, 150, 151, 152, len(ls) - 1]:
        a, b = int(ls[k]), int(lt[k])
        S = dense[off[k]:off[k] + a * b].reshape(a, b)
        ref = O.OracleResult(O.make_config(locality=LOCS[loc][1], gap=("affine", 3, 1), goal=O.GOAL_ONE,
                                           lim_s=4096, lim_t=4096), S)
        for r in res:
            assert np.float32(r[k].score).tobytes() == ref.score.tobytes(), k
            assert np.array_equal(np.asarray(r[k].path, dtype=np.int32), ref.path()), k
