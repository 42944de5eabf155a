"""hipGraph replay of repeated launch sequences.

A solver captures a launch sequence the second time it sees it (same
shapes, same buffers) and replays the graph afterwards.  Every replay is
checked against an eager solve of the same inputs (graphs disabled), with
other eager GPU work and synchronisations between the replays, and the
inputs changed in place between replays (the graph must read the new data);
a sample of every checked replay's scores and paths against the oracle.
"""

import ctypes

import numpy as np
import pytest
import torch

import oracle as O
import pyalign_amd as pa
from pyalign_amd import _lib, algorithm
from pyalign_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def local_affine_solver(codomain, lim):
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(codomain), locality=algorithm.Locality.LOCAL, gap_cost=pa.AffineGapCost(10, 1)))
    return algorithm.m16.create_solver(lim, lim, opts)


def batch(rng, n, lo, hi, K=20):
    ls = rng.integers(lo, hi + 1, n).astype(np.int32)
    lt = rng.integers(lo, hi + 1, n).astype(np.int32)
    cs = rng.integers(0, K, int(ls.sum())).astype(np.uint8)
    ct = rng.integers(0, K, int(lt.sum())).astype(np.uint8)
    os_ = np.concatenate([[0], np.cumsum(ls.astype(np.int64))[:-1]])
    ot = np.concatenate([[0], np.cumsum(lt.astype(np.int64))[:-1]])
    t = rng.integers(-4, 4, size=(K, K))
    t = np.triu(t) + np.triu(t, 1).T
    np.fill_diagonal(t, rng.integers(4, 12, K))
    return ls, lt, cs, ct, os_, ot, t.astype(np.float32)


def device_scores(solver, ls, lt, d_cs, d_ct, os_, ot, table, d_out):
    rg, keep = algorithm.make_ragged("indexed", ls, lt, codes_s=d_cs.data_ptr(), codes_t=d_ct.data_ptr(),
                                     off_s=os_, off_t=ot, table=table, on_device=True)
    check(lib().pa_solve_ragged_scores_device(solver._h, ctypes.byref(rg), _lib.PA_SCHED_AUTO,
                                              ctypes.c_void_p(d_out.data_ptr())))
    check(lib().pa_solver_sync(solver._h))
    del keep
    return d_out.cpu().numpy().copy()


@pytest.mark.parametrize("lengths", [(40, 400), (3000, 4000)], ids=["many-short", "band-long"])
def test_replayed_device_scores_match_eager(lengths):
    rng = np.random.default_rng(91)
    n = 2000 if lengths[1] < 1000 else 3
    ls, lt, cs, ct, os_, ot, table = batch(rng, n, *lengths)
    lim = 4096
    graph_solver = local_affine_solver(pa.Score, lim)
    eager_solver = local_affine_solver(pa.Score, lim)
    eager_solver.set_graphs(False)
    d_cs = torch.from_numpy(cs).cuda()
    d_ct = torch.from_numpy(ct).cuda()
    d_out = torch.empty(n, dtype=torch.float32, device="cuda")
    other = batch(rng, 50, 20, 90)
    for rep in range(6):
        if rep == 3:  # new data in the same buffers: replays must read it
            cs = rng.integers(0, 20, cs.size).astype(np.uint8)
            d_cs.copy_(torch.from_numpy(cs).cuda())
            torch.cuda.synchronize()
        got = device_scores(graph_solver, ls, lt, d_cs, d_ct, os_, ot, table, d_out)
        want = device_scores(eager_solver, ls, lt, d_cs, d_ct, os_, ot, table,
                             torch.empty(n, dtype=torch.float32, device="cuda"))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), rep
        if rep in (2, 3, 5):  # replays (and the capture after new data) against the oracle
            for k in range(0, n, max(1, n // 25)):
                S = table[np.ix_(cs[os_[k]:os_[k] + ls[k]], ct[ot[k]:ot[k] + lt[k]])]
                ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=("affine", 10, 1), goal=O.GOAL_SCORE,
                                                   lim_s=lim, lim_t=lim), S)
                assert got[k].tobytes() == ref.score.tobytes(), (rep, k)
        if lengths[1] >= 1000:
            assert graph_solver.last_schedule()[1] == n
        # unrelated eager work between replays
        o = eager_solver.solve_ragged("score", "indexed", other[0], other[1], codes_s=other[2], codes_t=other[3],
                                      off_s=other[4], off_t=other[5], table=other[6])
        assert len(o) == 50
        torch.cuda.synchronize()


def test_replayed_solutions_match_eager():
    rng = np.random.default_rng(92)
    ls, lt, cs, ct, os_, ot, table = batch(rng, 300, 20, 200)
    graph_solver = local_affine_solver(pa.Solution, 256)
    eager_solver = local_affine_solver(pa.Solution, 256)
    eager_solver.set_graphs(False)
    payload = dict(codes_s=cs, codes_t=ct, off_s=os_, off_t=ot, table=table)
    want = eager_solver.solve_ragged("solution", "indexed", ls, lt, **payload)
    for rep in range(5):
        got = graph_solver.solve_ragged("solution", "indexed", ls, lt, **payload)
        for g, w in zip(got, want):
            assert np.float32(g.score).tobytes() == np.float32(w.score).tobytes(), rep
            assert np.array_equal(np.asarray(g.path), np.asarray(w.path)), rep
        for k in range(rep, len(ls), 23):  # replayed paths against the oracle
            S = table[np.ix_(cs[os_[k]:os_[k] + ls[k]], ct[ot[k]:ot[k] + lt[k]])]
            ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=("affine", 10, 1), goal=O.GOAL_ONE,
                                               lim_s=256, lim_t=256), S)
            assert np.float32(got[k].score).tobytes() == ref.score.tobytes(), (rep, k)
            assert np.array_equal(np.asarray(got[k].path, dtype=np.int32), ref.path()), (rep, k)
        x = torch.randn(1000, device="cuda").sum().item()  # eager torch work between replays
        assert np.isfinite(x)
