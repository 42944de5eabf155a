"""The allocation-free score solve (pa_solve_ragged_scores_noalloc) against
the allocating one (pa_solve_ragged_scores_device) on the same device-resident
inputs, called through the C ABI: two calls in a row, the second on a smaller
batch, bit-identical scores; batches beyond the scratch size are refused.
"""

import ctypes

import numpy as np
import pytest
import torch

import pyalign_amd as pa
from pyalign_amd import _lib, algorithm
from pyalign_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _solver(loc, gap, lim_s, lim_t, direction=algorithm.Direction.MAXIMIZE, dtw=False):
    opts = dict(batch=True, direction=direction, return_dup=False, codomain=pa.solve.Codomain(pa.Score))
    if dtw:
        opts["solver"] = algorithm.Type.DTW
    else:
        opts.update(solver=algorithm.Type.ALIGNMENT, locality=loc, gap_cost=gap)
    return algorithm.m16.create_solver(lim_s, lim_t, algorithm.m16.create_options(opts))


def _both(solver, rg, n):
    """(allocating, allocation-free) scores of one ragged batch."""
    stream = torch.cuda.current_stream()
    ref = torch.full((n,), float("nan"), device="cuda")
    got = torch.full((n,), float("nan"), device="cuda")
    solver.set_stream(stream.cuda_stream)
    check(lib().pa_solve_ragged_scores_device(solver._h, ctypes.byref(rg), _lib.PA_SCHED_AUTO,
                                              ctypes.c_void_p(ref.data_ptr())))
    check(lib().pa_solve_ragged_scores_noalloc(solver._h, ctypes.byref(rg), _lib.PA_SCHED_AUTO,
                                               ctypes.c_void_p(got.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _indexed_bag(rng, n, lo, hi, K=20):
    ls = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(np.int32)
    lt = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(np.int32)
    cs = rng.integers(0, K, int(ls.sum())).astype(np.uint8)
    ct = rng.integers(0, K, int(lt.sum())).astype(np.uint8)
    off_s = np.concatenate([[0], np.cumsum(ls.astype(np.int64))[:-1]])
    off_t = np.concatenate([[0], np.cumsum(lt.astype(np.int64))[:-1]])
    return ls, lt, torch.from_numpy(cs).cuda(), torch.from_numpy(ct).cuda(), off_s, off_t


def test_noalloc_equals_allocating_entry_twice_then_smaller():
    rng = np.random.default_rng(11)
    K = 20
    table = rng.integers(-4, 4, size=(K, K))
    table = np.triu(table) + np.triu(table, 1).T
    np.fill_diagonal(table, rng.integers(4, 12, size=K))
    table = table.astype(np.float32)
    solver = _solver(algorithm.Locality.LOCAL, pa.AffineGapCost(10, 1), 2000, 2000)
    ls, lt, dcs, dct, off_s, off_t = _indexed_bag(rng, 3000, 50, 2000)
    torch.cuda.synchronize()
    for n in (3000, 700):  # the second call on a smaller batch
        rg, keep = algorithm.make_ragged("indexed", ls[:n], lt[:n], codes_s=dcs.data_ptr(), codes_t=dct.data_ptr(),
                                         off_s=off_s[:n], off_t=off_t[:n], table=table, on_device=True)
        ref, got = _both(solver, rg, n)
        assert not np.isnan(got).any()
        assert _same(ref, got), n
        del keep


@pytest.mark.parametrize("case", ["global_binary_linear", "semiglobal_dense_affine", "dtw_series", "wide_table"])
def test_noalloc_other_forms(case):
    rng = np.random.default_rng(len(case))
    n = 400
    ls = rng.integers(1, 300, n).astype(np.int32)
    lt = rng.integers(1, 300, n).astype(np.int32)
    off_s = np.concatenate([[0], np.cumsum(ls.astype(np.int64))[:-1]])
    off_t = np.concatenate([[0], np.cumsum(lt.astype(np.int64))[:-1]])
    if case == "global_binary_linear":
        solver = _solver(algorithm.Locality.GLOBAL, pa.LinearGapCost(0.5), 512, 512)
        cs = torch.from_numpy(rng.integers(0, 4, int(ls.sum())).astype(np.int32)).cuda()
        ct = torch.from_numpy(rng.integers(0, 4, int(lt.sum())).astype(np.int32)).cuda()
        mk = lambda m: algorithm.make_ragged("binary", ls[:m], lt[:m], codes_s=cs.data_ptr(),  # noqa: E731
                                             codes_t=ct.data_ptr(), off_s=off_s[:m], off_t=off_t[:m], eq=1.0,
                                             ne=-1.0, on_device=True)
    elif case == "semiglobal_dense_affine":
        solver = _solver(algorithm.Locality.SEMIGLOBAL, pa.AffineGapCost(2.0, 0.5), 512, 512)
        dense_off = np.concatenate([[0], np.cumsum(ls.astype(np.int64) * lt)[:-1]])
        S = torch.from_numpy(rng.standard_normal(int((ls.astype(np.int64) * lt).sum())).astype(np.float32)).cuda()
        mk = lambda m: algorithm.make_ragged("dense", ls[:m], lt[:m], dense=S.data_ptr(),  # noqa: E731
                                             dense_off=dense_off[:m], on_device=True)
    elif case == "dtw_series":
        solver = _solver(None, None, 512, 512, direction=algorithm.Direction.MINIMIZE, dtw=True)
        xs = torch.from_numpy(rng.standard_normal((int(ls.sum()), 3)).astype(np.float32)).cuda()
        xt = torch.from_numpy(rng.standard_normal((int(lt.sum()), 3)).astype(np.float32)).cuda()
        mk = lambda m: algorithm.make_ragged("series", ls[:m], lt[:m], codes_s=xs.data_ptr(),  # noqa: E731
                                             codes_t=xt.data_ptr(), off_s=off_s[:m], off_t=off_t[:m], dim=3,
                                             on_device=True)
    else:
        K = 300
        table = rng.integers(-4, 6, size=(K, K)).astype(np.float32)
        solver = _solver(algorithm.Locality.LOCAL, pa.AffineGapCost(6, 1), 512, 512)
        cs = torch.from_numpy(rng.integers(0, K, int(ls.sum())).astype(np.int16)).cuda()
        ct = torch.from_numpy(rng.integers(0, K, int(lt.sum())).astype(np.int16)).cuda()
        mk = lambda m: algorithm.make_ragged("indexed", ls[:m], lt[:m], codes_s=cs.data_ptr(),  # noqa: E731
                                             codes_t=ct.data_ptr(), off_s=off_s[:m], off_t=off_t[:m], table=table,
                                             on_device=True, code_bytes=2)
    torch.cuda.synchronize()
    for m in (n, n // 3):
        rg, keep = mk(m)
        ref, got = _both(solver, rg, m)
        assert _same(ref, got), (case, m)


def test_noalloc_refuses_batches_beyond_its_scratch():
    """1,200,000 one-cell problems need 86 MB of problem descriptors (72
    bytes each): refused before any launch.  (Every problem reads the same
    code: in bounds even if it were launched.)"""
    solver = _solver(algorithm.Locality.LOCAL, pa.LinearGapCost(1), 2000, 2000)
    codes = torch.zeros(2000, dtype=torch.uint8, device="cuda")
    n = 1200000
    ls = np.full(n, 1, np.int32)
    zeros = np.zeros(n, np.int64)
    rg, keep = algorithm.make_ragged("indexed", ls, ls, codes_s=codes.data_ptr(), codes_t=codes.data_ptr(),
                                     off_s=zeros, off_t=zeros, table=np.eye(4, dtype=np.float32), on_device=True)
    out = torch.empty(n, device="cuda")
    with pytest.raises(ValueError, match="bytes of device scratch"):
        check(lib().pa_solve_ragged_scores_noalloc(solver._h, ctypes.byref(rg), _lib.PA_SCHED_AUTO,
                                                   ctypes.c_void_p(out.data_ptr()), None))
    # a batch that fits still solves afterwards
    ls2 = np.full(4, 2000, np.int32)
    rg2, keep2 = algorithm.make_ragged("indexed", ls2, ls2, codes_s=codes.data_ptr(), codes_t=codes.data_ptr(),
                                       off_s=zeros[:4], off_t=zeros[:4], table=np.eye(4, dtype=np.float32),
                                       on_device=True)
    ref, got = _both(solver, rg2, 4)
    assert _same(ref, got)
    assert got[0] == 2000.0
