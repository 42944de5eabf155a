"""Checkpointed differentiation on the GPU.

Long pairs keep only strip boundary rows (and the path's end) in the
forward; the backward recomputes their tracebacks from those rows: whole
paths at once where the scratch budget allows, else one strip at a time
(from the seed's strip up), continuing the walk through each.  Scores and every gradient must be
bit-identical to the run that keeps the full traceback, with the default
budget and with a budget small enough to force the strip-by-strip passes,
and the schedule must show the pairs checkpointed.
"""

import numpy as np
import pytest
import torch

from pyalign_amd import torch_ops
from pyalign_amd.torch_ops import align_score, dtw_distance

pytestmark = pytest.mark.gpu


def last_schedule():
    return list(torch_ops._SOLVERS.values())[-1].last_schedule()


def walks(rng, n, lo, hi, dim=3):
    lx = rng.integers(lo, hi + 1, n).astype(np.int32)
    ly = rng.integers(lo, hi + 1, n).astype(np.int32)
    X = np.zeros((n, int(lx.max()), dim), np.float32)
    Y = np.zeros((n, int(ly.max()), dim), np.float32)
    for k in range(n):
        X[k, :lx[k]] = np.cumsum(rng.normal(0, 0.1, (lx[k], dim)), 0)
        Y[k, :ly[k]] = np.cumsum(rng.normal(0, 0.1, (ly[k], dim)), 0)
    return X, Y, lx, ly


def dtw_run(X, Y, lx, ly, checkpoint):
    Xt = torch.tensor(X, device="cuda", requires_grad=True)
    Yt = torch.tensor(Y, device="cuda", requires_grad=True)
    d = dtw_distance(Xt, Yt, lx, ly, checkpoint=checkpoint)
    sched = last_schedule()
    d.sum().backward()
    return d.detach().cpu().numpy(), Xt.grad.cpu().numpy(), Yt.grad.cpu().numpy(), sched


# scratch words: the default (free memory: one recomputation launch), room
# for a few whole paths at a time (chunks recomputed while the previous
# chunk's walk runs on a second stream), or less than the whole traceback of
# the longer pairs (those walk strip by strip)
SCRATCH = {"auto": None, "chunks": "1000000", "strips": "150000"}


def with_scratch(monkeypatch, scratch, fn):
    torch_ops._SOLVERS.clear()  # the budget is read when a solver is made
    with monkeypatch.context() as m:
        if SCRATCH[scratch]:
            m.setenv("PYALIGN_AMD_CKPT_SCRATCH_WORDS", SCRATCH[scratch])
        out = fn()
    torch_ops._SOLVERS.clear()
    return out


@pytest.mark.parametrize("scratch", list(SCRATCH))
@pytest.mark.parametrize("dim", [1, 3])
def test_dtw_checkpointed_equals_full_traceback(dim, scratch, monkeypatch):
    rng = np.random.default_rng(101 + dim)
    X, Y, lx, ly = walks(rng, 12, 300, 1700, dim)
    full = dtw_run(X, Y, lx, ly, -1)
    ck = with_scratch(monkeypatch, scratch, lambda: dtw_run(X, Y, lx, ly, 1000))
    assert full[3][2] == 0 and ck[3][2] > 0, (full[3], ck[3])
    for a, b in zip(full[:3], ck[:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def align_run(S, ls, lt, checkpoint, gap, locality="global"):
    St = torch.tensor(S, device="cuda", requires_grad=True)
    go = torch.tensor(1.0, requires_grad=True)
    ge = torch.tensor(0.2, requires_grad=True)
    s = align_score(St, ls, lt, go, ge if gap == "affine" else 0.0, locality=locality, gap=gap,
                    checkpoint=checkpoint)
    sched = last_schedule()
    s.sum().backward()
    out = [s.detach().cpu().numpy(), St.grad.cpu().numpy(), go.grad.numpy().reshape(1)]
    if gap == "affine":
        out.append(ge.grad.numpy().reshape(1))
    return out, sched


@pytest.mark.parametrize("scratch", list(SCRATCH))
@pytest.mark.parametrize("gap", ["affine", "linear"])
@pytest.mark.parametrize("locality", ["global", "local", "semiglobal"])
def test_align_checkpointed_equals_full_traceback(locality, gap, scratch, monkeypatch):
    """Local and semiglobal paths end where the forward's seed search put
    them (inside the matrix, or on its last row or column): the strip walk
    starts in the seed's strip."""
    rng = np.random.default_rng(111)
    n = 8
    ls = rng.integers(200, 1400, n).astype(np.int32)
    lt = rng.integers(200, 1400, n).astype(np.int32)
    S = (rng.standard_normal((n, int(ls.max()), int(lt.max()))) / 4).astype(np.float32)
    if locality == "local":
        S -= np.float32(0.1)  # local paths shorter than the matrix, ending inside it
    if scratch == "strips":  # some checkpointed pairs walk strip by strip: traceback words above the scratch
        ck_pairs = (np.maximum(ls, lt) > 600) & (ls > 512)
        words = (lt.astype(np.int64) + 63) * 64 * ((ls + 511) // 512)
        assert (words[ck_pairs] > int(SCRATCH[scratch])).sum() >= 2
    full, sf = align_run(S, ls, lt, -1, gap, locality)
    ck, sc = with_scratch(monkeypatch, scratch, lambda: align_run(S, ls, lt, 600, gap, locality))
    assert sf[2] == 0 and sc[2] > 0, (sf, sc)
    for a, b in zip(full, ck):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_local_checkpointed_without_a_positive_cell(monkeypatch):
    """A long local pair whose similarities are all negative has no path (no
    seed): its checkpointed backward adds nothing, strip by strip too."""
    rng = np.random.default_rng(113)
    ls = np.array([1500, 300], np.int32)
    lt = np.array([1800, 250], np.int32)  # 3 strips of 1862 x 64 words: more than the scratch
    S = (rng.standard_normal((2, 1500, 1800)) / 4).astype(np.float32)
    S[0] = -np.abs(S[0]) - np.float32(0.5)
    full, _ = align_run(S, ls, lt, -1, "affine", "local")
    ck, sc = with_scratch(monkeypatch, "strips", lambda: align_run(S, ls, lt, 600, "affine", "local"))
    assert sc[2] > 0
    assert full[0][0] == 0.0 and not full[1][0].any()
    for a, b in zip(full, ck):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_minimizing_linear_checkpointed():
    """Global minimisation with a free linear gap over a dense cost, through
    the checkpointed backward, equals the full-traceback run."""
    rng = np.random.default_rng(112)
    n = 4
    ls = rng.integers(900, 1300, n).astype(np.int32)
    lt = rng.integers(900, 1300, n).astype(np.int32)
    C = rng.random((n, int(ls.max()), int(lt.max()))).astype(np.float32)

    def run(ck):
        Ct = torch.tensor(C, device="cuda", requires_grad=True)
        s = align_score(Ct, ls, lt, 0.5, 0.0, locality="global", direction="minimize", gap="linear",
                        checkpoint=ck)
        sched = last_schedule()
        s.sum().backward()
        return (s.detach().cpu().numpy(), Ct.grad.cpu().numpy()), sched

    (a, sa), (b, sb) = run(-1), run(500)
    assert sa[2] == 0 and sb[2] > 0
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_dtw_full_size_batch():
    """8,192 DTW pairs, lengths U[128,4096] (3-D random walks): distances and
    gradients with checkpointed strips equal the full-traceback run, and 66
    pairs (the longest, the shortest and 64 spread over the batch, most of
    them checkpointed) match the oracle's distance and optimal path."""
    import bench
    import oracle as O
    import pyalign_amd as pa
    X, Y, lx, ly = bench.c5_inputs()
    full = dtw_run(X, Y, lx, ly, -1)
    ck = dtw_run(X, Y, lx, ly, 2048)
    assert full[3][2] == 0 and ck[3][2] > 4000
    for a, b in zip(full[:3], ck[:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    d, gx = ck[0], ck[1]
    pf = pa.series()
    lims = (pa.solve.SizeRounder()(int(lx.max())), pa.solve.SizeRounder()(int(ly.max())))
    longest = int(np.argmax(np.maximum(lx, ly)))
    short = int(np.argmin(np.maximum(lx, ly)))
    # the longest and shortest pairs and 64 more spread over the batch
    sample = [longest, short] + list(range(61, len(lx), len(lx) // 64))[:64]
    for k in sample:
        p = pf.new_problem(X[k, :lx[k]], Y[k, :ly[k]])
        ref = O.OracleResult(O.make_config(solver_type=O.DTW, direction=O.MINIMIZE, goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), p.matrix)
        assert np.float32(d[k]).tobytes() == ref.score.tobytes(), k
        ex = np.zeros_like(X[k, :lx[k]])
        for u, v in ref.path().tolist():
            if u >= 0 and v >= 0:
                ex[u] += 2 * (X[k, u] - Y[k, v])
        np.testing.assert_allclose(gx[k, :lx[k]], ex, atol=1e-3)
