"""Differentiable optimal alignment scores (torch.autograd.Function).

The optimal score of an alignment problem is piecewise linear in the
similarity values and the gap costs, so on the optimal path its gradients
are an indicator and counts (no reference counterpart; the reference
computes no gradients):

* ``align_score``: dScore/dS[k, i, j] = 1 on the matched cells of problem
  k's optimal alignment, dScore/d(gap_open) = -(number of gap runs) and
  dScore/d(gap_extend) = -(sum of run lengths - 1) when maximising, for the
  affine cost w(k) = open + extend (k - 1) of ``AffineGapCost(open, extend)``
  (signs flip when minimising).  With ``gap="linear"``, ``gap_open`` is the
  per-symbol cost and its gradient is minus the total gap length.
* ``dtw_distance``: DTW on vector series under the squared Euclidean cost;
  dD/dx[i] = sum over path cells (i, j) of 2 (x[i] - y[j]), likewise for y.

Forward and backward both run on the GPU: the forward keeps the packed
traceback on the device and the backward kernel walks every optimal path
once.  At ties the gradient is the one of the reference's tie-broken path.
"""

import ctypes

import numpy as np
import torch

from . import algorithm
from ._lib import check, lib
from .gaps import AffineGapCost, LinearGapCost
from .solve import Codomain, Solution, SizeRounder

_LOC = {"local": algorithm.Locality.LOCAL, "global": algorithm.Locality.GLOBAL,
        "semiglobal": algorithm.Locality.SEMIGLOBAL}


class _Tape:
    def __init__(self, h, solver, keep):
        self._free = lib().pa_tape_free
        self.h = h
        self.solver = solver  # the tape borrows the solver and the payload
        self.keep = keep

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None


def _lengths(x, n):
    a = np.asarray(x.cpu() if torch.is_tensor(x) else x, dtype=np.int32).reshape(-1)
    if a.shape[0] != n:
        raise ValueError(f"expected {n} lengths, got {a.shape[0]}")
    return a


_SOLVERS = {}  # recent solvers by configuration (a tape keeps its own buffers)
_MAX_SOLVERS = 8


def _gap_cost(gap, gap_open, gap_extend):
    return AffineGapCost(float(gap_open), float(gap_extend)) if gap == "affine" else LinearGapCost(float(gap_open))


def _solver(kind, direction, locality, gap, gap_open, gap_extend, ls, lt, device=None):
    """A cached solver for this configuration.  The gap values are not part
    of the key: learnable gap costs change every optimisation step, and the
    solver takes the new values in place (same buffers, plan and scratch;
    only its border vectors are recomputed, on torch's current stream)."""
    rounder = SizeRounder()
    lim_s = rounder(int(ls.max()) if len(ls) else 0)
    lim_t = rounder(int(lt.max()) if len(lt) else 0)
    key = (kind, direction, locality, gap, lim_s, lim_t)
    solver = _SOLVERS.pop(key, None)
    if solver is None:
        solver = _new_solver(kind, direction, locality, gap, gap_open, gap_extend, lim_s, lim_t)
    elif kind != "dtw":
        solver.set_stream(torch.cuda.current_stream(device).cuda_stream)
        solver.set_gap_cost(_gap_cost(gap, gap_open, gap_extend))
    _SOLVERS[key] = solver  # most recent last
    while len(_SOLVERS) > _MAX_SOLVERS:
        _SOLVERS.pop(next(iter(_SOLVERS)))
    return solver


def _new_solver(kind, direction, locality, gap, gap_open, gap_extend, lim_s, lim_t):
    bits = algorithm.m16 if (lim_s < 16000 and lim_t < 16000) else algorithm.m32
    opts = dict(batch=True, return_dup=False, codomain=Codomain(Solution),
                direction=algorithm.Direction.__members__[direction.upper()])
    if kind == "dtw":
        opts["solver"] = algorithm.Type.DTW
    else:
        opts["solver"] = algorithm.Type.ALIGNMENT
        opts["locality"] = _LOC[locality]
        opts["gap_cost"] = _gap_cost(gap, gap_open, gap_extend)
    return bits.create_solver(lim_s, lim_t, bits.create_options(opts))


def _forward(solver, rg, scores, keep):
    """The forward pass on torch's current stream: kernels are ordered after
    the inputs' producers and before the scores' consumers, with no host
    wait (except to check a band-mode run's strip pipeline)."""
    solver.set_stream(torch.cuda.current_stream(scores.device).cuda_stream)
    h = ctypes.c_void_p()
    check(lib().pa_forward_ragged(solver._h, ctypes.byref(rg), ctypes.c_void_p(scores.data_ptr()), ctypes.byref(h)))
    tape = _Tape(h, solver, keep)
    if solver.last_schedule()[1] > 0:
        check(lib().pa_solver_sync(solver._h))
    return tape


class _AlignScore(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S, gap_open, gap_extend, len_s, len_t, locality, direction, gap, checkpoint):
        if not S.is_cuda or S.dtype != torch.float32 or S.dim() != 3:
            raise ValueError("S must be a float32 CUDA tensor of shape [N, Ls, Lt]")
        S = S.contiguous()
        n, Ls, Lt = S.shape
        ls = _lengths(len_s, n)
        lt = _lengths(len_t, n)
        if (ls > Ls).any() or (lt > Lt).any():
            raise ValueError("lengths exceed the padded shape of S")
        solver = _solver("align", direction, locality, gap, float(gap_open), float(gap_extend), ls, lt, S.device)
        solver.set_checkpoint(checkpoint)
        dense_off = np.arange(n, dtype=np.int64) * (Ls * Lt)
        rg, keep = algorithm.make_ragged("dense", ls, lt, dense=S.data_ptr(), dense_off=dense_off,
                                         dense_ld=Lt, on_device=True)
        scores = torch.empty(n, dtype=torch.float32, device=S.device)
        ctx.tape = _forward(solver, rg, scores, (rg, keep, S))
        ctx.shape = S.shape
        ctx.gap = gap
        return scores

    @staticmethod
    def backward(ctx, grad):
        grad = grad.contiguous().to(torch.float32)
        n = ctx.shape[0]
        dS = torch.zeros(ctx.shape, dtype=torch.float32, device=grad.device)
        dgap = torch.zeros((n, 2), dtype=torch.float32, device=grad.device)
        ctx.tape.solver.set_stream(torch.cuda.current_stream(grad.device).cuda_stream)
        check(lib().pa_backward(ctx.tape.h, ctypes.c_void_p(grad.data_ptr()), ctypes.c_void_p(dS.data_ptr()),
                                None, None, ctypes.c_void_p(dgap.data_ptr())))
        d_open = dgap[:, 0].sum()
        d_ext = dgap[:, 1].sum() if ctx.gap == "affine" else None
        return dS, d_open, d_ext, None, None, None, None, None, None


def align_score(S, len_s, len_t, gap_open=1.0, gap_extend=0.2, locality="global", direction="maximize",
                gap="affine", checkpoint=2048):
    """Optimal alignment scores of N problems given padded similarities S[N, Ls, Lt].

    gap_open / gap_extend may be floats or 0-dim tensors (requires_grad for
    gap gradients).  Returns float32 scores [N], identical to the solvers'
    ``Score`` results on the same inputs.  Pairs whose longer side exceeds
    ``checkpoint`` keep only strip boundary rows (and the path's end the
    forward found) and are recomputed strip by strip in the backward (-1:
    never).
    """
    if gap not in ("affine", "linear"):
        raise ValueError(gap)
    go = gap_open if torch.is_tensor(gap_open) else torch.tensor(float(gap_open))
    ge = gap_extend if torch.is_tensor(gap_extend) else torch.tensor(float(gap_extend))
    return _AlignScore.apply(S, go, ge, len_s, len_t, locality, direction, gap, int(checkpoint))


class _DTW(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, len_x, len_y, direction, checkpoint):
        for t in (x, y):
            if not t.is_cuda or t.dtype != torch.float32 or t.dim() != 3:
                raise ValueError("series must be float32 CUDA tensors [N, L, dim]")
        x = x.contiguous()
        y = y.contiguous()
        n, Lx, dim = x.shape
        if y.shape[0] != n or y.shape[2] != dim or not 1 <= dim <= 4:
            raise ValueError("series batches must agree in N and dim (1..4)")
        Ly = y.shape[1]
        lx = _lengths(len_x, n)
        ly = _lengths(len_y, n)
        if (lx > Lx).any() or (ly > Ly).any():
            raise ValueError("lengths exceed the padded series")
        solver = _solver("dtw", direction, None, None, 0.0, 0.0, lx, ly)
        solver.set_checkpoint(checkpoint)
        rg, keep = algorithm.make_ragged("series", lx, ly, codes_s=x.data_ptr(), codes_t=y.data_ptr(),
                                         off_s=np.arange(n, dtype=np.int64) * Lx,
                                         off_t=np.arange(n, dtype=np.int64) * Ly, dim=dim, on_device=True)
        scores = torch.empty(n, dtype=torch.float32, device=x.device)
        ctx.tape = _forward(solver, rg, scores, (rg, keep, x, y))
        ctx.shapes = (x.shape, y.shape)
        return scores

    @staticmethod
    def backward(ctx, grad):
        grad = grad.contiguous().to(torch.float32)
        dx = torch.zeros(ctx.shapes[0], dtype=torch.float32, device=grad.device)
        dy = torch.zeros(ctx.shapes[1], dtype=torch.float32, device=grad.device)
        ctx.tape.solver.set_stream(torch.cuda.current_stream(grad.device).cuda_stream)
        check(lib().pa_backward(ctx.tape.h, ctypes.c_void_p(grad.data_ptr()), None,
                                ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), None))
        return dx, dy, None, None, None, None


def dtw_distance(x, y, len_x, len_y, direction="minimize", checkpoint=2048):
    """DTW distances of N series pairs x[N, Lx, dim], y[N, Ly, dim] (squared
    Euclidean cost), differentiable with respect to x and y.  Pairs whose
    longer side exceeds ``checkpoint`` are recomputed strip by strip in the
    backward instead of keeping their traceback (-1: never)."""
    return _DTW.apply(x, y, len_x, len_y, direction, int(checkpoint))
