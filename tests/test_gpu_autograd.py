"""Gradients of optimal scores (torch.autograd) on the GPU.

The backward pass is pinned three ways: the path indicator equals the
matched cells of the oracle's optimal alignment, the gap gradients equal
the gap-run counts read off the oracle's path, and central finite
differences (eps 1e-3) agree on tie-free random draws.
"""

import numpy as np
import pytest
import torch

import oracle as O
import pyalign_amd as pa
from pyalign_amd.torch_ops import align_score, dtw_distance

pytestmark = pytest.mark.gpu

LOCS = {"global": O.GLOBAL, "local": O.LOCAL, "semiglobal": O.SEMIGLOBAL}


def c4_batch(rng, n, lo, hi, dim=64):
    """Embedding pairs: S = A B^T with N(0,1)/8 embeddings (float32)."""
    ls = rng.integers(lo, hi + 1, n)
    lt = rng.integers(lo, hi + 1, n)
    Ls, Lt = int(ls.max()), int(lt.max())
    S = np.zeros((n, Ls, Lt), np.float32)
    for k in range(n):
        A = (rng.standard_normal((ls[k], dim)) / 8).astype(np.float32)
        B = (rng.standard_normal((lt[k], dim)) / 8).astype(np.float32)
        S[k, :ls[k], :lt[k]] = A @ B.T
    return S, ls.astype(np.int32), lt.astype(np.int32)


def gap_counts(path, locality):
    """(runs, sum of run lengths - 1) along a reference-format path."""
    runs = ext = 0
    for (u, v), (pu, pv) in zip(path[:-1], path[1:]):
        if u - pu == 1 and v - pv == 1:
            continue
        k = (u - pu) if v == pv else (v - pv)
        runs += 1
        ext += k - 1
    if locality == O.GLOBAL:
        lu, lv = path[-1]
        k = lu + 1 if (lu >= 0 and lv < 0) else (lv + 1 if (lv >= 0 and lu < 0) else 0)
        if k:
            runs += 1
            ext += k - 1
    return runs, ext


@pytest.mark.parametrize("locality", list(LOCS))
def test_align_score_grad_matches_oracle_path(locality):
    rng = np.random.default_rng(61)
    S, ls, lt = c4_batch(rng, 24, 5, 70)
    St = torch.tensor(S, device="cuda", requires_grad=True)
    go = torch.tensor(1.0, requires_grad=True)
    ge = torch.tensor(0.2, requires_grad=True)
    scores = align_score(St, ls, lt, go, ge, locality=locality)
    scores.sum().backward()
    dS = St.grad.cpu().numpy()
    lim_s, lim_t = pa.solve.SizeRounder()(int(ls.max())), pa.solve.SizeRounder()(int(lt.max()))
    runs_tot = ext_tot = 0
    for k in range(len(ls)):
        sub = S[k, :ls[k], :lt[k]]
        ref = O.OracleResult(O.make_config(locality=LOCS[locality], gap=("affine", 1.0, 0.2),
                                           goal=O.GOAL_ONE, lim_s=lim_s, lim_t=lim_t), sub)
        assert np.float32(scores[k].item()).tobytes() == ref.score.tobytes()
        ind = np.zeros_like(S[k])
        for u, v in ref.edges():
            ind[u, v] = 1.0
        assert np.array_equal(dS[k], ind), k
        r, e = gap_counts(ref.path().tolist(), LOCS[locality])
        runs_tot += r
        ext_tot += e
    assert go.grad.item() == -runs_tot
    assert abs(ge.grad.item() + ext_tot) < 1e-3


def test_align_score_forward_equals_score_solver():
    rng = np.random.default_rng(62)
    S, ls, lt = c4_batch(rng, 40, 64, 200)
    got = align_score(torch.tensor(S, device="cuda"), ls, lt, 1.0, 0.2, locality="global").cpu().numpy()

    class Dense(pa.MatrixProblem):
        def __init__(self, m):
            super().__init__(m.shape)
            self._m = m

        def build_matrix(self, out):
            out[:, :] = self._m

    probs = [Dense(S[k, :ls[k], :lt[k]]) for k in range(len(ls))]
    ref = pa.GlobalSolver(gap_cost=pa.AffineGapCost(1.0, 0.2), codomain=pa.Score).solve(probs)
    assert np.array_equal(np.asarray(ref, np.float32).view(np.uint32), got.view(np.uint32))


@pytest.mark.parametrize("locality,gap", [("global", "affine"), ("local", "affine"), ("semiglobal", "linear")])
def test_align_score_central_differences(locality, gap):
    rng = np.random.default_rng(63)
    S, ls, lt = c4_batch(rng, 6, 20, 60)
    S = S * 8  # spread values: ties between alignments are then negligible
    St = torch.tensor(S, device="cuda", requires_grad=True)
    go = torch.tensor(1.0, requires_grad=True)
    ge = torch.tensor(0.2, requires_grad=True)
    s = align_score(St, ls, lt, go, ge if gap == "affine" else 0.0, locality=locality, gap=gap)
    s.sum().backward()
    eps = 1e-3

    def f(Sx, o, e):
        return align_score(torch.tensor(Sx, device="cuda"), ls, lt, o, e if gap == "affine" else 0.0,
                           locality=locality, gap=gap).sum().item()

    dS = St.grad.cpu().numpy()
    for _ in range(20):
        k = rng.integers(len(ls))
        i, j = rng.integers(ls[k]), rng.integers(lt[k])
        if rng.random() < 0.5:  # probe the path as often as not
            nz = np.argwhere(dS[k] != 0)
            if len(nz):
                i, j = nz[rng.integers(len(nz))]
        Sp, Sm = S.copy(), S.copy()
        Sp[k, i, j] += eps
        Sm[k, i, j] -= eps
        fd = (f(Sp, 1.0, 0.2) - f(Sm, 1.0, 0.2)) / (2 * eps)
        assert abs(fd - dS[k, i, j]) < 2e-2, (k, i, j, fd, dS[k, i, j])
    fd_o = (f(S, 1.0 + eps, 0.2) - f(S, 1.0 - eps, 0.2)) / (2 * eps)
    assert abs(fd_o - go.grad.item()) < 5e-2 * max(1.0, abs(fd_o))
    if gap == "affine":
        fd_e = (f(S, 1.0, 0.2 + eps) - f(S, 1.0, 0.2 - eps)) / (2 * eps)
        assert abs(fd_e - ge.grad.item()) < 5e-2 * max(1.0, abs(fd_e))


def test_c4_full_size_forward_and_indicator():
    """1,024 pairs, lengths U[64,512], S = dot(N(0,1)/8 embeddings of dim 64)."""
    rng = np.random.default_rng(4)
    n = 1024
    ls = rng.integers(64, 513, n)
    lt = rng.integers(64, 513, n)
    dev = torch.device("cuda")
    g = torch.Generator(device="cpu").manual_seed(4)
    A = (torch.randn(n, 512, 64, generator=g) / 8).to(dev)
    B = (torch.randn(n, 512, 64, generator=g) / 8).to(dev)
    S = torch.bmm(A, B.transpose(1, 2).contiguous()).contiguous()
    mask = (torch.arange(512, device=dev)[None, :, None] < torch.tensor(ls, device=dev)[:, None, None]) & \
           (torch.arange(512, device=dev)[None, None, :] < torch.tensor(lt, device=dev)[:, None, None])
    S = (S * mask).requires_grad_(True)
    scores = align_score(S, ls, lt, 1.0, 0.2, locality="global")
    scores.sum().backward()
    Sh = S.detach().cpu().numpy()
    dS = S.grad.cpu().numpy()
    for k in range(0, n, 97):
        ref = O.OracleResult(O.make_config(locality=O.GLOBAL, gap=("affine", 1.0, 0.2), goal=O.GOAL_ONE,
                                           lim_s=512, lim_t=512), Sh[k, :ls[k], :lt[k]])
        assert np.float32(scores[k].item()).tobytes() == ref.score.tobytes()
        ind = np.zeros((512, 512), np.float32)
        for u, v in ref.edges():
            ind[u, v] = 1.0
        assert np.array_equal(dS[k], ind)


def random_walks(rng, n, lo, hi, dim=3, step=0.1):
    lx = rng.integers(lo, hi + 1, n)
    ly = rng.integers(lo, hi + 1, n)
    X = np.zeros((n, int(lx.max()), dim), np.float32)
    Y = np.zeros((n, int(ly.max()), dim), np.float32)
    for k in range(n):
        X[k, :lx[k]] = np.cumsum(rng.normal(0, step, (lx[k], dim)), 0)
        Y[k, :ly[k]] = np.cumsum(rng.normal(0, step, (ly[k], dim)), 0)
    return X, Y, lx.astype(np.int32), ly.astype(np.int32)


def test_dtw_forward_and_path_gradient():
    rng = np.random.default_rng(65)
    X, Y, lx, ly = random_walks(rng, 16, 5, 120)
    Xt = torch.tensor(X, device="cuda", requires_grad=True)
    Yt = torch.tensor(Y, device="cuda", requires_grad=True)
    d = dtw_distance(Xt, Yt, lx, ly)
    d.sum().backward()
    gx, gy = Xt.grad.cpu().numpy(), Yt.grad.cpu().numpy()
    pf = pa.series()
    for k in range(len(lx)):
        p = pf.new_problem(X[k, :lx[k]], Y[k, :ly[k]])
        ref = O.OracleResult(O.make_config(solver_type=O.DTW, direction=O.MINIMIZE, goal=O.GOAL_ONE,
                                           lim_s=pa.solve.SizeRounder()(int(lx.max())),
                                           lim_t=pa.solve.SizeRounder()(int(ly.max()))), p.matrix)
        assert np.float32(d[k].item()).tobytes() == ref.score.tobytes()
        ex = np.zeros_like(X[k])
        ey = np.zeros_like(Y[k])
        for u, v in ref.path().tolist():
            if u >= 0 and v >= 0:
                diff = X[k, u] - Y[k, v]
                ex[u] += 2 * diff
                ey[v] -= 2 * diff
        np.testing.assert_allclose(gx[k], ex, atol=1e-4)
        np.testing.assert_allclose(gy[k], ey, atol=1e-4)


def test_dtw_central_differences():
    """Central differences on probes whose optimal warping path is unchanged
    under the perturbation (along a fixed path the distance is a quadratic in
    x, so the central quotient equals the path gradient up to rounding)."""
    rng = np.random.default_rng(66)
    X, Y, lx, ly = random_walks(rng, 4, 10, 40, step=0.3)
    Xt = torch.tensor(X, device="cuda", requires_grad=True)
    d = dtw_distance(Xt, torch.tensor(Y, device="cuda"), lx, ly)
    d.sum().backward()
    gx = Xt.grad.cpu().numpy()
    eps = 1e-3
    solver = pa.ElasticSolver(codomain=pa.Solution)
    pf = pa.series()

    def path(Xk, k):
        return np.asarray(solver.solve(pf.new_problem(Xk[:lx[k]], Y[k, :ly[k]])).path)

    stable = 0
    for _ in range(20):
        k = rng.integers(len(lx))
        i, c = rng.integers(lx[k]), rng.integers(3)
        Xp, Xm = X.copy(), X.copy()
        Xp[k, i, c] += eps
        Xm[k, i, c] -= eps
        base = path(X[k], k)
        if not (np.array_equal(path(Xp[k], k), base) and np.array_equal(path(Xm[k], k), base)):
            continue
        stable += 1
        fp = dtw_distance(torch.tensor(Xp, device="cuda"), torch.tensor(Y, device="cuda"), lx, ly)[k].item()
        fm = dtw_distance(torch.tensor(Xm, device="cuda"), torch.tensor(Y, device="cuda"), lx, ly)[k].item()
        tol = 1e-3 + 16 * float(np.spacing(np.float32(max(abs(fp), abs(fm))))) / (2 * eps)
        assert abs((fp - fm) / (2 * eps) - gx[k, i, c]) < tol, (k, i, c, (fp - fm) / (2 * eps), gx[k, i, c])
    assert stable >= 12


def test_tapes_share_the_solvers_buffers():
    """Freed tapes hand their device buffers back to the solver.  Several live
    tapes, freed in either order between forwards of different sizes, give
    the gradients of fresh solvers; a tape whose solver went first frees
    without touching it and refuses a backward."""
    import ctypes
    from pyalign_amd import algorithm, torch_ops
    from pyalign_amd._lib import check, lib
    rng = np.random.default_rng(64)
    batches = [c4_batch(rng, n, lo, hi, dim=16) for n, lo, hi in ((30, 20, 120), (50, 10, 60), (20, 100, 200))]

    def grads(S, ls, lt):
        Sg = torch.tensor(S, device="cuda", requires_grad=True)
        align_score(Sg, ls, lt, 1.0, 0.2, locality="global").sum().backward()
        return Sg.grad.cpu().numpy()

    torch_ops._SOLVERS.clear()
    fresh = []
    for b in batches:
        fresh.append(grads(*b))
        torch_ops._SOLVERS.clear()
    # one solver (limits 256 x 256) for all: live tapes overlap
    keep = []
    for order in ((0, 1, 2), (2, 0, 1)):
        live = []
        for k in order:
            S, ls, lt = batches[k]
            Sg = torch.tensor(np.pad(S, ((0, 0), (0, 256 - S.shape[1]), (0, 256 - S.shape[2]))), device="cuda",
                              requires_grad=True)
            ls2, lt2 = ls.copy(), lt.copy()
            ls2[0], lt2[0] = 256, 256  # the same rounded limits for every batch
            Sg.data[0] = torch.randn(256, 256, device="cuda") / 8
            out = align_score(Sg, ls2, lt2, 1.0, 0.2, locality="global")
            live.append((k, Sg, out))
        for q in (reversed(range(len(live))) if order[0] else range(len(live))):
            k, Sg, out = live[q]
            live[q] = None
            out.sum().backward()
            got = Sg.grad.cpu().numpy()[1:, :batches[k][0].shape[1], :batches[k][0].shape[2]]
            assert np.array_equal(got.view(np.uint32), fresh[k][1:].view(np.uint32)), (order, k)
            del out  # frees the tape: its buffers go back to the solver
        keep.append(len(torch_ops._SOLVERS))
    assert keep == [1, 1]
    # a tape that outlives its solver
    S, ls, lt = batches[1]
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Solution), locality=algorithm.Locality.GLOBAL,
        gap_cost=pa.AffineGapCost(1.0, 0.2)))
    solver = algorithm.m16.create_solver(64, 64, opts)
    St = torch.tensor(S, device="cuda")
    rg, hold = algorithm.make_ragged("dense", ls, lt, dense=St.data_ptr(),
                                     dense_off=np.arange(len(ls), dtype=np.int64) * S.shape[1] * S.shape[2],
                                     dense_ld=S.shape[2], on_device=True)
    scores = torch.empty(len(ls), device="cuda")
    h = ctypes.c_void_p()
    check(lib().pa_forward_ragged(solver._h, ctypes.byref(rg), ctypes.c_void_p(scores.data_ptr()), ctypes.byref(h)))
    check(lib().pa_solver_sync(solver._h))
    del solver
    g = torch.ones(len(ls), device="cuda")
    dS = torch.zeros_like(St)
    with pytest.raises(ValueError, match="solver was freed"):
        check(lib().pa_backward(h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(dS.data_ptr()), None, None, None))
    lib().pa_tape_free(h)


def test_learnable_gap_costs_reuse_one_solver():
    """A training loop whose gap costs change every step keeps one cached
    solver (its border vectors recomputed in place), holds device memory
    flat and gives the scores and gradients of a fresh solver bit for bit,
    including a checkpointed pair (recomputed in the backward from the
    tape's own borders) and two tapes of different gap costs alive at once."""
    from pyalign_amd import torch_ops
    rng = np.random.default_rng(65)
    S, ls, lt = c4_batch(rng, 12, 40, 300, dim=16)
    # one long global pair, recomputed strip by strip in the backward
    S = np.pad(S, ((0, 1), (0, 2100 - S.shape[1]), (0, 0)))
    S[-1, :2100, :] = (rng.standard_normal((2100, S.shape[2])) / 8).astype(np.float32)
    ls = np.append(ls, 2100).astype(np.int32)
    lt = np.append(lt, S.shape[2]).astype(np.int32)
    St = torch.tensor(S, device="cuda")

    def step(o, e, backward=True):
        Sg = St.clone().requires_grad_(True)
        go = torch.tensor(o, requires_grad=True)
        ge = torch.tensor(e, requires_grad=True)
        out = align_score(Sg, ls, lt, go, ge, locality="global")
        if not backward:
            return out, Sg, go, ge
        out.sum().backward()
        return (out.detach().cpu().numpy(), Sg.grad.cpu().numpy(), go.grad.item(), ge.grad.item())

    def same(x, y):
        return (np.array_equal(x[0].view(np.uint32), y[0].view(np.uint32))
                and np.array_equal(x[1].view(np.uint32), y[1].view(np.uint32)) and x[2] == y[2] and x[3] == y[3])

    gaps = [(1.0 + 0.013 * k, 0.2 + 0.0031 * k) for k in range(50)]
    torch_ops._SOLVERS.clear()
    kept, free = {}, []
    for k, (o, e) in enumerate(gaps):
        r = step(o, e)
        if k in (0, 17, 49):
            kept[k] = r
        assert len(torch_ops._SOLVERS) == 1
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
    assert free[10] == free[-1], (free[10], free[-1])
    for k, r in kept.items():
        torch_ops._SOLVERS.clear()
        assert same(r, step(*gaps[k])), k
    # two live tapes of different gap costs, backwards in reverse order
    torch_ops._SOLVERS.clear()
    fresh_a = step(*gaps[3])
    torch_ops._SOLVERS.clear()
    fresh_b = step(*gaps[40])
    torch_ops._SOLVERS.clear()
    a = step(*gaps[3], backward=False)
    b = step(*gaps[40], backward=False)
    for (out, Sg, go, ge), fresh in ((b, fresh_b), (a, fresh_a)):
        out.sum().backward()
        got = (out.detach().cpu().numpy(), Sg.grad.cpu().numpy(), go.grad.item(), ge.grad.item())
        assert same(got, fresh)
