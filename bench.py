#!/usr/bin/env python3
"""Batched local Gotoh alignment throughput (GCUPS) on one MI355X.

Workload: 65,536 pairs, each length floor(exp(U(ln 50, ln 2000))) drawn
independently, 20-letter alphabet, symmetric integer substitution table
(diagonal U[4,11], off-diagonal U[-4,3]) as float32, AffineGapCost(open=10,
extend=1), local alignment, score only.  Seed 2.  Sequences, table and
output scores are resident in HBM before timing starts.

A step solves the whole batch once (one persistent, autotuned launch per
strip height).  Reported beside it: the plain GPU form (the reference's own
Morton batches, one problem per lane, no work queue or autotuning) on the
same inputs with a bitwise check of the scores, and the CPU baseline (the
oracle's 16-lane restatement of the reference's batched path, timed on the
host's cores on a bounded sample of the same workload).
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS batched Gotoh local align, fwd & fwd+bwd, vs plain GPU path; CPU ref"
OPS_PER_CELL = 11  # 5 add + 4 max + local zero clamp + running best (Gotoh, local score)
# MI355X's published fp32 vector peak, 157.3 TFLOP/s; each algorithmic add
# or max counts as one operation against it
VALU_PEAK_TFLOPS = 157.3


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def c2_inputs(n=65536, seed=2):
    rng = np.random.default_rng(seed)
    K = 20
    table = rng.integers(-4, 4, size=(K, K))
    table = np.triu(table) + np.triu(table, 1).T
    np.fill_diagonal(table, rng.integers(4, 12, size=K))
    table = table.astype(np.float32)
    ls = np.floor(np.exp(rng.uniform(np.log(50), np.log(2000), n))).astype(np.int32)
    lt = np.floor(np.exp(rng.uniform(np.log(50), np.log(2000), n))).astype(np.int32)
    codes_s = rng.integers(0, K, int(ls.sum()), dtype=np.uint8)
    codes_t = rng.integers(0, K, int(lt.sum()), dtype=np.uint8)
    off_s = np.zeros(n, np.int64)
    off_t = np.zeros(n, np.int64)
    off_s[1:] = np.cumsum(ls.astype(np.int64))[:-1]
    off_t[1:] = np.cumsum(lt.astype(np.int64))[:-1]
    return table, ls, lt, codes_s, codes_t, off_s, off_t


def api_bench(table, ls, lt, codes_s, codes_t, reps=2):
    """The same bag through the public API: LocalSolver.solve on 65,536
    alphabetic problems of str sequences (problem objects made before the
    timed region; joining, device encoding, the solve and the result list
    inside it)."""
    import pyalign_amd as pa
    letters = np.array(list("ACDEFGHIKLMNPQRSTVWY"))
    byte = np.frombuffer("ACDEFGHIKLMNPQRSTVWY".encode("latin-1"), dtype=np.uint8)
    text_s = byte[codes_s].tobytes().decode("latin-1")
    text_t = byte[codes_t].tobytes().decode("latin-1")
    lookup = {(x, y): float(table[i, j]) for i, x in enumerate(letters) for j, y in enumerate(letters)}
    pf = pa.alphabetic(list(letters), lambda x, y: lookup[(x, y)])
    os_ = np.concatenate([[0], np.cumsum(ls.astype(np.int64))])
    ot = np.concatenate([[0], np.cumsum(lt.astype(np.int64))])
    t0 = time.perf_counter()
    probs = [pf.new_problem(text_s[os_[k]:os_[k + 1]], text_t[ot[k]:ot[k + 1]]) for k in range(len(ls))]
    make_s = time.perf_counter() - t0
    solver = pa.LocalSolver(gap_cost=pa.AffineGapCost(10, 1), codomain=pa.Score)
    solver.solve(probs)
    t0 = time.perf_counter()
    for _ in range(reps):
        scores = solver.solve(probs)
    secs = (time.perf_counter() - t0) / reps
    cells = float((ls.astype(np.float64) * lt).sum())
    return {"workload": "the batched local Gotoh bag as 65,536 alphabetic str problems through "
                        "LocalSolver(AffineGapCost(10, 1), codomain=Score).solve (text joined on the host, "
                        "encoded on the device)",
            "solve_s": round(secs, 4), "gcups": round(cells / secs / 1e9, 2),
            "problem_objects_s": round(make_s, 3)}, np.asarray(scores, dtype=np.float32)


def c3_inputs(seed=3):
    """One long pair: B uniform over ACGT (10,000), A a 5,000-long window of B
    with 10% substitutions and indel events at p = 0.02 per position (half
    insertions, half deletions, lengths Geometric(1/3))."""
    rng = np.random.default_rng(seed)
    alphabet = np.array(list("ACGT"))
    B = rng.integers(0, 4, 10000)
    o = int(rng.integers(0, 10000 - 5000 + 1))
    window = B[o:o + 5000]
    A = []
    i = 0
    while i < len(window):
        x = int(window[i])
        if rng.random() < 0.10:
            x = (x + int(rng.integers(1, 4))) % 4
        if rng.random() < 0.02:
            k = int(rng.geometric(1 / 3))
            if rng.random() < 0.5:  # insertion of k random bases after this one
                A.append(x)
                A.extend(rng.integers(0, 4, k).tolist())
                i += 1
            else:  # deletion of this and the next k - 1 positions
                i += k
            continue
        A.append(x)
        i += 1
    return "".join(alphabet[A]), "".join(alphabet[B])


def c1_inputs(n=4096, length=100, seed=1):
    """4,096 pairs of uniform ACGT strings, both of length 100."""
    rng = np.random.default_rng(seed)
    letters = np.array(list("ACGT"))
    return [("".join(rng.choice(letters, length)), "".join(rng.choice(letters, length))) for _ in range(n)]


def c1_bench(reps):
    """Plain global alignment through the public API: GlobalSolver(LinearGapCost(0.2),
    codomain=Solution) on 4,096 binary(eq=1, ne=-1) problems (problem objects made
    before the timed region; each solve returns 4,096 Solutions with their paths)."""
    import pyalign_amd as pa
    pairs = c1_inputs()
    pf = pa.binary(eq=1, ne=-1)
    probs = [pf.new_problem(a, b) for a, b in pairs]
    solver = pa.GlobalSolver(gap_cost=pa.LinearGapCost(0.2), codomain=pa.Solution)
    solver.solve(probs)
    t0 = time.perf_counter()
    for _ in range(reps):
        sols = solver.solve(probs)
    secs = (time.perf_counter() - t0) / reps
    cells = float(len(pairs) * 100 * 100)
    return {"workload": "plain global alignment, binary(eq=1, ne=-1), LinearGapCost(0.2): 4,096 pairs of "
                        "100 x 100 ACGT through GlobalSolver(codomain=Solution).solve",
            "solve_s": round(secs, 4), "gcups": round(cells / secs / 1e9, 3),
            "first_score": float(sols[0].score)}


def c3_bench(reps):
    """Single long semiglobal affine pair, score-only and single Solution."""
    import pyalign_amd as pa
    a, b = c3_inputs()
    pf = pa.binary(eq=1, ne=-1)
    problem = pf.new_problem(a, b)
    cells = float(len(a)) * len(b)
    out = {"workload": f"single semiglobal affine pair (open 2.0, extend 0.5, binary eq 1 ne -1): "
                       f"|A| = {len(a)}, |B| = {len(b)}", "cells": cells}
    for name, codomain in (("score", pa.Score), ("solution", pa.Solution)):
        solver = pa.SemiglobalSolver(gap_cost=pa.AffineGapCost(2.0, 0.5), codomain=codomain)
        solver.solve(problem)  # warm: builds the solver and its buffers
        t0 = time.perf_counter()
        for _ in range(reps):
            r = solver.solve(problem)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[name] = {"ms": round(ms, 3), "gcups": round(cells / (ms * 1e-3) / 1e9, 3)}
        if name == "score":
            out["optimal_score"] = float(r)
    return out


def c4_bench(steps):
    """Differentiable global affine alignment: forward score (keeping the
    traceback) against forward + backward (dScore/dS, d/d(open, extend))."""
    import torch
    from pyalign_amd.torch_ops import align_score
    n, Lmax, dim = 1024, 512, 64
    rng = np.random.default_rng(4)
    ls = rng.integers(64, Lmax + 1, n).astype(np.int32)
    lt = rng.integers(64, Lmax + 1, n).astype(np.int32)
    g = torch.Generator(device="cpu").manual_seed(4)
    A = (torch.randn(n, Lmax, dim, generator=g) / 8).cuda()
    B = (torch.randn(n, Lmax, dim, generator=g) / 8).cuda()
    S = torch.bmm(A, B.transpose(1, 2).contiguous()).contiguous()
    cells = float((ls.astype(np.float64) * lt).sum())
    go = torch.tensor(1.0, requires_grad=True)
    ge = torch.tensor(0.2, requires_grad=True)
    Sg = S.clone().requires_grad_(True)

    def fwd():
        with torch.no_grad():
            align_score(S, ls, lt, 1.0, 0.2, locality="global")

    def fwd_bwd():
        Sg.grad = None
        align_score(Sg, ls, lt, go, ge, locality="global").sum().backward()

    res = {}
    for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) * 1e3 / steps
    return {"workload": "differentiable global affine (open 1.0, extend 0.2): 1024 pairs, lengths U[64,512], "
                        "S = dot products of N(0,1)/8 embeddings of dim 64 (float32, resident in HBM)",
            "cells": cells, "fwd_ms": round(res["fwd"], 3), "fwd_bwd_ms": round(res["fwd_bwd"], 3),
            "fwd_bwd_over_fwd": round(res["fwd_bwd"] / res["fwd"], 3),
            "fwd_gcups": round(cells / (res["fwd"] * 1e-3) / 1e9, 2),
            "fwd_bwd_gcups": round(cells / (res["fwd_bwd"] * 1e-3) / 1e9, 2)}


def c5_inputs(n=8192, lo=128, hi=4096, dim=3, seed=5):
    """3-D Gaussian random walks (steps N(0, 0.1^2) per axis), lengths U[lo, hi]."""
    rng = np.random.default_rng(seed)
    lx = rng.integers(lo, hi + 1, n).astype(np.int32)
    ly = rng.integers(lo, hi + 1, n).astype(np.int32)
    X = np.zeros((n, hi, dim), np.float32)
    Y = np.zeros((n, hi, dim), np.float32)
    for k in range(n):
        X[k, :lx[k]] = np.cumsum(rng.normal(0, 0.1, (lx[k], dim)), 0, dtype=np.float32)
        Y[k, :ly[k]] = np.cumsum(rng.normal(0, 0.1, (ly[k], dim)), 0, dtype=np.float32)
    return X, Y, lx, ly


def c5_bench(steps):
    """Batched DTW with backward through the squared Euclidean cost; pairs
    longer than 2048 are checkpointed (strip recomputation)."""
    import torch
    from pyalign_amd import torch_ops
    from pyalign_amd.torch_ops import dtw_distance
    X, Y, lx, ly = c5_inputs()
    cells = float((lx.astype(np.float64) * ly).sum())
    Xd = torch.from_numpy(X).cuda()
    Yd = torch.from_numpy(Y).cuda()
    Xg = Xd.clone().requires_grad_(True)
    Yg = Yd.clone().requires_grad_(True)

    def fwd():
        with torch.no_grad():
            dtw_distance(Xd, Yd, lx, ly)

    def fwd_bwd():
        Xg.grad = None
        Yg.grad = None
        dtw_distance(Xg, Yg, lx, ly).sum().backward()

    res = {}
    for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) * 1e3 / steps
    ckpt = list(torch_ops._SOLVERS.values())[-1].last_schedule()[2]
    return {"workload": "batched DTW (squared Euclidean on 3-D Gaussian random walks, minimize): 8192 pairs, "
                        "lengths U[128,4096], score + path, backward to both series; pairs above 2048 "
                        "recomputed in checkpointed strips",
            "cells": cells, "checkpointed_pairs": ckpt, "fwd_ms": round(res["fwd"], 2),
            "fwd_bwd_ms": round(res["fwd_bwd"], 2), "fwd_bwd_over_fwd": round(res["fwd_bwd"] / res["fwd"], 3),
            "fwd_gcups": round(cells / (res["fwd"] * 1e-3) / 1e9, 2),
            "fwd_bwd_gcups": round(cells / (res["fwd_bwd"] * 1e-3) / 1e9, 2)}


def traffic():
    """HBM bytes per launch of the score kernel from the committed PMC runs
    (FETCH_SIZE / WRITE_SIZE scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "pa_fast_kernel_local_gotoh_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["traffic_bytes"])


def make_solver(lim_s, lim_t):
    import pyalign_amd as pa
    from pyalign_amd import algorithm
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL,
        gap_cost=pa.AffineGapCost(10, 1)))
    return algorithm.m16.create_solver(lim_s, lim_t, opts)


def cpu_baseline(table, ls, lt, codes_s, codes_t, off_s, off_t, sample, lim, threads=None):
    """The oracle's 16-lane batched score path (reference batching) on the host."""
    import oracle
    import pyalign_amd as pa

    class _Shape:
        form = pa.Form.INDEXED_MATRIX_FORM
        direction = "maximize"
        dtype = np.float32

        def __init__(self, a, b):
            self.shape = (a, b)

    idx = np.arange(sample)
    bag = pa.ProblemBag([_Shape(int(ls[k]), int(lt[k])) for k in idx])
    members, boff = [], [0]
    for b in bag.batches(16):  # instance.py:230-275 with 16 AVX-512 lanes
        members += list(b.indices)
        boff.append(len(members))
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    cfg = oracle.make_config(locality=oracle.LOCAL, gap=("affine", 10, 1), goal=oracle.GOAL_SCORE,
                             lim_s=lim, lim_t=lim)
    secs, scores = oracle.baseline_indexed_score(
        cfg, codes_s, off_s[:sample], ls[:sample].astype(np.int64), codes_t, off_t[:sample],
        lt[:sample].astype(np.int64), table, np.array(members), np.array(boff), threads)
    cells = float((ls[:sample].astype(np.float64) * lt[:sample]).sum())
    return cells / secs / 1e9, threads, secs, scores


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a float32 elementwise copy of this many bytes (PMC byte-counter scale)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the long-pair and fwd+bwd measurements")
    ap.add_argument("--only", choices=["fwd_bwd", "dtw_fwd_bwd", "long_pair", "global_solution"], default=None,
                    help="only time one side workload (differentiable: 5 steps / 1 step; long pair: 5 solves) "
                         "and print its JSON (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(independent problems would shard as contiguous cost-balanced ranges, one per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from pyalign_amd import _lib, algorithm
    from pyalign_amd._lib import check, lib

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.add(src, 0.0, out=dst)  # reads and writes calibrate_copy bytes once
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy}))
        return
    if args.only:
        only = {"fwd_bwd": lambda: c4_bench(5), "dtw_fwd_bwd": lambda: c5_bench(1), "long_pair": lambda: c3_bench(5),
                "global_solution": lambda: c1_bench(10)}
        print(json.dumps(only[args.only]()))
        return
    table, ls, lt, codes_s, codes_t, off_s, off_t = c2_inputs()
    n = len(ls)
    cells = float((ls.astype(np.float64) * lt).sum())
    lim = int(max(ls.max(), lt.max()))  # SizeRounder keeps limits >= 1024 exact
    solver = make_solver(lim, lim)

    # inputs resident in HBM
    d_cs = torch.from_numpy(codes_s).cuda()
    d_ct = torch.from_numpy(codes_t).cuda()
    d_scores = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rg, keep = algorithm.make_ragged("indexed", ls, lt, codes_s=d_cs.data_ptr(), codes_t=d_ct.data_ptr(),
                                     off_s=off_s, off_t=off_t, table=table, on_device=True)

    def enqueue(schedule):
        check(lib().pa_solve_ragged_scores_device(solver._h, ctypes.byref(rg), schedule,
                                                  ctypes.c_void_p(d_scores.data_ptr())))

    def step(schedule):
        enqueue(schedule)
        check(lib().pa_solver_sync(solver._h))
        ms, launches, _ = solver.last_kernel_ms()
        return ms, launches

    # HIP events on the solver's own stream (its kernels run there)
    sh = ctypes.c_void_p()
    check(lib().pa_solver_stream(solver._h, ctypes.byref(sh)))
    sstream = torch.cuda.ExternalStream(sh.value) if sh.value else torch.cuda.current_stream()

    log(f"inputs ready: {n} pairs, {cells:.3e} cells")
    for _ in range(args.warmup):
        step(_lib.PA_SCHED_AUTO)
    log("warmup done")
    # the timed steps go back to back (no host wait between them: the host
    # enqueues the next solve while the GPU runs the last); events on the
    # solver's stream bracket them, so their mean is the launch sequence's
    # (queue-clearing fill and the DP kernel) average duration
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    check(lib().pa_solver_sync(solver._h))
    t0 = time.perf_counter()
    ev0.record(sstream)
    for _ in range(args.steps):
        enqueue(_lib.PA_SCHED_AUTO)
    ev1.record(sstream)
    check(lib().pa_solver_sync(solver._h))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kms = ev0.elapsed_time(ev1) / args.steps
    _, launches, _ = solver.last_kernel_ms()
    auto_scores = d_scores.cpu().numpy().copy()
    value = cells / (ms_per_step * 1e-3) / 1e9
    achieved = cells * OPS_PER_CELL / (kms * 1e-3) / 1e12

    log(f"batched: {ms_per_step:.2f} ms/step, {value:.1f} GCUPS")
    # the same solve from host buffers: codes uploaded and scores read back per call
    host = None
    if not args.no_extra:
        solver.solve_ragged("score", "indexed", ls, lt, codes_s=codes_s, codes_t=codes_t, off_s=off_s, off_t=off_t,
                            table=table)
        th0 = time.perf_counter()
        for _ in range(3):
            hs = solver.solve_ragged("score", "indexed", ls, lt, codes_s=codes_s, codes_t=codes_t, off_s=off_s,
                                     off_t=off_t, table=table)
        hms = (time.perf_counter() - th0) * 1e3 / 3
        hs = np.asarray(hs, dtype=np.float32)
        host = {"workload": "the same bag from host buffers (69 MB of codes uploaded, scores read back, per call)",
                "ms_per_solve": round(hms, 3), "gcups": round(cells / (hms * 1e-3) / 1e9, 2),
                "scores_bitwise_equal": bool(np.array_equal(hs.view(np.uint32), auto_scores.view(np.uint32)))}
        log(f"host buffers: {hms:.2f} ms per solve")
    plain = None
    if not args.no_plain:
        step(_lib.PA_SCHED_PLAIN)
        pms = []
        tp0 = time.perf_counter()
        for _ in range(args.plain_steps):
            pms.append(step(_lib.PA_SCHED_PLAIN)[0])
        torch.cuda.synchronize()
        plain_step = (time.perf_counter() - tp0) * 1e3 / args.plain_steps
        plain_scores = d_scores.cpu().numpy()
        plain = {"value": round(cells / (plain_step * 1e-3) / 1e9, 3), "unit": "GCUPS",
                 "ms_per_step": round(plain_step, 3), "kernel_ms": round(float(np.mean(pms)), 3),
                 "scores_bitwise_equal": bool(np.array_equal(plain_scores.view(np.uint32),
                                                             auto_scores.view(np.uint32))),
                 "speedup_of_batched": round(value / (cells / (plain_step * 1e-3) / 1e9), 2)}

    if plain:
        log(f"plain: {plain['ms_per_step']:.2f} ms/step")
    cpu = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        gcups, threads, secs, cpu_scores = cpu_baseline(table, ls, lt, codes_s, codes_t, off_s, off_t, sample, lim)
        g1, _, secs1, _ = cpu_baseline(table, ls, lt, codes_s, codes_t, off_s, off_t, min(2048, n), lim, threads=1)
        cpu = {"value": round(gcups, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "single_core": {"value": round(g1, 4), "sample": f"first {min(2048, n)} pairs ({secs1:.1f} s)"},
               "sample": (f"all {n} pairs" if sample == n else f"first {sample} of the {n} pairs")
                         + f" ({secs:.1f} s; oracle's 16-lane batched restatement, Morton batches of 16 as the "
                           "reference's ProblemBag.batches)",
               "scores_bitwise_equal": bool(np.array_equal(cpu_scores.view(np.uint32),
                                                           auto_scores[:sample].view(np.uint32)))}

    long_pair = fwd_bwd = dtw = api = c1 = None
    if not args.no_extra:
        c1 = c1_bench(10)
        log(f"global Solution batch: {c1['solve_s']:.3f} s per solve")
        api, api_scores = api_bench(table, ls, lt, codes_s, codes_t)
        api["scores_bitwise_equal"] = bool(np.array_equal(api_scores.view(np.uint32), auto_scores.view(np.uint32)))
        log(f"public API: {api['solve_s']:.3f} s per solve, {api['gcups']:.1f} GCUPS")
        fwd_bwd = c4_bench(5)
        log(f"fwd+bwd: {fwd_bwd['fwd_ms']:.2f} / {fwd_bwd['fwd_bwd_ms']:.2f} ms")
        long_pair = c3_bench(2)
        log(f"long pair: score {long_pair['score']['ms']:.1f} ms, solution {long_pair['solution']['ms']:.1f} ms")
        dtw = c5_bench(1)
        log(f"dtw: {dtw['fwd_ms']:.1f} / {dtw['fwd_bwd_ms']:.1f} ms")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched local Gotoh score-only: 65536 pairs, lengths log-uniform [50,2000], "
                               "20-letter integer table, affine open 10 extend 1",
                   "pairs": n, "cells": cells, "forward": "score", "launches_per_step": launches},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK_TFLOPS, 4), "traffic": traffic(),
                     "ops_per_cell": OPS_PER_CELL, "kernel_ms_per_step": round(kms, 3),
                     "peak_basis": "MI355X published fp32 vector peak; one add or max = one operation"},
        "plain_gpu": plain,
        "cpu_baseline": cpu,
        "fwd_bwd": fwd_bwd,
        "long_pair": long_pair,
        "dtw_fwd_bwd": dtw,
        "host_buffers": host,
        "public_api": api,
        "global_solution_batch": c1,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
