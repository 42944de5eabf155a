"""The native solver module, backed by libpyalign_hip.so through ctypes.

This module stands in for the reference's pybind11 extension
``pyalign.algorithm.<variant>.algorithm`` (pyalign/algorithm/module.cpp:1-18,
pyalign/algorithm/pyalign.h:12-102) and exposes the same names: the enums
``Type``, ``Direction``, ``Detail``, ``Count``, ``Locality``; the
submodules ``m16`` / ``m32`` with ``create_options`` and ``create_solver``;
the ``Solver`` methods ``solve_{,indexed_,binary_}for_{score,alignment,
alignment_iterator,solution,solution_iterator}``; and the result classes
``Alignment``, ``AlignmentIterator``, ``Solution``, ``SolutionIterator``,
``Algorithm``.  All DP work runs in the HIP kernels.
"""

import contextlib
import ctypes
import enum
import gc
import threading
import types

import numpy as np

from . import _lib
from ._lib import check, lib



@contextlib.contextmanager
def collector_paused():
    """Python's cyclic garbage collector paused while a bag's result objects
    are built (the core results here, the Solutions of solve.py around
    them): thousands of new objects otherwise trigger collections that
    traverse every live object of the process (4,096 Solutions of 100 x 100
    global alignments: 13 ms per solve paused, 34 ms beside the 170 K live
    objects that importing torch leaves, up to 110 ms beside 2 M).
    Reference counting still frees acyclic garbage at once; cycles wait for
    the next collection.

    The collector's switch is process-wide: builds in several threads share
    one pause (a count under a lock), which ends with the last of them, and
    a collector that was off at the first of them stays off."""
    global _pause_count, _pause_restore
    with _pause_lock:
        if _pause_count == 0:
            _pause_restore = gc.isenabled()
            gc.disable()
        _pause_count += 1
    try:
        yield
    finally:
        with _pause_lock:
            _pause_count -= 1
            if _pause_count == 0 and _pause_restore:
                gc.enable()


_pause_lock = threading.Lock()
_pause_count = 0
_pause_restore = False

class Type(enum.IntEnum):  # enum.h:6-9
    ALIGNMENT = 0
    DTW = 1


class Direction(enum.IntEnum):  # enum.h:11-14
    MINIMIZE = 0
    MAXIMIZE = 1


class Detail(enum.IntEnum):  # enum.h:16-20
    SCORE = 0
    ALIGNMENT = 1
    SOLUTION = 2


class Count(enum.IntEnum):  # enum.h:22-25
    ONE = 0
    ALL = 1


class Locality(enum.IntEnum):  # enum.h:27-31
    LOCAL = 0
    GLOBAL = 1
    SEMIGLOBAL = 2


class Algorithm:
    """AlgorithmMetaData (solver.h:154-177) as exposed by pyalign.h:235-241."""

    def __init__(self, name, runtime, memory):
        self._name, self._runtime, self._memory = name, runtime, memory

    @property
    def name(self):
        return self._name

    @property
    def runtime(self):
        return self._runtime

    @property
    def memory(self):
        return self._memory


def _algorithm_for(opts):
    # solver.h:2799-2802, 2914, 3107, 3203
    if opts.solver == Type.DTW:
        return Algorithm("DTW", "n^2", "n^2")
    if opts.gap_kind == "affine":
        return Algorithm("Gotoh", "n^2", "n^2")
    if opts.gap_kind == "general":
        return Algorithm("Waterman-Smith-Beyer", "n^3", "n^2")
    return Algorithm("Needleman-Wunsch" if opts.locality == Locality.GLOBAL else "Smith-Waterman", "n^2", "n^2")


def _gap_spec(gap):
    """GapCostOptions (factory.h:33-64): (kind, [u, v]) from to_special_case()."""
    if gap is None:
        return _lib.PA_GAP_NONE, (0.0, 0.0)
    special = gap.to_special_case()
    if "affine" in special:
        u, v = special["affine"]
        return _lib.PA_GAP_AFFINE, (float(u), float(v))
    if "linear" in special:
        return _lib.PA_GAP_LINEAR, (float(special["linear"]), 0.0)
    return _lib.PA_GAP_GENERAL, (0.0, 0.0)


class Options:
    """Parsed options dict (options.h:9-117)."""

    def __init__(self, options, index_bits):
        self._dict = options
        self.index_bits = index_bits
        self.solver = Type(int(options["solver"]))
        self.batch = bool(options["batch"])
        self.direction = Direction(int(options["direction"]))
        self.return_dup = bool(options["return_dup"])
        desc = _lib.OptionsDesc()
        desc.solver_type = int(self.solver)
        desc.batch = int(self.batch)
        desc.direction = int(self.direction)
        desc.return_dup = int(self.return_dup)
        desc.index_bits = index_bits
        self.locality = Locality.LOCAL
        self.gap_kind = None
        self.gap_objects = (None, None)
        if self.solver == Type.ALIGNMENT:
            codomain = options["codomain"]
            self.detail = Detail(int(codomain.detail))
            self.count = Count(int(codomain.count))
            self.locality = Locality(int(options["locality"])) if "locality" in options else Locality.LOCAL
            gap = options.get("gap_cost")
            if isinstance(gap, dict):
                gs, gt = gap.get("s"), gap.get("t")
            else:
                gs = gt = gap
            ks, vs = _gap_spec(gs)
            kt, vt = _gap_spec(gt)
            desc.detail, desc.count, desc.locality = int(self.detail), int(self.count), int(self.locality)
            desc.gap_kind_s, desc.gap_kind_t = ks, kt
            desc.gap_s[0], desc.gap_s[1] = vs
            desc.gap_t[0], desc.gap_t[1] = vt
            if ks == _lib.PA_GAP_AFFINE and kt == _lib.PA_GAP_AFFINE:
                self.gap_kind = "affine"
            elif ks == _lib.PA_GAP_GENERAL and kt == _lib.PA_GAP_GENERAL:
                self.gap_kind = "general"
            else:
                self.gap_kind = "linear"
            self.gap_objects = (gs, gt)
        else:
            self.detail, self.count = Detail.SOLUTION, Count.ONE
        h = ctypes.c_void_p()
        self._free = lib().pa_options_free  # held so that teardown at exit needs no module global
        check(lib().pa_create_options(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    def to_dict(self):
        return self._dict


class Alignment:
    """Alignment (factory.h:130-193): score and s_to_t / t_to_s maps."""

    def __init__(self, score, s_to_t, t_to_s):
        self._score = float(score)
        self._s_to_t = s_to_t
        self._t_to_s = t_to_s

    @property
    def score(self):
        return self._score

    @property
    def s_to_t(self):
        return self._s_to_t

    @property
    def t_to_s(self):
        return self._t_to_s


class Solution:
    """Solution (factory.h:230-352) of one problem."""

    def __init__(self, score, path, alignment, algorithm, max_degree_1=True, lazy=None):
        self._score = score
        self._path = path
        self._alignment = alignment
        self._algorithm = algorithm
        self._max_degree_1 = max_degree_1
        self._lazy = lazy

    @property
    def score(self):
        return self._score

    @property
    def path(self):
        return self._path

    @property
    def alignment(self):
        return self._alignment

    @property
    def algorithm(self):
        return self._algorithm

    @property
    def traceback_has_max_degree_1(self):
        return self._max_degree_1

    def _materialised(self):
        if self._lazy is None:
            raise NotImplementedError(
                "Solution matrices need the problem's data on the host (device-resident inputs keep none)")
        return self._lazy()

    @property
    def values(self):
        return self._materialised()["values"]

    @property
    def traceback_as_matrix(self):
        return self._materialised()["traceback_as_matrix"]

    @property
    def traceback_as_edges(self):
        return self._materialised()["traceback_as_edges"]


class _Matrices:
    """Solution.values / traceback_as_matrix / traceback_as_edges of one
    problem, computed on first use by solving it again with every layer kept
    (solver.h:2355-2471) and then cached.  Holds a private copy of the
    problem's data: the caller's batch buffers are reused between batches."""

    def __init__(self, solver, form, len_s, len_t, payload):
        self._solver = solver
        self._form = form
        self._ls, self._lt = int(len_s), int(len_t)
        self._payload = payload
        self._data = None

    def __call__(self):
        if self._data is None:
            self._data = self._solver._matrices(self._form, self._ls, self._lt, self._payload)
            self._payload = None
        return self._data


class _Result:
    """Owns one pa_result handle."""

    def __init__(self, h, n, index_dtype, len_s, len_t):
        self._free = lib().pa_result_free
        self.h = h
        self.n = n
        self.index_dtype = index_dtype
        self.len_s = np.asarray(len_s, dtype=np.int64)
        self.len_t = np.asarray(len_t, dtype=np.int64)

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def scores(self):
        out = np.empty(self.n, dtype=np.float32)
        if self.n:
            check(lib().pa_result_scores(self.h, out.ctypes.data))
        return out

    def path(self, k):
        m = lib().pa_result_path_len(self.h, k)
        if m < 0:
            return None
        uv = np.empty((m, 2), dtype=np.int32)
        check(lib().pa_result_path(self.h, k, uv.ctypes.data))
        return uv.astype(self.index_dtype)

    def all_paths(self):
        """Every problem's path length (-1: none) and the paths concatenated."""
        lens = np.empty(max(self.n, 1), dtype=np.int64)
        check(lib().pa_result_paths(self.h, lens.ctypes.data, None))
        total = int(np.maximum(lens[:self.n], 0).sum())
        uv = np.empty((max(total, 1), 2), dtype=np.int32)
        check(lib().pa_result_paths(self.h, lens.ctypes.data, uv.ctypes.data))
        return lens[:self.n], uv[:total].astype(self.index_dtype)

    def all_alignments(self):
        """Every problem's s_to_t and t_to_s, concatenated in problem order."""
        s2t = np.empty(max(int(self.len_s.sum()), 1), dtype=np.int32)
        t2s = np.empty(max(int(self.len_t.sum()), 1), dtype=np.int32)
        check(lib().pa_result_alignments(self.h, s2t.ctypes.data, t2s.ctypes.data))
        return s2t.astype(self.index_dtype), t2s.astype(self.index_dtype)

    def alignment(self, k, score):
        ls, lt = int(self.len_s[k]), int(self.len_t[k])
        s_to_t = np.empty(ls, dtype=np.int32)
        t_to_s = np.empty(lt, dtype=np.int32)
        check(lib().pa_result_alignment(self.h, k, s_to_t.ctypes.data, t_to_s.ctypes.data))
        return Alignment(score, s_to_t.astype(self.index_dtype), t_to_s.astype(self.index_dtype))

    def iter_next(self, k):
        score = ctypes.c_float()
        r = lib().pa_result_iter_next(self.h, k, ctypes.byref(score))
        if r < 0:
            check(-r)
        if r == 0:
            return None
        m = lib().pa_result_iter_path_len(self.h, k)
        uv = np.empty((m, 2), dtype=np.int32)
        check(lib().pa_result_iter_path(self.h, k, uv.ctypes.data))
        return float(np.float32(score.value)), uv.astype(self.index_dtype)


class AlignmentIterator:
    """AlignmentIterator (factory.h:195-216): next() returns None at the end."""

    def __init__(self, result, k):
        self._r, self._k = result, k

    def next(self):
        x = self._r.iter_next(self._k)
        if x is None:
            return None
        return self._r.alignment(self._k, x[0])


class SolutionIterator:
    """SolutionIterator (factory.h:354-385); solutions carry no algorithm
    record, as in the reference's iterator path (solver.h:1881)."""

    def __init__(self, result, k, lazy=None, max_degree_1=False):
        self._r, self._k, self._lazy = result, k, lazy
        self._max_degree_1 = max_degree_1

    def next(self):
        x = self._r.iter_next(self._k)
        if x is None:
            return None
        score, path = x
        return Solution(score, path, self._r.alignment(self._k, score), None,
                        max_degree_1=self._max_degree_1, lazy=self._lazy)


def device_count():
    """Number of HIP devices the library sees (0 without a GPU)."""
    return int(lib().pa_device_count())


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


class Solver:
    """Solver<Value, Index> (factory.h:414-508) on the MI355X."""

    def __init__(self, h, opts, max_len_s, max_len_t):
        self._free = lib().pa_solver_free
        self._h = h
        self._opts = opts
        self._index_dtype = np.int16 if opts.index_bits == 16 else np.int32
        self.max_len_s, self.max_len_t = max_len_s, max_len_t

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    @property
    def options(self):
        return self._opts.to_dict()

    @property
    def batch_size(self):
        return lib().pa_solver_batch_size(self._h)

    # ---- shared result conversion ------------------------------------
    def _wrap(self, h, len_s, len_t, kind, one=None):
        """one(k) -> (form, payload) copies problem k's data for its
        Solution matrices (None: not available)."""
        n = lib().pa_result_count(h)
        r = _Result(h, n, self._index_dtype, len_s, len_t)
        scores = r.scores()
        if kind == "score":
            return scores
        if kind == "alignment":
            return tuple(r.alignment(k, scores[k]) for k in range(n))

        def lazy(k):
            if one is None:
                return None
            a, b = int(r.len_s[k]), int(r.len_t[k])  # the closure keeps no result handle alive

            def make():
                form, payload = one(k)
                return _Matrices(self, form, a, b, payload)

            return _Deferred(make)

        if kind == "solution":
            with collector_paused():
                return self._solutions(r, n, scores, lazy)
        if kind == "alignment_iterator":
            return tuple(AlignmentIterator(r, k) for k in range(n))
        if kind == "solution_iterator":
            # traceback_n keeps every predecessor; DTW stays traceback_1
            deg1 = self._opts.solver == Type.DTW or self._opts.count == Count.ONE
            return tuple(SolutionIterator(r, k, lazy(k), deg1) for k in range(n))
        raise ValueError(kind)

    def _solutions(self, r, n, scores, lazy):
        """The Solutions of a solve (with their paths and alignments)."""
        alg = _algorithm_for(self._opts)
        lens, uv = r.all_paths()
        s2t, t2s = r.all_alignments()
        po = np.concatenate([[0], np.cumsum(np.maximum(lens, 0))])
        so = np.concatenate([[0], np.cumsum(r.len_s)])
        to = np.concatenate([[0], np.cumsum(r.len_t)])
        out = []
        for k in range(n):
            if lens[k] < 0:
                out.append(Solution(None, None, None, alg, lazy=lazy(k)))
            else:
                al = Alignment(scores[k], s2t[so[k]:so[k + 1]], t2s[to[k]:to[k + 1]])
                out.append(Solution(float(scores[k]), uv[po[k]:po[k + 1]], al, alg, lazy=lazy(k)))
        return tuple(out)

    def _matrices(self, form, ls, lt, payload):
        """Solve one problem again with its DP matrices kept."""
        rg, keep = make_ragged(form, [ls], [lt], **payload)
        h = ctypes.c_void_p()
        check(lib().pa_solve_ragged_matrices(self._h, ctypes.byref(rg), ctypes.byref(h)))
        del keep
        r = _Result(h, 1, self._index_dtype, [ls], [lt])
        layers = lib().pa_result_layers(h)
        values = np.empty((layers, ls + 1, lt + 1), dtype=np.float32)
        check(lib().pa_result_values(h, 0, values.ctypes.data))
        tb = np.empty((layers, ls + 1, lt + 1, 2), dtype=np.int32)
        check(lib().pa_result_traceback(h, 0, tb.ctypes.data))
        edges = []
        for layer in range(layers):
            m = lib().pa_result_edges_len(h, 0, layer)
            if m < 0:
                check(_lib.PA_EINVAL)
            e = np.empty((m, 2, 2), dtype=np.int32)
            if m:
                check(lib().pa_result_edges(h, 0, layer, e.ctypes.data))
            edges.append(e.astype(self._index_dtype))
        del r
        return {"values": values, "traceback_as_matrix": tb.astype(self._index_dtype),
                "traceback_as_edges": edges}

    def _dense(self, sim, length, kind):
        sim = np.ascontiguousarray(sim, dtype=np.float32)
        length = _i32(length)
        if sim.ndim != 3 or length.shape != (2, sim.shape[2]):
            raise ValueError(f"m_length has shape {length.shape}, expected (2, {sim.shape[2] if sim.ndim == 3 else '?'})")
        h = ctypes.c_void_p()
        check(lib().pa_solve_dense(self._h, sim.ctypes.data, sim.shape[0], sim.shape[1], sim.shape[2],
                                   length.ctypes.data, ctypes.byref(h)))

        def one(k):
            m = sim[:length[0, k], :length[1, k], k]
            return "dense", dict(dense=m.ravel().copy(), dense_off=[0])

        return self._wrap(h, length[0], length[1], kind, one)

    def _indexed(self, a, b, sim, length, kind):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        sim = np.ascontiguousarray(sim, dtype=np.float32)
        length = _i32(length)
        if length.shape != (2, a.shape[0]):
            raise ValueError(f"m_length has shape {length.shape}, expected (2, {a.shape[0]})")
        if b.shape[0] != a.shape[0]:
            raise ValueError(f"problem of batch size {b.shape[0]} given to solver that was configured to "
                             f"batch size {self.batch_size}")
        h = ctypes.c_void_p()
        check(lib().pa_solve_indexed(self._h, a.ctypes.data, a.shape[1], b.ctypes.data, b.shape[1], a.shape[0],
                                     sim.ctypes.data, sim.shape[0], sim.shape[1], length.ctypes.data,
                                     ctypes.byref(h)))
        table = sim.copy() if kind.startswith("solution") else None
        code_t = np.uint16 if max(sim.shape) > 128 else np.uint8

        def one(k):
            return "indexed", dict(codes_s=a[k, :length[0, k]].astype(code_t),
                                   codes_t=b[k, :length[1, k]].astype(code_t), off_s=[0], off_t=[0], table=table)

        return self._wrap(h, length[0], length[1], kind, one)

    def _binary(self, a, b, eq, ne, length, kind):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        length = _i32(length)
        if length.shape != (2, a.shape[0]):
            raise ValueError(f"m_length has shape {length.shape}, expected (2, {a.shape[0]})")
        h = ctypes.c_void_p()
        check(lib().pa_solve_binary(self._h, a.ctypes.data, a.shape[1], b.ctypes.data, b.shape[1], a.shape[0],
                                    float(eq), float(ne), length.ctypes.data, ctypes.byref(h)))

        def one(k):
            return "binary", dict(codes_s=a[k, :length[0, k]].copy(), codes_t=b[k, :length[1, k]].copy(),
                                  off_s=[0], off_t=[0], eq=eq, ne=ne)

        return self._wrap(h, length[0], length[1], kind, one)

    # ---- the reference's 15 entry points (pyalign.h:24-41) -----------
    def solve_for_score(self, sim, length):
        return self._dense(sim, length, "score")

    def solve_indexed_for_score(self, a, b, sim, length):
        return self._indexed(a, b, sim, length, "score")

    def solve_binary_for_score(self, a, b, eq, ne, length):
        return self._binary(a, b, eq, ne, length, "score")

    def solve_for_alignment(self, sim, length):
        return self._dense(sim, length, "alignment")

    def solve_for_alignment_iterator(self, sim, length):
        return self._dense(sim, length, "alignment_iterator")

    def solve_indexed_for_alignment(self, a, b, sim, length):
        return self._indexed(a, b, sim, length, "alignment")

    def solve_indexed_for_alignment_iterator(self, a, b, sim, length):
        return self._indexed(a, b, sim, length, "alignment_iterator")

    def solve_binary_for_alignment(self, a, b, eq, ne, length):
        return self._binary(a, b, eq, ne, length, "alignment")

    def solve_binary_for_alignment_iterator(self, a, b, eq, ne, length):
        return self._binary(a, b, eq, ne, length, "alignment_iterator")

    def solve_for_solution(self, sim, length):
        return self._dense(sim, length, "solution")

    def solve_for_solution_iterator(self, sim, length):
        return self._dense(sim, length, "solution_iterator")

    def solve_indexed_for_solution(self, a, b, sim, length):
        return self._indexed(a, b, sim, length, "solution")

    def solve_indexed_for_solution_iterator(self, a, b, sim, length):
        return self._indexed(a, b, sim, length, "solution_iterator")

    def solve_binary_for_solution(self, a, b, eq, ne, length):
        return self._binary(a, b, eq, ne, length, "solution")

    def solve_binary_for_solution_iterator(self, a, b, eq, ne, length):
        return self._binary(a, b, eq, ne, length, "solution_iterator")

    # ---- whole-bag entry (no reference counterpart: one call per bag) --
    def solve_ragged(self, kind, form, len_s, len_t, schedule=_lib.PA_SCHED_AUTO, *, payload_owned=False, **payload):
        """Solve n problems of individual lengths in one call.

        form: "indexed" (codes_s/codes_t uint8, off_s/off_t, table KxK),
        "binary" (uint32 codes, eq, ne), "dense" (dense float32 flat,
        dense_off) or "series" (float32 rows x dim).  kind as in _wrap.
        payload_owned: the payload's arrays belong to this call (nobody
        changes them later), so Solutions keep them without a copy.
        """
        rg, keep = make_ragged(form, len_s, len_t, **payload)
        h = ctypes.c_void_p()
        check(lib().pa_solve_ragged(self._h, ctypes.byref(rg), schedule, ctypes.byref(h)))
        del keep
        one = None
        if kind.startswith("solution") and not payload.get("on_device", False):
            one = _ragged_one(form, len_s, len_t, payload, copy=not payload_owned)
        return self._wrap(h, len_s, len_t, kind, one)

    def set_checkpoint(self, min_len):
        """Differentiation: recompute global pairs longer than min_len strip by
        strip in the backward instead of keeping their traceback (-1: never)."""
        check(lib().pa_solver_set_checkpoint(self._h, int(min_len)))

    def set_stream(self, stream):
        """Run on the caller's HIP stream (an int handle, 0 = null stream), or
        on the solver's own stream when stream is None."""
        ext = stream is not None
        check(lib().pa_solver_set_stream(self._h, ctypes.c_void_p(int(stream) if ext else 0), int(ext)))

    def set_gap_cost(self, gap_cost):
        """Replace the solver's linear or affine gap cost (a GapCost, or a dict
        {'s': ..., 't': ...}) in place: same limits, buffers and plan cache,
        new border vectors.  The kind must stay the same."""
        if isinstance(gap_cost, dict):
            gs, gt = gap_cost.get("s"), gap_cost.get("t")
        else:
            gs = gt = gap_cost
        (ks, vs), (kt, vt) = _gap_spec(gs), _gap_spec(gt)
        kind = {_lib.PA_GAP_NONE: "linear", _lib.PA_GAP_LINEAR: "linear", _lib.PA_GAP_AFFINE: "affine"}
        if kind.get(ks) != self._opts.gap_kind or kind.get(kt) != self._opts.gap_kind:
            raise ValueError("the gap cost kind of a solver cannot change")
        ds = (ctypes.c_double * 2)(*vs)
        dt = (ctypes.c_double * 2)(*vt)
        check(lib().pa_solver_set_gap_params(self._h, ds, dt))

    def set_graphs(self, enable):
        """Replay repeated launch sequences from a captured hipGraph (default on)."""
        check(lib().pa_solver_set_graphs(self._h, int(bool(enable))))

    def last_schedule(self):
        """(work groups, problems in band mode, checkpointed problems) of the last solve."""
        groups = ctypes.c_int32()
        band = ctypes.c_int32()
        ckpt = ctypes.c_int32()
        check(lib().pa_solver_last_schedule(self._h, ctypes.byref(groups), ctypes.byref(band), ctypes.byref(ckpt)))
        return groups.value, band.value, ckpt.value

    def last_bad_symbol(self):
        """First positions of symbols outside the alphabet in the last raw
        solve, (s, t), -1 for none."""
        ps, pt = ctypes.c_int64(), ctypes.c_int64()
        check(lib().pa_solver_last_bad_symbol(self._h, ctypes.byref(ps), ctypes.byref(pt)))
        return ps.value, pt.value

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        launches = ctypes.c_int32()
        cells = ctypes.c_double()
        check(lib().pa_solver_last_kernel_ms(self._h, ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(cells)))
        return ms.value, launches.value, cells.value


class _Deferred:
    """A problem's Solution matrices, built on first use."""

    __slots__ = ("_make", "_m")

    def __init__(self, make):
        self._make, self._m = make, None

    def __call__(self):
        if self._m is None:
            self._m = self._make()
        return self._m()


def _ragged_one(form, len_s, len_t, p, copy=True):
    """Problem k's data out of a ragged payload, as a one-problem payload
    (a caller's arrays are copied once here, so later changes by the caller
    do not reach solutions built from them; arrays the solve built itself
    are kept as they are)."""
    ls = np.array(len_s, dtype=np.int64)
    lt = np.array(len_t, dtype=np.int64)
    if copy:
        p = {key: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for key, v in p.items()}

    def one(k):
        a, b = int(ls[k]), int(lt[k])
        if form == "dense":
            o = int(np.asarray(p["dense_off"], dtype=np.int64)[k])
            ld = int(p.get("dense_ld", 0)) or b
            flat = np.asarray(p["dense"], dtype=np.float32).ravel()
            m = np.zeros((a, b), dtype=np.float32)
            for i in range(a):
                m[i] = flat[o + i * ld:o + i * ld + b]
            return "dense", dict(dense=m.ravel(), dense_off=[0])
        os_ = int(np.asarray(p["off_s"], dtype=np.int64)[k])
        ot = int(np.asarray(p["off_t"], dtype=np.int64)[k])
        if form == "series":
            d = int(p["dim"])
            xs = np.asarray(p["codes_s"], dtype=np.float32).ravel()
            xt = np.asarray(p["codes_t"], dtype=np.float32).ravel()
            return "series", dict(codes_s=xs[os_ * d:(os_ + a) * d].copy(), codes_t=xt[ot * d:(ot + b) * d].copy(),
                                  off_s=[0], off_t=[0], dim=d)
        cs, ct = np.asarray(p["codes_s"]).ravel(), np.asarray(p["codes_t"]).ravel()
        w = int(p.get("raw_width", 0)) or 1  # raw symbols: offsets count symbols of w bytes
        if p.get("raw_width", 0):
            q = dict(codes_s=cs[os_ * w:(os_ + a) * w].copy(), codes_t=ct[ot * w:(ot + b) * w].copy(),
                     off_s=[0], off_t=[0], raw_width=w)
            if p.get("lut") is not None:
                q["lut"] = p["lut"]
        else:
            q = dict(codes_s=cs[os_:os_ + a].copy(), codes_t=ct[ot:ot + b].copy(), off_s=[0], off_t=[0])
        if form == "indexed":
            q["table"] = p["table"]
        else:
            q["eq"], q["ne"] = p.get("eq", 0.0), p.get("ne", 0.0)
        return form, q

    return one


_FORMS = {"dense": _lib.PA_FORM_DENSE, "indexed": _lib.PA_FORM_INDEXED,
          "binary": _lib.PA_FORM_BINARY, "series": _lib.PA_FORM_SERIES}


def make_ragged(form, len_s, len_t, codes_s=None, codes_t=None, off_s=None, off_t=None, table=None,
                eq=0.0, ne=0.0, dense=None, dense_off=None, dense_ld=0, dim=0, on_device=False, raw_width=0,
                lut=None, code_bytes=0):
    """Build a pa_ragged descriptor; returns (struct, arrays to keep alive).

    With on_device=True the payload arguments are device pointers (ints).
    Indexed form: table is K x K_t (rows: s symbols, columns: t symbols);
    code_bytes 1 or 2 fixes the code width (0: uint8 up to 128 symbols,
    uint16 beyond).
    raw_width 1 or 4: codes_s/codes_t are raw symbols (bytes of latin-1 or
    UTF-32-LE text) that the device encodes, through lut for the indexed
    form.
    """
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    def payload(x, dt):
        if x is None:
            return None
        if on_device:
            return int(x)
        return arr(x, dt)

    rg = _lib.Ragged()
    ls = np.ascontiguousarray(len_s, dtype=np.int32)
    lt = np.ascontiguousarray(len_t, dtype=np.int32)
    keep += [ls, lt]
    rg.n = len(ls)
    rg.form = _FORMS[form]
    rg.on_device = int(bool(on_device))
    rg.len_s, rg.len_t = ls.ctypes.data, lt.ctypes.data
    if off_s is not None:
        rg.off_s = arr(off_s, np.int64)
        rg.off_t = arr(off_t, np.int64)
    if raw_width:
        rg.raw_width = int(raw_width)
        rg.codes_s = payload(codes_s, np.uint8)
        rg.codes_t = payload(codes_t, np.uint8)
        if lut is not None:
            rg.lut = arr(lut, np.uint8)
            rg.lut_len = len(lut)
    if form == "indexed":
        t = np.ascontiguousarray(table, dtype=np.float32)
        if t.ndim != 2:
            raise ValueError(f"similarity table has shape {t.shape}, expected two dimensions")
        keep.append(t)
        rg.table = t.ctypes.data
        rg.K, rg.K_t = t.shape
        if not raw_width:
            # uint8 codes for tables of up to 128 symbols (staged in LDS),
            # uint16 beyond (read from global memory); device-resident codes
            # are taken in the width the caller gives
            wide = code_bytes == 2 if code_bytes else max(t.shape) > 128
            rg.code_bytes = 2 if wide else 1
            rg.codes_s = payload(codes_s, np.uint16 if wide else np.uint8)
            rg.codes_t = payload(codes_t, np.uint16 if wide else np.uint8)
    elif form == "binary":
        if not raw_width:
            rg.codes_s = payload(codes_s, np.uint32)
            rg.codes_t = payload(codes_t, np.uint32)
        rg.eq, rg.ne = float(eq), float(ne)
    elif form == "series":
        rg.codes_s = payload(codes_s, np.float32)
        rg.codes_t = payload(codes_t, np.float32)
        rg.dim = int(dim)
    elif form == "dense":
        rg.dense = payload(dense, np.float32)
        rg.dense_off = arr(dense_off, np.int64)
        rg.dense_ld = int(dense_ld)
    return rg, keep


def _make_module(index_bits):
    mod = types.SimpleNamespace()

    def create_options(options):
        return Options(dict(options), index_bits)

    def create_solver(max_len_s, max_len_t, options):
        h = ctypes.c_void_p()
        check(lib().pa_create_solver(int(max_len_s), int(max_len_t), options._h, ctypes.byref(h)))
        solver = Solver(h, options, int(max_len_s), int(max_len_t))
        if options.gap_kind == "general":
            # GeneralGapCostSolver takes costs(max_len + 1) from each side's
            # gap object, as float32 (factory.h:20-31, solver.h:3093-3125)
            gs, gt = options.gap_objects
            ws = np.ascontiguousarray(gs.costs(int(max_len_s) + 1), dtype=np.float32)
            wt = np.ascontiguousarray(gt.costs(int(max_len_t) + 1), dtype=np.float32)
            if ws.ndim != 1 or wt.ndim != 1:
                raise ValueError("gap costs must be one-dimensional")
            check(lib().pa_solver_set_gap_costs(h, ws.ctypes.data, ws.shape[0], wt.ctypes.data, wt.shape[0]))
        return solver

    mod.create_options = create_options
    mod.create_solver = create_solver
    mod.Solver = Solver
    mod.Alignment = Alignment
    mod.AlignmentIterator = AlignmentIterator
    mod.Solution = Solution
    mod.SolutionIterator = SolutionIterator
    mod.Options = Options
    return mod


m16 = _make_module(16)  # Options<float, int16_t> (module.cpp:10-11)
m32 = _make_module(32)  # Options<float, int32_t> (module.cpp:13-14)
