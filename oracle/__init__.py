"""CPU oracle for the DP alignment solver -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``oracle/pa_oracle.cpp``, a restatement of the
reference solver (pyalign/algorithm/solver.h).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module, and only as the checker or the timed CPU baseline.  The product
package ``pyalign_amd`` never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# enum values: pyalign/algorithm/enum.h:6-31
ALIGNMENT, DTW = 0, 1
MINIMIZE, MAXIMIZE = 0, 1
LOCAL, GLOBAL, SEMIGLOBAL = 0, 1, 2
GOAL_SCORE, GOAL_ONE, GOAL_ALL = 0, 1, 2
GAP_LINEAR, GAP_AFFINE, GAP_GENERAL = 0, 1, 2


class OcConfig(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32), ("direction", ctypes.c_int32),
        ("locality", ctypes.c_int32), ("goal", ctypes.c_int32),
        ("gap_kind", ctypes.c_int32), ("index_bits", ctypes.c_int32),
        ("remove_dup", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("lin_s", ctypes.c_float), ("lin_t", ctypes.c_float),
        ("aff_u_s", ctypes.c_float), ("aff_v_s", ctypes.c_float),
        ("aff_u_t", ctypes.c_float), ("aff_v_t", ctypes.c_float),
        ("lim_s", ctypes.c_int64), ("lim_t", ctypes.c_int64),
        ("gen_s", ctypes.c_void_p), ("gen_t", ctypes.c_void_p),
    ]


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = line.split()
                    return all(x in flags for x in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))
    except OSError:
        pass
    return False


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        name = "libpa_oracle_v4.so" if _cpu_has_avx512() else "libpa_oracle.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.oc_solve_dense.restype = ctypes.c_void_p
        L.oc_solve_dense.argtypes = [P(OcConfig), P(ctypes.c_float), ctypes.c_int64, ctypes.c_int64]
        L.oc_free.argtypes = [ctypes.c_void_p]
        L.oc_last_error.restype = ctypes.c_char_p
        for fn in ("oc_score",):
            getattr(L, fn).restype = ctypes.c_float
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        for fn in ("oc_has_seed", "oc_all_next", "oc_layers"):
            getattr(L, fn).restype = ctypes.c_int32
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        for fn in ("oc_path_len", "oc_edges_len"):
            getattr(L, fn).restype = ctypes.c_int64
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        for fn in ("oc_path", "oc_edges"):
            getattr(L, fn).restype = None
            getattr(L, fn).argtypes = [ctypes.c_void_p, P(ctypes.c_int32)]
        L.oc_values.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_float)]
        L.oc_traceback.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_int32)]
        L.oc_tb_edges_len.restype = ctypes.c_int64
        L.oc_tb_edges_len.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.oc_tb_edges.restype = None
        L.oc_tb_edges.argtypes = [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_int32)]
        L.oc_baseline_indexed_score.restype = ctypes.c_double
        L.oc_baseline_indexed_score.argtypes = [
            P(OcConfig), P(ctypes.c_uint8), P(ctypes.c_int64), P(ctypes.c_int64),
            P(ctypes.c_uint8), P(ctypes.c_int64), P(ctypes.c_int64),
            P(ctypes.c_float), ctypes.c_int64,
            P(ctypes.c_int32), P(ctypes.c_int64), ctypes.c_int64, ctypes.c_int32,
            P(ctypes.c_float)]
        _lib = L
    return _lib


def next_power_of_2(x):
    return 1 if x == 0 else 2 ** (int(x) - 1).bit_length()


def rounded_limit(x, limit=1024):
    """SizeRounder (pyalign/solve.py:141-153)."""
    return x if x >= limit else next_power_of_2(x)


def make_config(locality=GLOBAL, direction=MAXIMIZE, goal=GOAL_ONE, gap=("linear", 0.0),
                solver_type=ALIGNMENT, index_bits=16, remove_dup=True, lim_s=1, lim_t=1):
    """gap: ("linear", u) | ("affine", open, extend) | ("general", w_s, w_t)
    | None (= linear 0).  General costs are the vectors w(k), k = 0 .. lim,
    of each side (GapCost.costs(lim + 1) as float32).

    Affine follows pyalign.gaps.AffineGapCost (u = open - extend in float64,
    v = extend) and the u/v flip of factory.h:48-52 into
    AffineCost(u = float32(extend), v = float32(open - extend)).
    """
    c = OcConfig()
    c.type, c.direction, c.locality, c.goal = solver_type, direction, locality, goal
    c.index_bits, c.remove_dup = index_bits, int(bool(remove_dup))
    c.lim_s, c.lim_t = lim_s, lim_t
    if gap is None:
        gap = ("linear", 0.0)
    if gap[0] == "linear":
        c.gap_kind = GAP_LINEAR
        c.lin_s = c.lin_t = float(np.float32(gap[1]))
    elif gap[0] == "affine":
        open_, extend = gap[1], gap[2]
        c.gap_kind = GAP_AFFINE
        u = float(np.float32(extend))
        v = float(np.float32(float(open_) - float(extend)))
        c.aff_u_s = c.aff_u_t = u
        c.aff_v_s = c.aff_v_t = v
    elif gap[0] == "general":
        ws = np.ascontiguousarray(gap[1], dtype=np.float32)
        wt = np.ascontiguousarray(gap[2], dtype=np.float32)
        if ws.shape != (lim_s + 1,) or wt.shape != (lim_t + 1,):
            raise ValueError("general gap costs need lim + 1 entries per side")
        c.gap_kind = GAP_GENERAL
        c._keep = (ws, wt)
        c.gen_s, c.gen_t = ws.ctypes.data, wt.ctypes.data
    else:
        raise ValueError(gap)
    return c


class OracleResult:
    """One problem solved by the oracle (dense similarity)."""

    def __init__(self, cfg, sim):
        sim = np.ascontiguousarray(sim, dtype=np.float32)
        self._sim = sim
        self.len_s, self.len_t = sim.shape
        L = lib()
        h = L.oc_solve_dense(ctypes.byref(cfg), sim.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                             self.len_s, self.len_t)
        if not h:
            raise RuntimeError(L.oc_last_error().decode())
        self._h = h
        self._cfg = cfg

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oc_free(self._h)
            self._h = None

    @property
    def has_seed(self):
        return bool(lib().oc_has_seed(self._h))

    @property
    def score(self):
        return np.float32(lib().oc_score(self._h))

    def path(self):
        L = lib()
        n = L.oc_path_len(self._h)
        out = np.empty((n, 2), dtype=np.int32)
        if n:
            L.oc_path(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def edges(self):
        L = lib()
        n = L.oc_edges_len(self._h)
        out = np.empty((n, 2), dtype=np.int32)
        if n:
            L.oc_edges(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def all_next(self):
        """Advance the all-optimal iterator; returns (score, path, edges) or None."""
        if not lib().oc_all_next(self._h):
            return None
        return self.score, self.path(), self.edges()

    def values(self):
        L = lib()
        k = L.oc_layers(self._h)
        out = np.empty((k, self.len_s + 1, self.len_t + 1), dtype=np.float32)
        for layer in range(k):
            L.oc_values(self._h, layer, out[layer].ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out

    def traceback_matrix(self):
        L = lib()
        k = L.oc_layers(self._h)
        out = np.empty((k, self.len_s + 1, self.len_t + 1, 2), dtype=np.int32)
        for layer in range(k):
            L.oc_traceback(self._h, layer, out[layer].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return out

    def traceback_edges(self):
        """Per layer: int32[n, 2, 2] of ((i-1, j-1), (u, v)) traceback entries."""
        L = lib()
        out = []
        for layer in range(L.oc_layers(self._h)):
            n = L.oc_tb_edges_len(self._h, layer)
            e = np.empty((n, 2, 2), dtype=np.int32)
            if n:
                L.oc_tb_edges(self._h, layer, e.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
            out.append(e)
        return out


def solve(sim, **kw):
    """Solve one problem given its dense similarity (len_s x len_t)."""
    sim = np.asarray(sim, dtype=np.float32)
    kw.setdefault("lim_s", rounded_limit(sim.shape[0]))
    kw.setdefault("lim_t", rounded_limit(sim.shape[1]))
    return OracleResult(make_config(**kw), sim)


def all_alignments(sim, **kw):
    """All optimal alignments (DFS order of the reference) as list of (score, path, edges)."""
    kw["goal"] = GOAL_ALL
    r = solve(sim, **kw)
    out = []
    while True:
        x = r.all_next()
        if x is None:
            break
        out.append(x)
    return out


def baseline_indexed_score(cfg, codes_s, off_s, len_s, codes_t, off_t, len_t, sim,
                           batch_members, batch_off, threads):
    """Time the reference-style batched (16-lane) score path on `threads` threads."""
    L = lib()
    P = ctypes.POINTER
    n = len(len_s)
    scores = np.empty(n, dtype=np.float32)
    sim = np.ascontiguousarray(sim, dtype=np.float32)
    args = [np.ascontiguousarray(x) for x in (codes_s, off_s, len_s, codes_t, off_t, len_t)]
    bm = np.ascontiguousarray(batch_members, dtype=np.int32)
    bo = np.ascontiguousarray(batch_off, dtype=np.int64)
    t = L.oc_baseline_indexed_score(
        ctypes.byref(cfg),
        args[0].ctypes.data_as(P(ctypes.c_uint8)), args[1].ctypes.data_as(P(ctypes.c_int64)),
        args[2].ctypes.data_as(P(ctypes.c_int64)),
        args[3].ctypes.data_as(P(ctypes.c_uint8)), args[4].ctypes.data_as(P(ctypes.c_int64)),
        args[5].ctypes.data_as(P(ctypes.c_int64)),
        sim.ctypes.data_as(P(ctypes.c_float)), sim.shape[1],
        bm.ctypes.data_as(P(ctypes.c_int32)), bo.ctypes.data_as(P(ctypes.c_int64)),
        len(bo) - 1, threads, scores.ctypes.data_as(P(ctypes.c_float)))
    if t < 0:
        raise RuntimeError(L.oc_last_error().decode())
    return t, scores
