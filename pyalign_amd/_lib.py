"""Loader for libpyalign_hip.so (the C ABI declared in include/pyalign_hip.h).

There is one backend: the HIP library built for gfx950.  If it is missing
the import fails loudly; there is no CPU fallback.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpyalign_hip.so")

PA_OK, PA_EINVAL, PA_ERUNTIME, PA_ELENGTH, PA_EHIP = 0, 1, 2, 3, 4

PA_FORM_DENSE, PA_FORM_INDEXED, PA_FORM_BINARY, PA_FORM_SERIES = 0, 1, 2, 3
PA_SCHED_AUTO, PA_SCHED_PLAIN = 0, 1
PA_GAP_NONE, PA_GAP_LINEAR, PA_GAP_AFFINE, PA_GAP_GENERAL = 0, 1, 2, 3


class OptionsDesc(ctypes.Structure):
    _fields_ = [
        ("solver_type", ctypes.c_int32), ("batch", ctypes.c_int32),
        ("direction", ctypes.c_int32), ("return_dup", ctypes.c_int32),
        ("detail", ctypes.c_int32), ("count", ctypes.c_int32),
        ("locality", ctypes.c_int32), ("gap_kind_s", ctypes.c_int32),
        ("gap_kind_t", ctypes.c_int32), ("index_bits", ctypes.c_int32),
        ("gap_s", ctypes.c_double * 2), ("gap_t", ctypes.c_double * 2),
    ]


class Ragged(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("form", ctypes.c_int32), ("on_device", ctypes.c_int32),
        ("codes_s", ctypes.c_void_p), ("codes_t", ctypes.c_void_p),
        ("off_s", ctypes.c_void_p), ("off_t", ctypes.c_void_p),
        ("len_s", ctypes.c_void_p), ("len_t", ctypes.c_void_p),
        ("table", ctypes.c_void_p), ("K", ctypes.c_int32), ("dim", ctypes.c_int32),
        ("eq", ctypes.c_float), ("ne", ctypes.c_float),
        ("dense", ctypes.c_void_p), ("dense_off", ctypes.c_void_p),
        ("dense_ld", ctypes.c_int64),
        ("raw_width", ctypes.c_int32), ("lut_len", ctypes.c_int32), ("lut", ctypes.c_void_p),
        ("K_t", ctypes.c_int32), ("code_bytes", ctypes.c_int32),
    ]


# every symbol the header declares, with (restype, argtypes)
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F = ctypes.c_float
SIGNATURES = {
    "pa_last_error": (ctypes.c_char_p, []),
    "pa_device_count": (ctypes.c_int, []),
    "pa_create_options": (ctypes.c_int, [ctypes.POINTER(OptionsDesc), ctypes.POINTER(_P)]),
    "pa_options_free": (None, [_P]),
    "pa_create_solver": (ctypes.c_int, [_I64, _I64, _P, ctypes.POINTER(_P)]),
    "pa_solver_free": (None, [_P]),
    "pa_solver_batch_size": (ctypes.c_int, [_P]),
    "pa_solve_dense": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, ctypes.POINTER(_P)]),
    "pa_solve_indexed": (ctypes.c_int, [_P, _P, _I64, _P, _I64, _I64, _P, _I64, _I64, _P, ctypes.POINTER(_P)]),
    "pa_solve_binary": (ctypes.c_int, [_P, _P, _I64, _P, _I64, _I64, _F, _F, _P, ctypes.POINTER(_P)]),
    "pa_solve_ragged": (ctypes.c_int, [_P, ctypes.POINTER(Ragged), _I32, ctypes.POINTER(_P)]),
    "pa_solve_ragged_scores_device": (ctypes.c_int, [_P, ctypes.POINTER(Ragged), _I32, _P]),
    "pa_solver_sync": (ctypes.c_int, [_P]),
    "pa_solve_ragged_scores_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(Ragged), _I32, _P, _P]),
    "pa_solver_last_kernel_ms": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_double)]),
    "pa_solver_set_graphs": (ctypes.c_int, [_P, _I32]),
    "pa_solver_set_checkpoint": (ctypes.c_int, [_P, _I64]),
    "pa_solver_set_stream": (ctypes.c_int, [_P, _P, _I32]),
    "pa_solver_stream": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_void_p)]),
    "pa_solver_set_gap_costs": (ctypes.c_int, [_P, _P, _I64, _P, _I64]),
    "pa_solver_set_gap_params": (ctypes.c_int, [_P, _P, _P]),
    "pa_solver_last_bad_symbol": (ctypes.c_int, [_P, _P, _P]),
    "pa_solver_last_schedule": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "pa_forward_ragged": (ctypes.c_int, [_P, ctypes.POINTER(Ragged), _P, ctypes.POINTER(_P)]),
    "pa_backward": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "pa_tape_free": (None, [_P]),
    "pa_result_count": (_I64, [_P]),
    "pa_result_scores": (ctypes.c_int, [_P, _P]),
    "pa_result_has_path": (ctypes.c_int, [_P, _I64]),
    "pa_result_path_len": (_I64, [_P, _I64]),
    "pa_result_path": (ctypes.c_int, [_P, _I64, _P]),
    "pa_result_alignment": (ctypes.c_int, [_P, _I64, _P, _P]),
    "pa_result_iter_next": (ctypes.c_int, [_P, _I64, ctypes.POINTER(_F)]),
    "pa_result_iter_path_len": (_I64, [_P, _I64]),
    "pa_result_iter_path": (ctypes.c_int, [_P, _I64, _P]),
    "pa_result_free": (None, [_P]),
    "pa_solve_ragged_matrices": (ctypes.c_int, [_P, ctypes.POINTER(Ragged), ctypes.POINTER(_P)]),
    "pa_result_layers": (_I32, [_P]),
    "pa_result_values": (ctypes.c_int, [_P, _I64, _P]),
    "pa_result_paths": (ctypes.c_int, [_P, _P, _P]),
    "pa_result_alignments": (ctypes.c_int, [_P, _P, _P]),
    "pa_result_traceback": (ctypes.c_int, [_P, _I64, _P]),
    "pa_result_edges_len": (_I64, [_P, _I64, _I32]),
    "pa_result_edges": (ctypes.c_int, [_P, _I64, _I32, _P]),
}

_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(hipcc for gfx950); pyalign_amd has no CPU fallback")
        # One HIP runtime per process: torch's build bundles its own (same
        # soname, loaded through a different name), and a process that loads
        # this library's first cannot open the GPU from torch afterwards.
        # Loading torch first makes this library bind to torch's copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Raise the reference's exception type for a non-zero status."""
    if status == PA_OK:
        return
    msg = lib().pa_last_error().decode("utf-8", "replace")
    if status == PA_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)
