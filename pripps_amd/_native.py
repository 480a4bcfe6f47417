"""ctypes binding of the C ABI in include/pripps_debye.h and include/pripps_direct.h.

The shared library ``pripps_amd/libpripps_debye.so`` is built in-tree by
``make -C pripps_amd/csrc`` (``__graft_entry__.build()``).  There is no
fallback: if the library is missing every engine call raises
``ImportError``.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from .errors import from_status

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpripps_debye.so")

FP32_COMPENSATED = 0
FP64 = 1
FP64_PLAIN = 2
PRECISIONS = {"fp32": FP32_COMPENSATED, "fp32_compensated": FP32_COMPENSATED,
              "fp64": FP64, "fp64_plain": FP64_PLAIN}

FIT_NELDER_MEAD = 0
FIT_GRID = 1

# bytes of the fixed device scratch of pripps_debye_build_noalloc (per GPU)
NOALLOC_SCRATCH_BYTES = 64 << 20

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)


class DebyeInput(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int64),
        ("x", _D), ("y", _D), ("z", _D),
        ("kind", _I32),
        ("n_kinds", ctypes.c_int32),
        ("nq", ctypes.c_int32),
        ("q", _D),
        ("atom_ff", _D),
        ("excluded_ff", _D),
        ("water_ff", _D),
        ("hydration_weight", _D),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("q_per_lane", ctypes.c_int32), ("block", ctypes.c_int32),
                ("span", ctypes.c_int32), ("variant", ctypes.c_int32)]


class FitResultC(ctypes.Structure):
    _fields_ = [("volume_scale", ctypes.c_double), ("contrast_density", ctypes.c_double),
                ("chi2", ctypes.c_double), ("scale", ctypes.c_double)]


class BuildStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("pair_q_terms", ctypes.c_double), ("q_per_lane", ctypes.c_int32),
                ("block", ctypes.c_int32), ("span", ctypes.c_int32),
                ("work_items", ctypes.c_int64), ("frames", ctypes.c_int32),
                ("frames_per_launch", ctypes.c_int32), ("graph_launches", ctypes.c_int32),
                ("cheb_order", ctypes.c_int32)]


class TuningEntry(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("uniform", ctypes.c_int32),
                ("log2_atoms", ctypes.c_int32), ("log2_nq", ctypes.c_int32),
                ("q_per_lane", ctypes.c_int32), ("block", ctypes.c_int32),
                ("span_shift", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("pair_ms", ctypes.c_double)]


class DirectGrid(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("n_vectors", ctypes.c_int32),
                ("indices", _I32), ("num_molecules", ctypes.c_int32),
                ("n_groups", ctypes.c_int32), ("group", _I32), ("group_q", _D)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_in_flight", ctypes.c_int32),
                ("memory_budget", ctypes.c_uint64)]


# every exported symbol of include/pripps_debye.h and include/pripps_direct.h
EXPORTS = (
    "pripps_last_error", "pripps_version", "pripps_engine_create", "pripps_engine_destroy",
    "pripps_engine_memory", "pripps_debye_model_build", "pripps_debye_model_build_batch",
    "pripps_debye_model_nq", "pripps_debye_model_q_values", "pripps_debye_model_intensity",
    "pripps_debye_model_profiles", "pripps_debye_model_fit", "pripps_debye_model_fit_curve",
    "pripps_debye_model_from_profiles",
    "pripps_debye_model_stats", "pripps_debye_model_recompute",
    "pripps_debye_model_free", "pripps_debye_build_noalloc", "pripps_engine_autotune", "pripps_engine_tuning_table",
    "pripps_engine_set_tuning_entry", "pripps_debye_frames_average", "pripps_server_create",
    "pripps_server_submit", "pripps_server_wait", "pripps_server_destroy",
    "pripps_direct_model_build", "pripps_direct_model_n_groups", "pripps_direct_model_q_values",
    "pripps_direct_model_intensity", "pripps_direct_model_self_term",
    "pripps_direct_model_profiles", "pripps_direct_model_amplitudes", "pripps_direct_model_stats",
    "pripps_direct_model_recompute", "pripps_direct_model_free", "pripps_direct_frames_average",
)

_lib = None
_lock = threading.Lock()


def load():
    """Load the engine library (no compute happens here)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C pripps_amd/csrc` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.pripps_last_error.restype = ctypes.c_char_p
        L.pripps_version.restype = ctypes.c_char_p
        L.pripps_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.pripps_engine_destroy.argtypes = [vp]
        L.pripps_engine_memory.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64)]
        L.pripps_debye_model_build.argtypes = [vp, ctypes.POINTER(DebyeInput), ctypes.c_int32,
                                               ctypes.POINTER(Tuning), ctypes.POINTER(vp)]
        L.pripps_debye_model_build_batch.argtypes = [vp, ctypes.c_int32,
                                                     ctypes.POINTER(DebyeInput), ctypes.c_int32,
                                                     ctypes.POINTER(Tuning), ctypes.POINTER(vp)]
        L.pripps_debye_model_nq.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.pripps_debye_model_q_values.argtypes = [vp, _D]
        L.pripps_debye_model_intensity.argtypes = [vp, ctypes.c_double, ctypes.c_double, _D, _D,
                                                   _D, _D]
        L.pripps_debye_model_profiles.argtypes = [vp, _D]
        L.pripps_debye_model_fit.argtypes = [vp, _D, _D, _D, ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32, _D, _D,
                                             ctypes.POINTER(FitResultC)]
        L.pripps_debye_model_fit_curve.argtypes = [vp, ctypes.c_double, ctypes.c_double, _D,
                                                   ctypes.c_int64, _D]
        L.pripps_debye_model_from_profiles.argtypes = [vp, _D, ctypes.c_int32, _D,
                                                       ctypes.POINTER(vp)]
        L.pripps_debye_model_recompute.argtypes = [vp, ctypes.c_int32,
                                                   ctypes.POINTER(ctypes.c_double),
                                                   ctypes.POINTER(ctypes.c_double)]
        L.pripps_debye_model_stats.argtypes = [vp, ctypes.POINTER(BuildStats)]
        L.pripps_debye_model_free.argtypes = [vp]
        L.pripps_debye_build_noalloc.argtypes = [vp, ctypes.POINTER(DebyeInput), ctypes.c_int32,
                                                 ctypes.POINTER(Tuning), _D]
        L.pripps_engine_autotune.argtypes = [vp, ctypes.POINTER(DebyeInput), ctypes.c_int32,
                                             ctypes.c_int32, ctypes.POINTER(Tuning),
                                             ctypes.POINTER(ctypes.c_double)]
        L.pripps_engine_tuning_table.argtypes = [vp, ctypes.POINTER(TuningEntry), ctypes.c_int32,
                                                 ctypes.POINTER(ctypes.c_int32)]
        L.pripps_engine_set_tuning_entry.argtypes = [vp, ctypes.POINTER(TuningEntry)]
        L.pripps_debye_frames_average.argtypes = [vp, ctypes.POINTER(DebyeInput), ctypes.c_int32,
                                                  _D, _D, ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.POINTER(Tuning), ctypes.c_int32,
                                                  ctypes.POINTER(vp)]
        L.pripps_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.pripps_server_submit.argtypes = [vp, ctypes.POINTER(DebyeInput),
                                           ctypes.POINTER(ctypes.c_int64)]
        L.pripps_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(vp),
                                         ctypes.POINTER(ctypes.c_double)]
        L.pripps_server_destroy.argtypes = [vp]
        L.pripps_direct_model_build.argtypes = [vp, ctypes.POINTER(DebyeInput),
                                                ctypes.POINTER(DirectGrid), ctypes.c_int32,
                                                ctypes.POINTER(vp)]
        L.pripps_direct_model_n_groups.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.pripps_direct_model_q_values.argtypes = [vp, _D]
        L.pripps_direct_model_intensity.argtypes = [vp, ctypes.c_double, ctypes.c_double, _D, _D,
                                                    _D, _D]
        L.pripps_direct_model_self_term.argtypes = [vp, _D]
        L.pripps_direct_model_profiles.argtypes = [vp, _D]
        L.pripps_direct_model_amplitudes.argtypes = [vp, _D]
        L.pripps_direct_model_stats.argtypes = [vp, ctypes.POINTER(BuildStats)]
        L.pripps_direct_model_recompute.argtypes = [vp, ctypes.c_int32,
                                                    ctypes.POINTER(ctypes.c_double)]
        L.pripps_direct_model_free.argtypes = [vp]
        L.pripps_direct_frames_average.argtypes = [vp, ctypes.POINTER(DebyeInput),
                                                   ctypes.POINTER(DirectGrid), ctypes.c_int32,
                                                   _D, _D, ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.c_double, ctypes.c_double, _D, _D]
        for name in EXPORTS:
            if name not in ("pripps_last_error", "pripps_version"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
        return L


def check(rc: int) -> None:
    if rc != 0:
        msg = load().pripps_last_error().decode("utf-8", "replace")
        raise from_status(rc, msg)


def dptr(a):
    return None if a is None else a.ctypes.data_as(_D)


class PreparedInput:
    """Keeps the numpy arrays of one DebyeInput alive."""

    def __init__(self, pos, kind, q, atom_ff, excluded_ff=None, water_ff=None, weights=None):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        self.x = np.ascontiguousarray(pos[:, 0])
        self.y = np.ascontiguousarray(pos[:, 1])
        self.z = np.ascontiguousarray(pos[:, 2])
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        self.q = np.ascontiguousarray(q, dtype=np.float64)
        self.atom_ff = np.ascontiguousarray(atom_ff, dtype=np.float64)
        self.excluded_ff = None if excluded_ff is None else np.ascontiguousarray(excluded_ff, dtype=np.float64)
        self.water_ff = None if water_ff is None else np.ascontiguousarray(water_ff, dtype=np.float64)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        n_kinds = self.atom_ff.shape[0] if self.atom_ff.ndim == 2 else 1
        self.c = DebyeInput(len(self.x), dptr(self.x), dptr(self.y), dptr(self.z),
                            self.kind.ctypes.data_as(_I32), n_kinds, len(self.q), dptr(self.q),
                            dptr(self.atom_ff), dptr(self.excluded_ff), dptr(self.water_ff),
                            dptr(self.weights))


def tuning(q_per_lane=0, block=0, span=0, variant=0):
    return Tuning(int(q_per_lane), int(block), int(span), int(variant))
