"""Engine and model handles over the C ABI.

``Engine`` owns one HIP device stream and the device arenas;
``NativeModel`` is one cached Debye model whose six partial profiles stay on
the device.  ``intensity`` and ``fit`` run device kernels and return host
arrays; they may be called from several threads on one model.
"""
from __future__ import annotations

import ctypes
import threading
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N

_default = None
_default_lock = threading.Lock()


class Engine:
    def __init__(self, device: int = 0):
        L = N.load()
        h = ctypes.c_void_p()
        N.check(L.pripps_engine_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def memory(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        N.check(N.load().pripps_engine_memory(self._h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def build(self, inp: N.PreparedInput, precision="fp32", tuning: Optional[N.Tuning] = None):
        prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        h = ctypes.c_void_p()
        tp = ctypes.byref(tuning) if tuning is not None else None
        N.check(N.load().pripps_debye_model_build(self._h, ctypes.byref(inp.c), prec, tp,
                                                  ctypes.byref(h)))
        return NativeModel(h, self)

    def build_batch(self, inputs: Sequence[N.PreparedInput], precision="fp32",
                    tuning: Optional[N.Tuning] = None) -> List["NativeModel"]:
        prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        arr = (N.DebyeInput * len(inputs))(*[i.c for i in inputs])
        outs = (ctypes.c_void_p * len(inputs))()
        tp = ctypes.byref(tuning) if tuning is not None else None
        N.check(N.load().pripps_debye_model_build_batch(self._h, len(inputs), arr, prec, tp, outs))
        return [NativeModel(ctypes.c_void_p(outs[i]), self) for i in range(len(inputs))]

    def autotune(self, inp: N.PreparedInput, precision="fp32", repeats: int = 3):
        """Time candidate plans on ``inp`` and record the fastest in this
        engine's tuning table (used by later builds without explicit tuning of
        the same size class).  Returns (tuning, pair-kernel ms)."""
        prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        best = N.Tuning()
        ms = ctypes.c_double()
        N.check(N.load().pripps_engine_autotune(self._h, ctypes.byref(inp.c), prec, int(repeats),
                                                ctypes.byref(best), ctypes.byref(ms)))
        return best, ms.value

    def tuning_table(self) -> List[dict]:
        """The engine's tuning table as a list of dicts (JSON-serialisable)."""
        L = N.load()
        count = ctypes.c_int32()
        N.check(L.pripps_engine_tuning_table(self._h, None, 0, ctypes.byref(count)))
        arr = (N.TuningEntry * max(1, count.value))()
        N.check(L.pripps_engine_tuning_table(self._h, arr, count.value, ctypes.byref(count)))
        return [{f: getattr(arr[k], f) for f, _ in N.TuningEntry._fields_}
                for k in range(count.value)]

    def set_tuning_entry(self, entry: dict) -> None:
        e = N.TuningEntry(*[entry[f] for f, _ in N.TuningEntry._fields_])
        N.check(N.load().pripps_engine_set_tuning_entry(self._h, ctypes.byref(e)))

    def frames_average(self, topology: N.PreparedInput, coords, weights=None, stride: int = 1,
                       precision="fp32", tuning: Optional[N.Tuning] = None,
                       max_frames_per_launch: int = 0) -> "NativeModel":
        """Weighted frame average of the six profiles; ``coords`` is
        [n_frames][n_atoms][3] in Angstrom."""
        prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        c = np.ascontiguousarray(coords, dtype=np.float64)
        n_atoms = len(topology.x)
        if c.ndim != 3 or c.shape[1:] != (n_atoms, 3):
            from .errors import ValidationError
            raise ValidationError(f"frames must be [n_frames][{n_atoms}][3], got {list(c.shape)}")
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w is not None and len(w) != c.shape[0]:
            from .errors import ValidationError
            raise ValidationError(f"number of weights ({len(w)}) does not match number of "
                                  f"frames in trajectory ({c.shape[0]})")
        h = ctypes.c_void_p()
        tp = ctypes.byref(tuning) if tuning is not None else None
        N.check(N.load().pripps_debye_frames_average(self._h, ctypes.byref(topology.c), c.shape[0],
                                                     N.dptr(c), N.dptr(w), int(stride), prec, tp,
                                                     int(max_frames_per_launch), ctypes.byref(h)))
        return NativeModel(h, self)

    def from_profiles(self, q, profiles) -> "NativeModel":
        q = np.ascontiguousarray(q, dtype=np.float64)
        p = np.ascontiguousarray(profiles, dtype=np.float64).reshape(6, len(q))
        h = ctypes.c_void_p()
        N.check(N.load().pripps_debye_model_from_profiles(self._h, N.dptr(q), len(q), N.dptr(p),
                                                          ctypes.byref(h)))
        return NativeModel(h, self)

    def close(self):
        if self._h:
            N.load().pripps_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_engine() -> Engine:
    global _default
    with _default_lock:
        if _default is None:
            _default = Engine(0)
        return _default


class NativeModel:
    def __init__(self, handle, engine: Engine):
        self._h = handle
        self._engine = engine  # keeps the engine alive
        nq = ctypes.c_int32()
        N.check(N.load().pripps_debye_model_nq(self._h, ctypes.byref(nq)))
        self.nq = nq.value

    def q_values(self) -> np.ndarray:
        out = np.empty(self.nq)
        N.check(N.load().pripps_debye_model_q_values(self._h, N.dptr(out)))
        return out

    def profiles(self) -> np.ndarray:
        out = np.empty((6, self.nq))
        N.check(N.load().pripps_debye_model_profiles(self._h, N.dptr(out)))
        return out

    def intensity(self, volume_scale: float, contrast_density: float):
        t, a, e, h = (np.empty(self.nq) for _ in range(4))
        N.check(N.load().pripps_debye_model_intensity(self._h, float(volume_scale),
                                                      float(contrast_density), N.dptr(t),
                                                      N.dptr(a), N.dptr(e), N.dptr(h)))
        return t, a, e, h

    def fit(self, q, intensity, sigma=None, method=N.FIT_NELDER_MEAD, c1_points=151,
            c2_points=101, c1_bounds=None, c2_bounds=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        it = np.ascontiguousarray(intensity, dtype=np.float64)
        sg = None if sigma is None else np.ascontiguousarray(sigma, dtype=np.float64)
        from .errors import ValidationError
        if len(q) == 0:
            raise ValidationError("experimental data is empty")
        if len(it) != len(q):
            raise ValidationError(f"q ({len(q)}) and intensity ({len(it)}) length mismatch")
        if sg is not None and len(sg) == 0:
            sg = None
        if sg is not None and len(sg) != len(q):
            raise ValidationError(f"q ({len(q)}) and sigma ({len(sg)}) length mismatch")
        b1 = None if c1_bounds is None else np.ascontiguousarray(c1_bounds, dtype=np.float64)
        b2 = None if c2_bounds is None else np.ascontiguousarray(c2_bounds, dtype=np.float64)
        res = N.FitResultC()
        N.check(N.load().pripps_debye_model_fit(self._h, N.dptr(q), N.dptr(it), N.dptr(sg),
                                                len(q), int(method), int(c1_points),
                                                int(c2_points), N.dptr(b1), N.dptr(b2),
                                                ctypes.byref(res)))
        return res

    def fit_curve(self, q, volume_scale: float, contrast_density: float) -> np.ndarray:
        """Natural spline of the combined I(q) evaluated at ``q`` (device)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty(len(q))
        N.check(N.load().pripps_debye_model_fit_curve(self._h, float(volume_scale),
                                                      float(contrast_density), N.dptr(q), len(q),
                                                      N.dptr(out)))
        return out

    def recompute(self, repeats: int, graph: bool = False):
        """Re-run the last build's kernels on resident inputs; returns
        (mean pair-kernel ms, total device ms).  graph=True replays them as
        one captured HIP graph launch per run (the pair-kernel time is then
        not measured and returned as None)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        N.check(N.load().pripps_debye_model_recompute(self._h, int(repeats),
                                                      None if graph else ctypes.byref(a),
                                                      ctypes.byref(b)))
        return (None if graph else a.value), b.value

    def stats(self) -> N.BuildStats:
        s = N.BuildStats()
        N.check(N.load().pripps_debye_model_stats(self._h, ctypes.byref(s)))
        return s

    def close(self):
        if self._h:
            N.load().pripps_debye_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
