"""Serving loop for Debye solves (continuous batching, admission by free HBM).

``Server.submit`` copies a prepared input and returns a request id at once;
``Server.wait`` returns the finished ``NativeModel`` (profiles resident on the
GPU) and the request's latency from submit to completion.  Requests that
arrive while a batch runs start in the next batch, on another of three
device lanes, and overlap on the GPU.  A structure's profiles are bitwise the
same whether it runs alone or batched with others.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

from . import _native as N
from .engine import Engine, NativeModel


class Server:
    def __init__(self, engine: Engine, precision: str = "fp32", max_in_flight: int = 0,
                 memory_budget: int = 0):
        prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        cfg = N.ServerConfig(prec, int(max_in_flight), int(memory_budget))
        h = ctypes.c_void_p()
        N.check(N.load().pripps_server_create(engine.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._engine = engine

    def submit(self, inp: N.PreparedInput) -> int:
        rid = ctypes.c_int64()
        N.check(N.load().pripps_server_submit(self._h, ctypes.byref(inp.c), ctypes.byref(rid)))
        return rid.value

    def wait(self, request_id: int, timeout_ms: int = -1) -> Tuple[NativeModel, float]:
        h = ctypes.c_void_p()
        lat = ctypes.c_double()
        N.check(N.load().pripps_server_wait(self._h, int(request_id), int(timeout_ms),
                                            ctypes.byref(h), ctypes.byref(lat)))
        return NativeModel(h, self._engine), lat.value

    def close(self):
        if self._h:
            N.load().pripps_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
