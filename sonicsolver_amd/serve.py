"""Served solves: ctypes over include/sonic_amd_serve.h.

A :class:`Server` owns a device arena and a loop thread (native code) that
admits submitted problems as memory and slots free up, advances them
together in the persistent batch kernels and hands each back when its step
count is reached.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import _check, load
from .batch import BatchItem, _lib as _batch_lib

_P = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p
_LL = ctypes.c_longlong


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", _LL), ("completed", _LL), ("failed", _LL), ("queued", _LL),
                ("running", _LL), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("solves_per_s", ctypes.c_double), ("cell_updates", ctypes.c_double),
                ("busy_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("retire_s", ctypes.c_double), ("ticks", _LL), ("peak_items", ctypes.c_int),
                ("peak_nodes", ctypes.c_size_t), ("arena_nodes", ctypes.c_size_t)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_declared = False


def _lib():
    global _declared
    lib = load()
    _batch_lib()
    if not _declared:
        sig = {
            "sf_server_create": [ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(_VP)],
            "sf_server_destroy": [_VP],
            "sf_server_submit": [_VP, ctypes.POINTER(BatchItem), _P, _LL, ctypes.POINTER(_LL)],
            "sf_server_poll": [_VP, _LL, ctypes.POINTER(ctypes.c_int)],
            "sf_server_wait": [_VP, _LL, ctypes.c_double, ctypes.POINTER(ctypes.c_int)],
            "sf_server_result": [_VP, _LL, _P, _P, ctypes.POINTER(_LL), _P],
            "sf_server_drain": [_VP, ctypes.c_double],
            "sf_server_stats": [_VP, ctypes.POINTER(ServeStats)],
            "sf_server_reset_stats": [_VP],
            "sf_server_set_shape": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
            "sf_server_set_launch": [_VP, ctypes.c_int],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _declared = True
    return lib


class Server:
    def __init__(self, device_bytes: int, max_items: int = 1024, steps_per_tick: int = 25,
                 gamma: float = 1.4, device: int = 0):
        self._lib = _lib()
        self._h = _VP()
        _check(self._lib.sf_server_create(device, gamma, int(device_bytes), int(max_items),
                                          int(steps_per_tick), ctypes.byref(self._h)))
        self._shapes = {}

    def close(self):
        if self._h:
            self._lib.sf_server_destroy(self._h)
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, item: BatchItem, soa: np.ndarray, steps: int) -> int:
        n = item.n[0] * max(1, item.n[1])
        a = np.ascontiguousarray(soa, dtype=np.float64).reshape(5, n)
        t = _LL(0)
        _check(self._lib.sf_server_submit(self._h, ctypes.byref(item), a.ctypes.data_as(_P),
                                          int(steps), ctypes.byref(t)))
        self._shapes[t.value] = n
        return t.value

    def poll(self, ticket: int) -> int:
        st = ctypes.c_int(0)
        _check(self._lib.sf_server_poll(self._h, ticket, ctypes.byref(st)))
        return st.value

    def wait(self, ticket: int, timeout: float = 60.0) -> int:
        st = ctypes.c_int(0)
        _check(self._lib.sf_server_wait(self._h, ticket, float(timeout), ctypes.byref(st)))
        return st.value

    def result(self, ticket: int):
        """(state, time, steps, latency_ms); releases the ticket."""
        n = self._shapes.pop(ticket)
        out = np.empty((5, n))
        t = ctypes.c_double(0.0)
        k = _LL(0)
        lat = ctypes.c_double(0.0)
        _check(self._lib.sf_server_result(self._h, ticket, out.ctypes.data_as(_P), ctypes.byref(t),
                                          ctypes.byref(k), ctypes.byref(lat)))
        return out, t.value, k.value, lat.value

    def drain(self, timeout: float = 600.0):
        _check(self._lib.sf_server_drain(self._h, float(timeout)))

    def stats(self) -> dict:
        s = ServeStats()
        _check(self._lib.sf_server_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _check(self._lib.sf_server_reset_stats(self._h))

    def set_launch(self, use_graph: int = 2):
        """0 direct launches, 1 replayed graphs, 2 automatic (default)."""
        _check(self._lib.sf_server_set_launch(self._h, int(use_graph)))

    def set_shape(self, xlen: int = 255, ylen: int = 32, blocks_per_cu: int = 16,
                  fused: bool = False):
        _check(self._lib.sf_server_set_shape(self._h, int(xlen), int(ylen), int(blocks_per_cu),
                                             1 if fused else 0))
