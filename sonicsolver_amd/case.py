"""SonicFile case directories: ctypes over include/sonic_amd_case.h.

    c = Case("path/to/case")      # reads case.yaml and the registries
    print(c.describe())
    state, steps, t = c.run(output_dir="path/to/case/result")   # device run
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import GridDesc, _check, load
from ._native import RunControl

_P = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p

_declared = False


def _lib():
    global _declared
    lib = load()
    if not _declared:
        sig = {
            "sf_case_open": [ctypes.c_char_p, ctypes.POINTER(_VP)],
            "sf_case_close": [_VP],
            "sf_case_grid": [_VP, ctypes.POINTER(GridDesc)],
            "sf_case_state": [_VP, _P],
            "sf_case_numerics": [_VP, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p,
                                 ctypes.c_size_t, _P, _P],
            "sf_case_run_control": [_VP, ctypes.POINTER(RunControl)],
            "sf_case_describe": [_VP, ctypes.c_char_p, ctypes.c_size_t],
            "sf_case_run": [_VP, ctypes.c_char_p, _P, ctypes.POINTER(ctypes.c_int), _P],
            "sf_case_write_vts": [_VP, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), _P, _P],
            "sf_case_mesh": [_VP, ctypes.POINTER(ctypes.c_int), _P, _P],
            "sf_case_boundary_sets": [_VP, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t],
            "sf_case_set_mask": [_VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint8)],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _declared = True
    return lib


class Case:
    def __init__(self, directory: str):
        self._lib = _lib()
        self._h = _VP()
        _check(self._lib.sf_case_open(str(directory).encode(), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            self._lib.sf_case_close(self._h)
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

    def grid(self) -> GridDesc:
        d = GridDesc()
        _check(self._lib.sf_case_grid(self._h, ctypes.byref(d)))
        return d

    def nodes(self) -> int:
        d = self.grid()
        return d.n[0] * d.n[1] * d.n[2]

    def state(self) -> np.ndarray:
        out = np.empty((5, self.nodes()))
        _check(self._lib.sf_case_state(self._h, out.ctypes.data_as(_P)))
        return out

    def numerics(self):
        conv = ctypes.create_string_buffer(64)
        time = ctypes.create_string_buffer(64)
        cfl = ctypes.c_double(0.0)
        max_dt = ctypes.c_double(0.0)
        _check(self._lib.sf_case_numerics(self._h, conv, 64, time, 64, ctypes.byref(cfl),
                                          ctypes.byref(max_dt)))
        return conv.value.decode(), time.value.decode(), cfl.value, max_dt.value

    def run_control(self) -> dict:
        rc = RunControl()
        _check(self._lib.sf_case_run_control(self._h, ctypes.byref(rc)))
        return {name: getattr(rc, name) for name, _ in rc._fields_}

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(1 << 16)
        _check(self._lib.sf_case_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    def write_vts(self, output_dir: str, states, steps, times):
        """Write states (each (5, N) as state() returns it) as the output
        files of the given steps / times, as run() writes them (host only)."""
        arr = np.ascontiguousarray(np.stack([np.asarray(s, dtype=np.float64).reshape(5, -1)
                                             for s in states]))
        st = (ctypes.c_int * len(steps))(*[int(s) for s in steps])
        tm = np.ascontiguousarray(times, dtype=np.float64)
        _check(self._lib.sf_case_write_vts(self._h, str(output_dir).encode(), len(steps), st,
                                           tm.ctypes.data_as(_P), arr.ctypes.data_as(_P)))

    def mesh(self):
        """(block nodes per axis, (8, 3) scaled vertices, gas constant)."""
        nodes = (ctypes.c_int * 3)()
        vert = np.empty(24)
        r = ctypes.c_double(0.0)
        _check(self._lib.sf_case_mesh(self._h, nodes, vert.ctypes.data_as(_P), ctypes.byref(r)))
        return tuple(nodes), vert.reshape(8, 3), r.value

    def boundary_sets(self, field: str) -> list:
        buf = ctypes.create_string_buffer(4096)
        _check(self._lib.sf_case_boundary_sets(self._h, field.encode(), buf, len(buf)))
        text = buf.value.decode()
        return text.split(",") if text else []

    def set_mask(self, name: str) -> np.ndarray:
        nodes, _, _ = self.mesh()
        out = np.zeros(nodes[0] * nodes[1] * nodes[2], dtype=np.uint8)
        _check(self._lib.sf_case_set_mask(self._h, name.encode(),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def run(self, output_dir: str | None = None):
        """Run on the device; returns (final state, steps, end time)."""
        out = np.empty((5, self.nodes()))
        steps = ctypes.c_int(0)
        t = ctypes.c_double(0.0)
        _check(self._lib.sf_case_run(self._h, output_dir.encode() if output_dir else None,
                                     out.ctypes.data_as(_P), ctypes.byref(steps), ctypes.byref(t)))
        return out, steps.value, t.value
