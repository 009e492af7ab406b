"""Batched solves of independent 1D / 2D problems (ctypes over the sf_batch_*
entry points of include/sonic_amd.h).

Every problem has its own size, scheme, flux, CFL number, dt and clock;
one call of :meth:`Batch.step` advances every live problem by SSPRK3 steps
with persistent kernels over a tile table, optionally replaying a captured
hipGraph.  Boundaries are transmissive on every side.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import FLUXES, SCHEMES, _check, load

_P = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p

BYTES_PER_NODE = 4 * 5 * 8  # three state buffers + flux accumulator, 5 FP64 each


class Tuning(ctypes.Structure):
    _fields_ = [("xlen", ctypes.c_int), ("ylen", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("fused", ctypes.c_int), ("ms_per_step", ctypes.c_double),
                ("candidates", ctypes.c_int)]


class BatchItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 2), ("d", ctypes.c_double * 2), ("scheme", ctypes.c_int),
                ("flux", ctypes.c_int), ("cfl", ctypes.c_double)]


def make_item(n, d, scheme="weno5", flux="rusanov", cfl=0.5) -> BatchItem:
    n = list(n) + [1] * (2 - len(n))
    d = list(d) + [1.0] * (2 - len(d))
    it = BatchItem()
    it.n[:] = [int(x) for x in n]
    it.d[:] = [float(x) for x in d]
    it.scheme = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
    it.flux = FLUXES[flux] if isinstance(flux, str) else int(flux)
    it.cfl = float(cfl)
    return it


_declared = False


def _lib():
    global _declared
    lib = load()
    if not _declared:
        sig = {
            "sf_batch_item_nodes": [ctypes.POINTER(BatchItem), ctypes.POINTER(ctypes.c_size_t)],
            "sf_batch_create": [ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_int,
                                ctypes.POINTER(_VP)],
            "sf_batch_destroy": [_VP],
            "sf_batch_add": [_VP, ctypes.POINTER(BatchItem), _P, ctypes.c_longlong,
                             ctypes.POINTER(ctypes.c_int)],
            "sf_batch_remove": [_VP, ctypes.c_int],
            "sf_batch_download": [_VP, ctypes.c_int, _P],
            "sf_batch_step": [_VP, ctypes.c_int, ctypes.c_int],
            "sf_batch_clock": [_VP, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_longlong)],
            "sf_batch_steps": [_VP, ctypes.POINTER(ctypes.c_longlong)],
            "sf_batch_memory": [_VP, ctypes.POINTER(ctypes.c_size_t),
                                ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)],
            "sf_batch_set_tiles": [_VP, ctypes.c_int, ctypes.c_int],
            "sf_batch_set_occupancy": [_VP, ctypes.c_int],
            "sf_batch_set_schedule": [_VP, ctypes.c_int],
            "sf_batch_faults": [_VP, ctypes.POINTER(ctypes.c_uint32)],
            "sf_batch_autotune": [_VP, ctypes.c_int, ctypes.POINTER(Tuning)],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _declared = True
    return lib


def item_nodes(item: BatchItem) -> int:
    """Padded device nodes the problem occupies in a batch arena."""
    n = ctypes.c_size_t(0)
    _check(_lib().sf_batch_item_nodes(ctypes.byref(item), ctypes.byref(n)))
    return n.value


class Batch:
    """A device arena of `node_capacity` padded nodes holding up to
    `item_capacity` problems."""

    def __init__(self, node_capacity: int, item_capacity: int, gamma: float = 1.4,
                 device: int = 0):
        self._lib = _lib()
        self._h = _VP()
        _check(self._lib.sf_batch_create(device, gamma, int(node_capacity), int(item_capacity),
                                         ctypes.byref(self._h)))
        self.item_capacity = int(item_capacity)
        self._items = {}

    @classmethod
    def for_budget(cls, device_bytes: int, item_capacity: int, **kw) -> "Batch":
        """A batch whose arena takes `device_bytes` of device memory."""
        return cls(max(64, int(device_bytes) // BYTES_PER_NODE), item_capacity, **kw)

    def close(self):
        if self._h:
            self._lib.sf_batch_destroy(self._h)
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

    def add(self, item: BatchItem, soa: np.ndarray, max_steps: int = 0) -> int:
        n = item.n[0] * max(1, item.n[1])
        a = np.ascontiguousarray(soa, dtype=np.float64).reshape(5, n)
        slot = ctypes.c_int(-1)
        _check(self._lib.sf_batch_add(self._h, ctypes.byref(item), a.ctypes.data_as(_P),
                                      int(max_steps), ctypes.byref(slot)))
        self._items[slot.value] = (item.n[0], max(1, item.n[1]))
        return slot.value

    def try_add(self, item: BatchItem, soa: np.ndarray, max_steps: int = 0):
        """Admit the problem if the arena has room; None otherwise."""
        try:
            return self.add(item, soa, max_steps)
        except RuntimeError as e:
            if getattr(e, "code", 0) == -4:
                return None
            raise

    def remove(self, slot: int):
        _check(self._lib.sf_batch_remove(self._h, slot))
        self._items.pop(slot, None)

    def download(self, slot: int) -> np.ndarray:
        nx, ny = self._items[slot]
        out = np.empty((5, nx * ny))
        _check(self._lib.sf_batch_download(self._h, slot, out.ctypes.data_as(_P)))
        return out

    def step(self, nsteps: int = 1, graph: bool = True):
        _check(self._lib.sf_batch_step(self._h, int(nsteps), 1 if graph else 0))

    def clock(self, slot: int):
        t = ctypes.c_double(0.0)
        k = ctypes.c_longlong(0)
        _check(self._lib.sf_batch_clock(self._h, slot, ctypes.byref(t), ctypes.byref(k)))
        return t.value, k.value

    def steps(self) -> np.ndarray:
        out = np.empty(self.item_capacity, dtype=np.int64)
        _check(self._lib.sf_batch_steps(self._h, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_longlong))))
        return out

    def memory(self):
        used = ctypes.c_size_t(0)
        free = ctypes.c_size_t(0)
        n = ctypes.c_int(0)
        _check(self._lib.sf_batch_memory(self._h, ctypes.byref(used), ctypes.byref(free),
                                         ctypes.byref(n)))
        return used.value, free.value, n.value

    def set_tiles(self, xlen: int, ylen: int):
        _check(self._lib.sf_batch_set_tiles(self._h, int(xlen), int(ylen)))

    def set_occupancy(self, blocks_per_cu: int):
        _check(self._lib.sf_batch_set_occupancy(self._h, int(blocks_per_cu)))

    def set_schedule(self, fused: bool):
        _check(self._lib.sf_batch_set_schedule(self._h, 1 if fused else 0))

    def faults(self) -> np.ndarray:
        """Per-slot device fault flags (0 for healthy or free slots)."""
        out = np.zeros(self.item_capacity, dtype=np.uint32)
        _check(self._lib.sf_batch_faults(self._h, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint32))))
        return out

    def autotune(self, trial_steps: int = 3) -> dict:
        """Time every launch shape on the admitted problems and keep the
        fastest; the problems' states and clocks are left unchanged."""
        t = Tuning()
        _check(self._lib.sf_batch_autotune(self._h, int(trial_steps), ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in t._fields_}
