"""ctypes binding of the checker libraries.  TEST INFRASTRUCTURE ONLY.

  liboracle.so      C restatement of the hot path (oracle/sfcpu.c)
  _ref/libsfref.so  harness over the reference's own translation units
                    (oracle/ref_harness.cpp); present where it was built
  _ref/libsfvts.so  the reference's structured-grid writer over a reference
                    Field (oracle/vts_harness.cpp); likewise

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libsfref.so"
REF_FMA_SO = HERE / "_ref" / "libsfref_fma.so"
VTS_SO = HERE / "_ref" / "libsfvts.so"

_P = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint)

SCHEMES = {"weno5": 0, "teno5": 1, "weno3": 2, "weno7": 3}
FLUXES = {"rusanov": 0, "hllc": 1, "steger": 2, "laxfriedrichs": 3, "roe": 4,
          "laxwendroff": 5}
TIMES = {"euler": 0, "forwardeuler": 0, "ssprk3": 1, "rk4": 2, "classicalrk4": 2}
# per axis; "none_lo" / "none_hi" / "none": that end carries no condition
BCS = {"transmissive": 0, "periodic": 1, "none_lo": 2, "none_hi": 3, "none": 4}
STENCIL = {0: 6, 1: 6, 2: 4, 3: 8}


class Desc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("active", ctypes.c_int * 3),
                ("bc", ctypes.c_int * 3), ("d", ctypes.c_double * 3),
                ("gamma", ctypes.c_double), ("scheme", ctypes.c_int),
                ("flux", ctypes.c_int), ("time", ctypes.c_int)]


def make_desc(n, active=None, bc=None, d=(1.0, 1.0, 1.0), gamma=1.4, scheme="weno5",
              flux="rusanov", time="ssprk3") -> Desc:
    n = list(n) + [1] * (3 - len(n))
    active = active if active is not None else [1 if x > 1 else 0 for x in n]
    bc = bc or ["transmissive"] * 3
    ds = Desc()
    ds.n[:] = [int(x) for x in n]
    ds.active[:] = [int(x) for x in active]
    ds.bc[:] = [BCS[b] if isinstance(b, str) else int(b) for b in bc]
    ds.d[:] = [float(x) for x in d]
    ds.gamma = float(gamma)
    ds.scheme = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
    ds.flux = FLUXES[flux] if isinstance(flux, str) else int(flux)
    ds.time = TIMES[time] if isinstance(time, str) else int(time)
    return ds


_oracle = None
_ref = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)


def oracle():
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            build_oracle()
        lib = ctypes.CDLL(str(ORACLE_SO))
        lib.sfo_grid_create.restype = ctypes.c_void_p
        lib.sfo_grid_create.argtypes = [ctypes.POINTER(Desc)]
        for name in ("sfo_grid_destroy",):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
            getattr(lib, name).restype = None
        for name in ("sfo_set_state", "sfo_get_state", "sfo_residual"):
            getattr(lib, name).argtypes = [ctypes.c_void_p, _P]
        lib.sfo_apply_bc.argtypes = [ctypes.c_void_p]
        lib.sfo_compute_dt.argtypes = [ctypes.c_void_p, ctypes.c_double, _P]
        lib.sfo_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.sfo_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, _P]
        lib.sfo_error.argtypes = [ctypes.c_void_p]
        lib.sfo_error.restype = ctypes.c_uint
        lib.sfo_weno5_core.argtypes = [_P, _U]
        lib.sfo_weno5_core.restype = ctypes.c_double
        lib.sfo_teno5_core.argtypes = [_P, _U]
        lib.sfo_teno5_core.restype = ctypes.c_double
        for name in ("sfo_weno3_core", "sfo_weno7_core"):
            getattr(lib, name).argtypes = [_P, _U]
            getattr(lib, name).restype = ctypes.c_double
        lib.sfo_weno7_smoothness.argtypes = [_P, ctypes.c_double, ctypes.c_double]
        lib.sfo_weno7_smoothness.restype = ctypes.c_double
        for name in ("sfo_steger_warming", "sfo_lax_friedrichs", "sfo_roe"):
            getattr(lib, name).argtypes = [_P, _P, _P, ctypes.c_double, _P]
        lib.sfo_face_flux.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_double, _P]
        lib.sfo_char_face_flux_n.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, _P]
        lib.sfo_char_face_flux_dt.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              _P]
        lib.sfo_lax_wendroff.argtypes = [_P, _P, _P, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _P]
        lib.sfo_char_matrix.argtypes = [_P, _P, ctypes.c_double, _P, _P]
        for name in ("sfo_rusanov", "sfo_hllc"):
            getattr(lib, name).argtypes = [_P, _P, _P, ctypes.c_double, _P]
        lib.sfo_physical_flux.argtypes = [_P, _P, ctypes.c_double, _P]
        lib.sfo_char_face_flux.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _P]
        lib.sfo_gemv5.argtypes = [_P, _P, _P]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return REF_SO.exists()


def ref_fma_available() -> bool:
    return REF_FMA_SO.exists()


def vts_available() -> bool:
    return VTS_SO.exists()


def ref_write_vts(nodes, vertex, active, gamma, gas_constant, states, steps, times,
                  sets: dict, field_sets: dict, directory: str, job: str, by_step: bool):
    """The reference's ResultWriter over a reference Field
    (oracle/vts_harness.cpp): states is a list of (5, N) states on the
    unique nodes of the active axes, sets maps a node-set name to its mask
    over the block nodes (i fastest), field_sets maps rho / U / p to the set
    names of their boundary settings, in order."""
    lib = ctypes.CDLL(str(VTS_SO))
    fn = lib.sfref_write_vts
    fn.restype = ctypes.c_int
    names = list(sets)
    masks = np.ascontiguousarray(np.stack([np.asarray(sets[n], dtype=np.uint8) for n in names])
                                 if names else np.zeros(1, dtype=np.uint8))
    arr = np.ascontiguousarray(np.stack([np.asarray(s, dtype=np.float64).reshape(5, -1)
                                         for s in states]))
    err = ctypes.create_string_buffer(1024)
    rc = fn((ctypes.c_int * 3)(*nodes), _ptr(np.ascontiguousarray(vertex, dtype=np.float64)),
            (ctypes.c_int * 3)(*active), ctypes.c_double(gamma), ctypes.c_double(gas_constant),
            ctypes.c_int(len(steps)), (ctypes.c_int * len(steps))(*steps),
            _ptr(np.ascontiguousarray(times, dtype=np.float64)), _ptr(arr),
            ctypes.c_int(len(names)), (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names]),
            masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            ",".join(field_sets.get("rho", [])).encode(), ",".join(field_sets.get("U", [])).encode(),
            ",".join(field_sets.get("p", [])).encode(), str(directory).encode(), job.encode(),
            ctypes.c_int(1 if by_step else 0), err, ctypes.c_size_t(len(err)))
    if rc != 0:
        raise RuntimeError("reference writer: " + err.value.decode())


_ref_fma = None


def ref(fma: bool = False):
    """The reference harness; fma=True loads the build with FMA contraction
    (the reference's release flags)."""
    global _ref, _ref_fma
    if fma:
        if _ref_fma is None:
            if not REF_FMA_SO.exists():
                raise FileNotFoundError(f"{REF_FMA_SO} has not been built (make -C oracle ref_fma)")
            _ref_fma = _declare_ref(ctypes.CDLL(str(REF_FMA_SO)))
        return _ref_fma
    if _ref is None:
        if not REF_SO.exists():
            raise FileNotFoundError(f"{REF_SO} has not been built (make -C oracle ref)")
        _ref = _declare_ref(ctypes.CDLL(str(REF_SO)))
    return _ref


def _declare_ref(lib):
    lib.sfref_run.argtypes = [ctypes.POINTER(Desc), _P, ctypes.c_int, ctypes.c_double,
                              ctypes.c_double, _P]
    lib.sfref_run_slabs.argtypes = [ctypes.POINTER(Desc), _P, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, _P, _P]
    lib.sfref_residual.argtypes = [ctypes.POINTER(Desc), _P, _P]
    lib.sfref_weno5.argtypes = [_P]
    lib.sfref_weno5.restype = ctypes.c_double
    lib.sfref_teno5.argtypes = [_P]
    lib.sfref_teno5.restype = ctypes.c_double
    lib.sfref_char_matrix.argtypes = [_P, _P, ctypes.c_double, _P, _P]
    lib.sfref_rusanov.argtypes = [_P, _P, _P, ctypes.c_double, _P]
    lib.sfref_char_face_flux.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, _P]
    lib.sfref_char_face_flux_n.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _P]
    for name in ("sfref_weno3", "sfref_weno7"):
        getattr(lib, name).argtypes = [_P]
        getattr(lib, name).restype = ctypes.c_double
    lib.sfref_split_flux.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_double, _P]
    lib.sfref_lax_wendroff.argtypes = [_P, _P, _P, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _P]
    lib.sfref_char_face_flux_dt.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _P]
    lib.sfref_run_driver.argtypes = [ctypes.POINTER(Desc), _P, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, _P,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return lib


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_P)


class OracleError(RuntimeError):
    def __init__(self, flags):
        super().__init__(f"oracle raised error flags 0x{flags:x}")
        self.flags = flags


def run(desc: Desc, q: np.ndarray, nsteps: int, cfl: float, max_dt: float = 1e300):
    """Run the C restatement; returns (state, dt sequence)."""
    lib = oracle()
    g = lib.sfo_grid_create(ctypes.byref(desc))
    if not g:
        raise ValueError("oracle rejected the grid description")
    try:
        a = _arr(q).reshape(5, -1).copy()
        lib.sfo_set_state(g, _ptr(a))
        dts = np.zeros(max(nsteps, 1))
        rc = lib.sfo_run(g, nsteps, cfl, max_dt, _ptr(dts))
        if rc != 0:
            raise OracleError(lib.sfo_error(g))
        lib.sfo_get_state(g, _ptr(a))
        return a, dts[:nsteps]
    finally:
        lib.sfo_grid_destroy(g)


def run_driver(desc: Desc, q: np.ndarray, start: float, end: float, cfl: float,
               max_dt: float = 1e300, end_step: int = 0, write_time: float = 1e300,
               write_steps: int = 0, write_by_step: bool = False, cap: int = 100000):
    """Time::Driver loop of the C restatement; returns (state, dt sequence)."""
    lib = oracle()
    lib.sfo_run_driver.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, _P, ctypes.c_int]
    g = lib.sfo_grid_create(ctypes.byref(desc))
    try:
        a = _arr(q).reshape(5, -1).copy()
        lib.sfo_set_state(g, _ptr(a))
        dts = np.zeros(cap)
        n = lib.sfo_run_driver(g, start, end, end_step, int(write_by_step), write_steps,
                               write_time, cfl, max_dt, _ptr(dts), cap)
        if n < 0:
            raise OracleError(lib.sfo_error(g))
        lib.sfo_get_state(g, _ptr(a))
        return a, dts[:n]
    finally:
        lib.sfo_grid_destroy(g)


def residual(desc: Desc, q: np.ndarray):
    lib = oracle()
    g = lib.sfo_grid_create(ctypes.byref(desc))
    try:
        a = _arr(q).reshape(5, -1).copy()
        lib.sfo_set_state(g, _ptr(a))
        out = np.zeros_like(a)
        if lib.sfo_residual(g, _ptr(out)) != 0:
            raise OracleError(lib.sfo_error(g))
        return out
    finally:
        lib.sfo_grid_destroy(g)


def compute_dt(desc: Desc, q: np.ndarray, cfl: float) -> float:
    lib = oracle()
    g = lib.sfo_grid_create(ctypes.byref(desc))
    try:
        a = _arr(q).reshape(5, -1).copy()
        lib.sfo_set_state(g, _ptr(a))
        dt = ctypes.c_double(0.0)
        if lib.sfo_compute_dt(g, cfl, ctypes.byref(dt)) != 0:
            raise OracleError(lib.sfo_error(g))
        return dt.value
    finally:
        lib.sfo_grid_destroy(g)


def ref_run(desc: Desc, q: np.ndarray, nsteps: int, cfl: float, max_dt: float = 1e300,
            fma: bool = False):
    """Run the reference's own code through the harness (fma=True: the
    build with FMA contraction)."""
    a = _arr(q).reshape(5, -1).copy()
    dts = np.zeros(max(nsteps, 1))
    if ref(fma).sfref_run(ctypes.byref(desc), _ptr(a), nsteps, cfl, max_dt, _ptr(dts)) != 0:
        raise RuntimeError("the reference harness raised an error")
    return a, dts[:nsteps]


def ref_run_slabs(desc: Desc, q: np.ndarray, nsteps: int, cfl: float, nslabs: int,
                  max_dt: float = 1e300):
    """The reference's own step on `nslabs` threads: the grid cut into z
    slabs (z periodic, at least the halo width of nodes each) with the
    halo planes exchanged before every ghost fill and the dt reduced to the
    minimum over slabs.  Returns (state, dt sequence, wall seconds of the
    steps alone)."""
    a = _arr(q).reshape(5, -1).copy()
    dts = np.zeros(max(nsteps, 1))
    sec = np.zeros(1)
    if ref().sfref_run_slabs(ctypes.byref(desc), _ptr(a), nsteps, cfl, max_dt, nslabs,
                             _ptr(dts), _ptr(sec)) != 0:
        raise RuntimeError("the reference harness raised an error")
    return a, dts[:nsteps], float(sec[0])


def ref_run_driver(desc: Desc, q: np.ndarray, start: float, end: float, cfl: float,
                   max_dt: float = 1e300, end_step: int = 0, write_time: float = 1e300,
                   write_steps: int = 0, write_by_step: bool = False, cap: int = 100000,
                   fma: bool = False):
    """Time::Driver over the reference's own step (harness); returns (state,
    dt sequence)."""
    a = _arr(q).reshape(5, -1).copy()
    dts = np.zeros(cap)
    n = ctypes.c_int(0)
    rc = ref(fma).sfref_run_driver(ctypes.byref(desc), _ptr(a), start, end, end_step,
                                   int(write_by_step), write_steps, write_time, cfl, max_dt,
                                   _ptr(dts), cap, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError("the reference harness raised an error")
    return a, dts[:n.value]


def ref_residual(desc: Desc, q: np.ndarray):
    a = _arr(q).reshape(5, -1).copy()
    out = np.zeros_like(a)
    ref().sfref_residual(ctypes.byref(desc), _ptr(a), _ptr(out))
    return out


# ---- point kernels --------------------------------------------------------

def weno5(v):
    e = ctypes.c_uint(0)
    r = oracle().sfo_weno5_core(_ptr(_arr(v)), ctypes.byref(e))
    return r, e.value


def teno5(v):
    e = ctypes.c_uint(0)
    r = oracle().sfo_teno5_core(_ptr(_arr(v)), ctypes.byref(e))
    return r, e.value


def weno3(v):
    e = ctypes.c_uint(0)
    r = oracle().sfo_weno3_core(_ptr(_arr(v)), ctypes.byref(e))
    return r, e.value


def weno7(v):
    e = ctypes.c_uint(0)
    r = oracle().sfo_weno7_core(_ptr(_arr(v)), ctypes.byref(e))
    return r, e.value


def weno7_smoothness(v, a, b):
    return oracle().sfo_weno7_smoothness(_ptr(_arr(v)), float(a), float(b))


def face_flux(flux, qL, qR, n, gamma=1.4):
    """Any of the five face fluxes by name; (rc, flux)."""
    out = np.zeros(5)
    rc = oracle().sfo_face_flux(FLUXES[flux], _ptr(_arr(qL)), _ptr(_arr(qR)), _ptr(_arr(n)),
                                gamma, _ptr(out))
    return rc, out


def char_matrix(q, n, gamma=1.4):
    L = np.zeros(25)
    R = np.zeros(25)
    rc = oracle().sfo_char_matrix(_ptr(_arr(q)), _ptr(_arr(n)), gamma, _ptr(L), _ptr(R))
    return rc, L.reshape(5, 5), R.reshape(5, 5)


def rusanov(qL, qR, n, gamma=1.4):
    out = np.zeros(5)
    rc = oracle().sfo_rusanov(_ptr(_arr(qL)), _ptr(_arr(qR)), _ptr(_arr(n)), gamma, _ptr(out))
    return rc, out


def hllc(qL, qR, n, gamma=1.4):
    out = np.zeros(5)
    rc = oracle().sfo_hllc(_ptr(_arr(qL)), _ptr(_arr(qR)), _ptr(_arr(n)), gamma, _ptr(out))
    return rc, out


def physical_flux(q, n, gamma=1.4):
    out = np.zeros(5)
    rc = oracle().sfo_physical_flux(_ptr(_arr(q)), _ptr(_arr(n)), gamma, _ptr(out))
    return rc, out


def char_face_flux(stencil, cofactor, scheme="weno5", flux="rusanov", gamma=1.4, dt=0.0,
                   spacing=0.0):
    """stencil: (ns, 5) nodes at offsets 1-ns/2 .. ns/2 of the face; dt and
    spacing are read by the Lax-Wendroff flux only."""
    out = np.zeros(5)
    sid = SCHEMES[scheme]
    st = _arr(stencil).reshape(STENCIL[sid] * 5)
    rc = oracle().sfo_char_face_flux_dt(_ptr(st), _ptr(_arr(cofactor)), sid, FLUXES[flux],
                                        gamma, dt, spacing, _ptr(out))
    return rc, out


def lax_wendroff(qL, qR, n, dt, spacing, gamma=1.4):
    out = np.zeros(5)
    rc = oracle().sfo_lax_wendroff(_ptr(_arr(qL)), _ptr(_arr(qR)), _ptr(_arr(n)), gamma, dt,
                                   spacing, _ptr(out))
    return rc, out
