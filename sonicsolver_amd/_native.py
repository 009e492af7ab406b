"""ctypes binding of libsonic_amd.so (the C ABI in include/sonic_amd.h).

The shared library is built in-tree (``__graft_entry__.build()`` or
``python -m sonicsolver_amd.build``).  Loading fails loudly when it is
missing: there is no CPU fallback for any operation.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libsonic_amd.so"

SCHEMES = {"weno5": 0, "teno5": 1, "weno3": 2, "weno7": 3}
FLUXES = {"rusanov": 0, "hllc": 1, "steger": 2, "stegerwarming": 2, "laxfriedrichs": 3, "roe": 4,
          "laxwendroff": 5}
TIMES = {"forwardeuler": 0, "euler": 0, "ssprk3": 1, "classicalrk4": 2, "rk4": 2}
# per axis; "none_lo" / "none_hi" / "none": the end carries no condition
# (its node is advanced, its ghosts keep the uploaded end-node state)
BCS = {"transmissive": 0, "zerogradient": 0, "periodic": 1, "none_lo": 2, "none_hi": 3,
       "none": 4}

FLAG_NAMES = {1: "state", 2: "face-average", 4: "teno-cutoff", 8: "flux", 16: "dt"}

KERNELS = {"sweep_x": 0, "sweep_y": 1, "sweep_z": 2, "ghost": 3, "cfl": 4}


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("active", ctypes.c_int * 3),
        ("bc", ctypes.c_int * 3),
        ("d", ctypes.c_double * 3),
        ("gamma", ctypes.c_double),
        ("scheme", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("time", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class SonicError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None

_P = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


def _declare(lib):
    sig = {
        "sf_last_error": [ctypes.c_char_p, ctypes.c_size_t],
        "sf_version": [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "sf_grid_create": [ctypes.POINTER(GridDesc), ctypes.POINTER(_VP)],
        "sf_grid_destroy": [_VP],
        "sf_grid_device_bytes": [_VP, ctypes.POINTER(ctypes.c_size_t)],
        "sf_grid_stream": [_VP, ctypes.POINTER(_VP)],
        "sf_state_upload": [_VP, _P],
        "sf_state_download": [_VP, _P],
        "sf_state_upload_device": [_VP, _VP],
        "sf_state_download_device": [_VP, _VP],
        "sf_step_prepare": [_VP],
        "sf_compute_dt": [_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P],
        "sf_set_dt": [_VP, ctypes.c_double],
        "sf_step_begin": [_VP],
        "sf_stage": [_VP, ctypes.c_int],
        "sf_step_commit": [_VP],
        "sf_stage_count": [_VP, ctypes.POINTER(ctypes.c_int)],
        "sf_run": [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P],
        "sf_clock": [_VP, _P, ctypes.POINTER(ctypes.c_longlong)],
        "sf_residual": [_VP, _P],
        "sf_error_flags": [_VP, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int],
        "sf_fault_check": [_VP, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_longlong),
                           ctypes.POINTER(ctypes.c_int)],
        "sf_grid_set_march_segment": [_VP, ctypes.c_int],
        "sf_grid_set_x_walk": [_VP, ctypes.c_int],
        "sf_run_static": [ctypes.POINTER(GridDesc), _VP, ctypes.c_int, ctypes.c_double,
                          ctypes.c_double, _VP, _VP, ctypes.POINTER(ctypes.c_size_t)],
        "sf_riemann_flux": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            ctypes.c_double, ctypes.c_int, _P, _P, _P,
                            ctypes.POINTER(ctypes.c_uint32)],
        "sf_face_flux": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                         ctypes.c_double, ctypes.c_int, _P, _P, ctypes.POINTER(ctypes.c_uint32)],
        "sf_synchronize": [_VP],
        "sf_profile_enable": [_VP, ctypes.c_int],
        "sf_profile_read": [_VP, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_longlong)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def load():
    """Load libsonic_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m sonicsolver_amd.build` "
                "(there is no CPU fallback)")
        _lib = _declare(ctypes.CDLL(str(LIB_PATH)))
        _declare_host(_lib)
    return _lib


def _check(rc: int):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        load().sf_last_error(buf, len(buf))
        raise SonicError(rc, buf.value.decode(errors="replace"))


def make_desc(n, active=None, bc=None, d=(1.0, 1.0, 1.0), gamma=1.4, scheme="weno5",
              flux="rusanov", time="ssprk3", device=0) -> GridDesc:
    n = list(n) + [1] * (3 - len(n))
    if active is None:
        active = [1 if x > 1 else 0 for x in n]
    bc = bc or ["transmissive"] * 3
    desc = GridDesc()
    desc.n[:] = [int(x) for x in n]
    desc.active[:] = [int(x) for x in active]
    desc.bc[:] = [BCS[b] if isinstance(b, str) else int(b) for b in bc]
    desc.d[:] = [float(x) for x in d]
    desc.gamma = float(gamma)
    desc.scheme = SCHEMES[scheme.lower()] if isinstance(scheme, str) else int(scheme)
    desc.flux = FLUXES[flux.lower()] if isinstance(flux, str) else int(flux)
    desc.time = TIMES[time.lower()] if isinstance(time, str) else int(time)
    desc.device = int(device)
    return desc


class Grid:
    """A device-resident structured grid (one sf_grid handle)."""

    def __init__(self, desc: GridDesc):
        lib = load()
        self.desc = desc
        self.n = tuple(desc.n)
        self.N = int(np.prod(self.n))
        h = _VP()
        _check(lib.sf_grid_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().sf_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def upload(self, soa: np.ndarray):
        a = np.ascontiguousarray(soa, dtype=np.float64).reshape(5, self.N)
        _check(load().sf_state_upload(self._h, a.ctypes.data_as(_P)))

    def download(self) -> np.ndarray:
        out = np.empty((5, self.N), dtype=np.float64)
        _check(load().sf_state_download(self._h, out.ctypes.data_as(_P)))
        return out

    def upload_device(self, ptr: int):
        _check(load().sf_state_upload_device(self._h, _VP(ptr)))

    def download_device(self, ptr: int):
        _check(load().sf_state_download_device(self._h, _VP(ptr)))

    def run(self, nsteps: int, cfl: float, max_dt: float = 1.0e300, record_dt: bool = True):
        dts = np.zeros(max(nsteps, 1), dtype=np.float64)
        _check(load().sf_run(self._h, int(nsteps), float(cfl), float(max_dt),
                             dts.ctypes.data_as(_P) if record_dt else None))
        return dts[:nsteps]

    def residual(self) -> np.ndarray:
        out = np.empty((5, self.N), dtype=np.float64)
        _check(load().sf_residual(self._h, out.ctypes.data_as(_P)))
        return out

    def prepare(self):
        _check(load().sf_step_prepare(self._h))

    def compute_dt(self, cfl: float, max_dt: float = 1.0e300, limit: float = 1.0e300) -> float:
        dt = ctypes.c_double(0.0)
        _check(load().sf_compute_dt(self._h, cfl, max_dt, limit, ctypes.byref(dt)))
        return dt.value

    def set_dt(self, dt: float):
        _check(load().sf_set_dt(self._h, float(dt)))

    def begin(self):
        _check(load().sf_step_begin(self._h))

    def stage(self, s: int):
        _check(load().sf_stage(self._h, int(s)))

    def commit(self):
        _check(load().sf_step_commit(self._h))

    def stage_count(self) -> int:
        c = ctypes.c_int(0)
        _check(load().sf_stage_count(self._h, ctypes.byref(c)))
        return c.value

    def clock(self):
        t = ctypes.c_double(0.0)
        s = ctypes.c_longlong(0)
        _check(load().sf_clock(self._h, ctypes.byref(t), ctypes.byref(s)))
        return t.value, s.value

    def flags(self, clear: bool = False) -> int:
        f = ctypes.c_uint32(0)
        _check(load().sf_error_flags(self._h, ctypes.byref(f), int(clear)))
        return f.value

    def fault(self):
        """(flags, faulting step on the grid's clock, phase) after waiting for
        the queued work; (0, 0, -1) without a fault.  With a fault the grid
        is left holding the state at the start of the faulting step."""
        f = ctypes.c_uint32(0)
        st = ctypes.c_longlong(0)
        ph = ctypes.c_int(-1)
        rc = load().sf_fault_check(self._h, ctypes.byref(f), ctypes.byref(st), ctypes.byref(ph))
        if rc not in (0, -3):
            _check(rc)
        return f.value, st.value, ph.value

    def set_march_segment(self, nodes: int):
        _check(load().sf_grid_set_march_segment(self._h, int(nodes)))

    def set_x_walk(self, mode: int):
        """x sweep by whole lines (1), by flattened face tiles (0) or the
        automatic choice (-1)."""
        _check(load().sf_grid_set_x_walk(self._h, int(mode)))

    def profile(self, enable: bool = True):
        _check(load().sf_profile_enable(self._h, int(enable)))

    def profile_read(self, kernel: int):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_longlong(0)
        _check(load().sf_profile_read(self._h, int(kernel), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        _check(load().sf_synchronize(self._h))

    def stream(self) -> int:
        s = _VP()
        _check(load().sf_grid_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def device_bytes(self) -> int:
        b = ctypes.c_size_t(0)
        _check(load().sf_grid_device_bytes(self._h, ctypes.byref(b)))
        return b.value


class RunControl(ctypes.Structure):
    _fields_ = [("start_time", ctypes.c_double), ("end_time", ctypes.c_double),
                ("end_step", ctypes.c_int), ("write_by_step", ctypes.c_int),
                ("write_interval_steps", ctypes.c_int), ("write_interval_time", ctypes.c_double)]


OUTPUT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                             ctypes.c_double)


def _declare_host(lib):
    sig = {
        "sf_recipe_resolve": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "sf_recipe_from_pair": [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "sf_time_recipe_resolve": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)],
        "sf_plan_describe": [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t],
        "sf_solver_create": [ctypes.POINTER(GridDesc), ctypes.c_char_p, ctypes.c_char_p,
                             ctypes.c_double, ctypes.c_double, ctypes.POINTER(_VP)],
        "sf_solver_destroy": [_VP],
        "sf_solver_grid": [_VP, ctypes.POINTER(_VP)],
        "sf_solver_run": [_VP, ctypes.POINTER(RunControl), OUTPUT_FN, _VP,
                          ctypes.POINTER(ctypes.c_int), _P],
        "sf_solver_dt_history": [_VP, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int


def resolve_recipe(name: str):
    """(scheme, flux, halo width) of a convection term recipe name."""
    s, f, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(load().sf_recipe_resolve(name.encode(), ctypes.byref(s), ctypes.byref(f),
                                    ctypes.byref(h)))
    return s.value, f.value, h.value


def recipe_from_pair(scheme: str, flux: str):
    s, f, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(load().sf_recipe_from_pair(scheme.encode(), flux.encode(), ctypes.byref(s),
                                      ctypes.byref(f), ctypes.byref(h)))
    return s.value, f.value, h.value


def resolve_time_recipe(token: str):
    t, n = ctypes.c_int(), ctypes.c_int()
    _check(load().sf_time_recipe_resolve(token.encode(), ctypes.byref(t), ctypes.byref(n)))
    return t.value, n.value


def describe_plan(convection: str, time: str) -> str:
    buf = ctypes.create_string_buffer(4096)
    _check(load().sf_plan_describe(convection.encode(), time.encode(), buf, len(buf)))
    return buf.value.decode()


class Solver:
    """Provider-level solver: recipe names, the compiled explicit plan and the
    time-driver loop over one device grid."""

    def __init__(self, geometry: GridDesc, convection: str, time: str, cfl: float,
                 max_dt: float = 1.0e300):
        lib = load()
        h = _VP()
        _check(lib.sf_solver_create(ctypes.byref(geometry), convection.encode(), time.encode(),
                                    float(cfl), float(max_dt), ctypes.byref(h)))
        self._h = h
        g = _VP()
        _check(lib.sf_solver_grid(h, ctypes.byref(g)))
        self._grid = g
        self.N = int(np.prod(list(geometry.n)))

    def close(self):
        if getattr(self, "_h", None):
            load().sf_solver_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, soa: np.ndarray):
        a = np.ascontiguousarray(soa, dtype=np.float64).reshape(5, self.N)
        _check(load().sf_state_upload(self._grid, a.ctypes.data_as(_P)))

    def download(self) -> np.ndarray:
        out = np.empty((5, self.N), dtype=np.float64)
        _check(load().sf_state_download(self._grid, out.ctypes.data_as(_P)))
        return out

    def run(self, end_time: float, start_time: float = 0.0, end_step: int = 0,
            write_interval_time: float = 1e300, write_interval_steps: int = 0,
            write_by_step: bool = False, on_output=None):
        rc = RunControl(start_time, end_time, end_step, int(write_by_step),
                        write_interval_steps, write_interval_time)
        outputs = []

        def cb(_user, step, t, dt):
            outputs.append((step, t, dt))
            if on_output is not None:
                return 0 if on_output(step, t, dt) in (None, True) else 1
            return 0

        fn = OUTPUT_FN(cb)
        steps = ctypes.c_int(0)
        t = ctypes.c_double(0.0)
        _check(load().sf_solver_run(self._h, ctypes.byref(rc), fn, None, ctypes.byref(steps),
                                    ctypes.byref(t)))
        return steps.value, t.value, outputs

    def dt_history(self) -> np.ndarray:
        n = ctypes.c_int(0)
        _check(load().sf_solver_dt_history(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value)
        _check(load().sf_solver_dt_history(self._h, out.ctypes.data_as(_P), n.value,
                                           ctypes.byref(n)))
        return out


def run_static(desc: GridDesc, soa_device: int, nsteps: int, cfl: float, max_dt: float = 1.0e300,
               dt_seq_device: int = 0, stream: int = 0) -> int:
    """sf_run_static on device pointers (integers); returns the scratch bytes
    the grid needs.  Raises SonicError on failure (SF_E_NUMERIC after a
    fault, with the state at the start of the faulting step in place)."""
    need = ctypes.c_size_t(0)
    _check(load().sf_run_static(ctypes.byref(desc), ctypes.c_void_p(soa_device), int(nsteps),
                                float(cfl), float(max_dt), ctypes.c_void_p(dt_seq_device or None),
                                ctypes.c_void_p(stream or None), ctypes.byref(need)))
    return need.value


def riemann_flux(flux, axis: int, qL, qR, gamma: float = 1.4, lw: float = 0.0, device: int = 0):
    """Device face fluxes of state pairs (count x 5 arrays); (flux, flags)."""
    a = np.ascontiguousarray(qL, dtype=np.float64).reshape(-1, 5)
    b = np.ascontiguousarray(qR, dtype=np.float64).reshape(-1, 5)
    out = np.empty_like(a)
    flags = np.zeros(len(a), dtype=np.uint32)
    f = FLUXES[flux.lower()] if isinstance(flux, str) else int(flux)
    _check(load().sf_riemann_flux(device, f, int(axis), float(gamma), float(lw), len(a),
                                  a.ctypes.data_as(_P), b.ctypes.data_as(_P),
                                  out.ctypes.data_as(_P),
                                  flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return out, flags


def face_flux(scheme, flux, axis: int, stencils, area: float = 1.0, gamma: float = 1.4,
              device: int = 0):
    """Device characteristic WENO face fluxes of six-node stencils
    (count x 6 x 5); (flux, flags)."""
    st = np.ascontiguousarray(stencils, dtype=np.float64).reshape(-1, 6, 5)
    out = np.empty((len(st), 5))
    flags = np.zeros(len(st), dtype=np.uint32)
    sc = SCHEMES[scheme.lower()] if isinstance(scheme, str) else int(scheme)
    f = FLUXES[flux.lower()] if isinstance(flux, str) else int(flux)
    _check(load().sf_face_flux(device, sc, f, int(axis), float(gamma), float(area), len(st),
                               st.ctypes.data_as(_P), out.ctypes.data_as(_P),
                               flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return out, flags


def describe_flags(flags: int) -> str:
    names = [v for k, v in FLAG_NAMES.items() if flags & k]
    return ",".join(names) if names else "none"


def env_device() -> int:
    return int(os.environ.get("SONIC_AMD_DEVICE", "0"))
