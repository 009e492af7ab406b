"""MI355X-native explicit Euler hot path of sonicSolver (FP64, gfx950).

The compute path is libsonic_amd.so (hand-written HIP kernels behind the C
ABI in include/sonic_amd.h).  This Python package is a thin ctypes layer for
tests and benchmarks; it has no CPU fallback.
"""
from ._native import (Grid, GridDesc, Solver, SonicError, load, make_desc, describe_flags,
                      describe_plan, resolve_recipe, recipe_from_pair, resolve_time_recipe,
                      SCHEMES, FLUXES, TIMES, BCS)

__all__ = ["Grid", "GridDesc", "Solver", "SonicError", "load", "make_desc", "describe_flags",
           "describe_plan", "resolve_recipe", "recipe_from_pair", "resolve_time_recipe",
           "SCHEMES", "FLUXES", "TIMES", "BCS"]
