"""Byte parity of the structured-grid output with the reference's own
writer: the reference's ResultWriter (infrastructure/io/SF_vtk*.cpp,
compiled into oracle/_ref/libsfvts.so with its block-point generator, Field
and perfect-gas state model) and the library's (host/sf_vts.cpp, through
sf_case_write_vts) write the same states of the reference's own case trees
under tests/cases; every .vts file and the .pvd index must be identical
byte for byte.  The states are each case's initial state and a seeded
perturbation of it (smooth, positive density and pressure), so the point
arrays carry full 17-digit values and the closure's temperature, sound
speed and velocity are compared bit for bit as the writer prints them.

What the harness restates rather than runs: the list of point arrays the
single-fluid application hands the writer (names and order), cited in
oracle/vts_harness.cpp.  Needs the reference tree (the harness is built
from it); skipped where it was not built."""
from pathlib import Path

import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd.case import Case

CASES = Path(__file__).resolve().parent / "cases"

pytestmark = pytest.mark.skipif(not ob.vts_available(),
                                reason="oracle/_ref/libsfvts.so not built (needs the reference tree)")


def _perturbed(q, seed):
    """A smooth, positive perturbation of a state (5, N): density and
    pressure scaled by 1 + 0.2 sin, momenta shifted."""
    rng = np.random.default_rng(seed)
    n = q.shape[1]
    x = np.linspace(0.0, 1.0, n)
    out = q.copy()
    s = 1.0 + 0.2 * np.sin(2 * np.pi * (x * rng.uniform(1, 3) + rng.uniform()))
    rho = q[0] * s
    u = q[1:4] / q[0] + 0.1 * np.cos(2 * np.pi * x * rng.uniform(1, 4))[None, :] * \
        rng.uniform(-1, 1, 3)[:, None]
    gamma = 1.4
    p = (gamma - 1.0) * (q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0]) * s[::-1]
    out[0] = rho
    out[1:4] = rho * u
    out[4] = p / (gamma - 1.0) + 0.5 * rho * (u ** 2).sum(axis=0)
    return out


@pytest.mark.parametrize("case", ["sod2d", "sod_teno5_steger", "sod_weno7_t0p2"])
def test_vts_bytes_match_the_reference_writer(case, tmp_path):
    with Case(CASES / case) as c:
        d = c.grid()
        q0 = c.state()
        rc = c.run_control()
        nodes, vertex, gas_constant = c.mesh()
        field_sets = {f: c.boundary_sets(f) for f in ("rho", "U", "p")}
        names = sorted({s for v in field_sets.values() for s in v})
        sets = {s: c.set_mask(s) for s in names}
        states = [q0, _perturbed(q0, 11), _perturbed(q0, 12)]
        steps = [0, 7, 140]
        times = [0.0, 0.0123456789, 0.025]
        ours, theirs = tmp_path / "ours", tmp_path / "reference"
        ours.mkdir()
        theirs.mkdir()
        c.write_vts(str(ours), states, steps, times)
    job = sorted(p.stem for p in ours.glob("*.pvd"))
    assert len(job) == 1
    ob.ref_write_vts(nodes, vertex, list(d.active), d.gamma, gas_constant, states, steps, times,
                     sets, field_sets, str(theirs), job[0], by_step=bool(rc["write_by_step"]))
    mine = sorted(p.name for p in ours.iterdir())
    ref = sorted(p.name for p in theirs.iterdir())
    assert mine == ref and len(mine) == 4, (mine, ref)
    for name in mine:
        a, b = (ours / name).read_bytes(), (theirs / name).read_bytes()
        if a != b:
            la, lb = a.split(b"\n"), b.split(b"\n")
            k = next(i for i, (x, y) in enumerate(zip(la, lb)) if x != y) if la != lb else -1
            pytest.fail(f"{case}/{name} differs from the reference writer at line {k}: "
                        f"{la[k][:200]!r} vs {lb[k][:200]!r}")
