"""SonicFile case reader (host side, no GPU): the YAML subset, the lowering
of a case tree onto the grid description / recipes / run control / initial
state, and the rejection of what this path does not serve."""
import shutil
from pathlib import Path

import numpy as np
import pytest

from sonicsolver_amd import SonicError
from sonicsolver_amd.case import Case

ROOT = Path(__file__).resolve().parent.parent
SOD2D = ROOT / "tests" / "cases" / "sod2d"
REFERENCE_SOD = Path("/root/reference/test/Sod/sodCase")


def test_sod2d_lowering():
    with Case(SOD2D) as c:
        d = c.grid()
        assert list(d.n) == [101, 9, 1]
        assert list(d.active) == [1, 1, 0]
        assert list(d.bc) == [0, 0, 0]
        assert d.d[0] == pytest.approx(0.01, rel=1e-15) and d.d[1] == pytest.approx(0.01, rel=1e-15)
        assert (d.scheme, d.flux, d.time) == (0, 0, 1)
        assert d.gamma == 1.4
        conv, time, cfl, max_dt = c.numerics()
        assert (conv, time, cfl) == ("weno5Rusanov", "SSPRK3", 0.5)
        assert max_dt == np.finfo(float).max
        rc = c.run_control()
        assert rc["end_time"] == 0.05 and rc["write_interval_time"] == 0.025
        assert not rc["write_by_step"]
        q = c.state().reshape(5, 9, 101)
        # leftHalf = nodes i < 101 // 2 = 50
        assert (q[0, :, :50] == 1.0).all() and (q[0, :, 50:] == 0.125).all()
        assert (q[1:4] == 0.0).all()
        # E = rho e + kinetic with e = p / ((gamma - 1) rho)
        g1 = 1.4 - 1.0
        assert q[4, 0, 0] == 1.0 * (0.0 + 1.0 / (g1 * 1.0))
        assert q[4, 0, 100] == 0.125 * (0.0 + 0.1 / (g1 * 0.125))
        assert "field p: all=0.10000000000000001 leftHalf=1" in c.describe()


def _copy(tmp_path):
    dst = tmp_path / "case"
    shutil.copytree(SOD2D, dst)
    return dst


def _edit(path, old, new):
    text = path.read_text()
    assert old in text
    path.write_text(text.replace(old, new))


@pytest.mark.parametrize("file,old,new,message", [
    ("solvers/numerics.yaml", "weno5Rusanov", "weno7Rusanov", "weno7"),
    ("solvers/numerics.yaml", "terms:\n  convection: weno5Rusanov",
     "convection:\n  default: TENO5\n  flux: StegerWarming", "StegerWarming"),
    ("fields/boundaries.yaml", "rho:\n  Left: zeroGradient\n", "rho:\n", "lower end of axis x"),
    ("fields/boundaries.yaml", "p:\n  Left: zeroGradient", "p:\n  Left: fixedValue",
     "fixedValue"),
    ("solvers/runtime.yaml", "writeControl: runTime", "writeControl: adjustable",
     "unsupported writeControl"),
    ("solvers/algorithm.yaml", "densityBase", "pressureBase", "densityBase"),
    ("models/thermoDynamics.yaml", "perfectGas", "rhoConst", "perfectGas"),
    ("mesh/generation.yaml", "[1.0, 0.08, 0.0]", "[1.1, 0.08, 0.0]", "axis-aligned box"),
    ("fields/internalField.yaml", "leftHalf: 1.0     #", "bubble: 1.0     #", "bubble"),
    ("case.yaml", "formatVersion: 1", "formatVersion: 2", "formatVersion"),
    ("solvers/solvers.yaml", "file: solvers/runtime.yaml", "file: fields/runtime.yaml",
     "must stay in solvers/"),
])
def test_rejections_name_the_problem(tmp_path, file, old, new, message):
    case = _copy(tmp_path)
    _edit(case / file, old, new)
    with pytest.raises(SonicError) as e:
        Case(case)
    assert message in str(e.value)
    assert e.value.code == -1


def test_timestep_output_and_pair_form(tmp_path):
    case = _copy(tmp_path)
    _edit(case / "solvers/runtime.yaml", "writeControl: runTime\nwriteInterval: 0.025",
          "writeControl: timeStep\nwriteInterval: 7")
    _edit(case / "solvers/numerics.yaml", "terms:\n  convection: weno5Rusanov",
          "convection:\n  default: TENO5\n  flux: HLLC\n  reconstruction: characteristic")
    with Case(case) as c:
        rc = c.run_control()
        assert rc["write_by_step"] == 1 and rc["write_interval_steps"] == 7
        conv, time, _, _ = c.numerics()
        assert conv == "teno5HLLC" and time == "SSPRK3"
        d = c.grid()
        assert (d.scheme, d.flux) == (1, 1)


def test_three_dimensional_and_one_dimensional(tmp_path):
    case = _copy(tmp_path)
    _edit(case / "mesh/generation.yaml", '"cells": [100, 8, 1]', '"cells": [20, 8, 6]')
    text = (case / "fields/boundaries.yaml").read_text().replace("frontBack: empty",
                                                                  "frontBack: zeroGradient")
    (case / "fields/boundaries.yaml").write_text(text)
    with Case(case) as c:
        d = c.grid()
        assert list(d.n) == [21, 9, 7] and list(d.active) == [1, 1, 1]
    text = text.replace("sideWalls: zeroGradient", "sideWalls: empty").replace(
        "sideWalls: {type: zeroGradient}", "sideWalls: empty").replace(
        "frontBack: zeroGradient", "frontBack: empty")
    (case / "fields/boundaries.yaml").write_text(text)
    with Case(case) as c:
        d = c.grid()
        assert list(d.n) == [21, 1, 1] and list(d.active) == [1, 0, 0]


@pytest.mark.skipif(not REFERENCE_SOD.exists(), reason="reference case tree not present")
def test_reference_sod_case_is_read_and_rejected_with_reason():
    """The reference's own Sod case parses (every registry and document),
    and is rejected for what this path does not serve: its density,
    velocity and pressure have no condition on the Left set."""
    with pytest.raises(SonicError) as e:
        Case(REFERENCE_SOD)
    assert "lower end of axis x" in str(e.value)
