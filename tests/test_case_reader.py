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
     "convection:\n  default: TENO5\n  flux: upwindAdvection", "upwindAdvection"),
    ("solvers/numerics.yaml", "weno5Rusanov", "primitiveUpwind1", "not served"),
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
def test_reference_sod_case_is_read():
    """The reference's own Sod case (TENO5 + Steger-Warming + SSPRK3 in the
    pair form) lowers onto the grid: its density, velocity and pressure
    name no condition for the Left set, so that end carries none."""
    with Case(REFERENCE_SOD) as c:
        d = c.grid()
        assert list(d.n) == [101, 21, 1] and list(d.active) == [1, 1, 0]
        assert list(d.bc) == [2, 0, 0]
        assert (d.scheme, d.flux, d.time) == (1, 2, 1)
        assert c.numerics()[:3] == ("teno5Steger", "SSPRK3", 0.5)


def test_recipes_beyond_weno5_and_ends_without_conditions(tmp_path):
    case = _copy(tmp_path)
    # WENO7 needs four ghost layers; with nGhost 4 every recipe id resolves
    _edit(case / "mesh/generation.yaml", '"nGhost": 3', '"nGhost": 4')
    for recipe, scheme, flux in (("weno7Rusanov", 3, 0), ("weno3Steger", 2, 2),
                                 ("teno5LaxFriedrichs", 1, 3), ("weno7Roe", 3, 4),
                                 ("weno3HLLC", 2, 1)):
        text = (case / "solvers/numerics.yaml").read_text()
        import re
        (case / "solvers/numerics.yaml").write_text(
            re.sub(r"convection: \w+", "convection: " + recipe, text))
        with Case(case) as c:
            d = c.grid()
            assert (d.scheme, d.flux) == (scheme, flux), recipe
            assert c.numerics()[0] == recipe
    # the pair form without a flux selects Steger-Warming, as the reference does
    text = (case / "solvers/numerics.yaml").read_text()
    (case / "solvers/numerics.yaml").write_text(
        re.sub(r"terms:\n  convection: \w+", "convection:\n  default: WENO5", text))
    with Case(case) as c:
        assert c.numerics()[0] == "weno5Steger"
    # no Left condition for any of rho, U, p: the lower x end carries none
    bnd = case / "fields/boundaries.yaml"
    text = bnd.read_text()
    text = text.replace("U:\n  Left: zeroGradient\n", "U:\n").replace(
        "p:\n  Left: zeroGradient\n", "p:\n").replace("rho:\n  Left: zeroGradient\n", "rho:\n")
    bnd.write_text(text)
    with Case(case) as c:
        assert list(c.grid().bc) == [2, 0, 0]
        assert "no rho/U/p condition" in c.describe()
