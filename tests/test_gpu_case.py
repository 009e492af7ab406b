"""A SonicFile case run on the device: the case tree lowered by the reader,
advanced by Time::Driver on the HIP path, against the oracle's restatement
of the same driver on the same lowered inputs (relative L1 / Linf <= 1e-11),
and the structured-grid output files it writes."""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd.case import Case
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
SOD2D = ROOT / "tests" / "cases" / "sod2d"


def _array(text, name):
    m = re.search(r'Name="%s"[^>]*>\n(.*?)\n\s*</DataArray>' % re.escape(name), text, re.S)
    assert m, name
    return np.array(m.group(1).split(), dtype=float)


def test_sod2d_case_run_matches_oracle_and_writes_vts(tmp_path):
    print("case sod2d", flush=True)
    with Case(SOD2D) as c:
        d = c.grid()
        q0 = c.state()
        conv, time, cfl, max_dt = c.numerics()
        rc = c.run_control()
        out, steps, t_end = c.run(output_dir=str(tmp_path))
    desc = ob.make_desc(n=list(d.n), active=list(d.active), d=list(d.d), gamma=d.gamma,
                        scheme="weno5", flux="rusanov", time="ssprk3")
    ref, dts = ob.run_driver(desc, q0, rc["start_time"], rc["end_time"], cfl, max_dt=1e300,
                             write_time=rc["write_interval_time"])
    assert steps == len(dts)
    assert abs(t_end - rc["end_time"]) <= 64 * np.finfo(float).eps * rc["end_time"]
    assert_close(out, ref, what="sod2d case")
    # output instants 0, 0.025, 0.05 and the collection file
    files = sorted(p.name for p in tmp_path.glob("*.vts"))
    assert files == ["sod2d_t0.025.vts", "sod2d_t0.05.vts", "sod2d_t0.vts"]
    pvd = (tmp_path / "sod2d.pvd").read_text()
    assert pvd.count("<DataSet ") == 3 and 'file="sod2d_t0.05.vts"' in pvd
    text = (tmp_path / "sod2d_t0.05.vts").read_text()
    assert 'WholeExtent="0 100 0 8 0 1"' in text
    names = re.findall(r'Name="([^"]+)"', text)
    assert names == ["BC_ID", "BC_ID_rho", "BC_ID_U", "BC_ID_p", "MomentumX", "MomentumY",
                     "MomentumZ", "SoundSpeed", "Temperature", "TotalEnergyDensity", "U", "p",
                     "rho"]
    rho = _array(text, "rho").reshape(2, 9, 101)
    # both node layers of the empty z axis carry the state
    np.testing.assert_array_equal(rho[0], out[0].reshape(9, 101))
    np.testing.assert_array_equal(rho[1], out[0].reshape(9, 101))
    energy = _array(text, "TotalEnergyDensity").reshape(2, 9, 101)
    np.testing.assert_array_equal(energy[0], out[4].reshape(9, 101))
