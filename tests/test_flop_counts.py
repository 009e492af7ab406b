"""The operation counts used for the roofline (sonicsolver_amd/flops.py)
are those of the CPU restatement compiled with a counting double type
(oracle/flopcount.cpp)."""
import json
import subprocess
from pathlib import Path

from sonicsolver_amd import flops

ORACLE = Path(__file__).resolve().parent.parent / "oracle"


def test_counts_match_instrumented_oracle():
    subprocess.run(["make", "-s", "-C", str(ORACLE), "flopcount"], check=True)
    out = subprocess.run([str(ORACLE / "_build" / "flopcount")], capture_output=True,
                         text=True, check=True).stdout
    c = json.loads(out)
    for (scheme, flux), v in flops.FACE.items():
        assert c["face"][f"{scheme}_{flux}"] == v
    for dim in (1, 2, 3):
        d = c[f"dim{dim}"]
        assert d["residual_per_node"] == flops.RESIDUAL[dim]
        assert d["dt_per_node"] == flops.DT[dim]
        for t in ("euler", "ssprk3", "rk4"):
            assert d[f"{t}_update_per_node_step"] == flops.UPDATE[t]


def test_3d_weno5_hllc_ssprk3_cell_step():
    assert flops.flops_per_cell_step(3, "weno5", "hllc", "ssprk3") == 12266
