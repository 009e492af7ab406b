"""Golden fixtures written from the reference's own code
(tests/golden/make_golden.py): the oracle must reproduce them bit for bit
(CPU), and the HIP path to relative L1/Linf 1e-11 (GPU)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import bind as ob

GOLDEN = sorted((Path(__file__).parent / "golden").glob("*.npz"))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        return meta, z["q_in"], z["q_out"], z["dt"]


def desc_kw(meta):
    return {k: meta[k] for k in ("n", "active", "bc", "d", "scheme", "flux", "time")}


def test_fixtures_present():
    assert len(GOLDEN) >= 6


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.stem)
def test_oracle_reproduces_fixture(path):
    meta, q, ref, dts = load(path)
    out, odts = ob.run(ob.make_desc(**desc_kw(meta)), q, meta["steps"], meta["cfl"])
    assert np.array_equal(odts, dts)
    assert np.array_equal(out, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.stem)
def test_hip_path_matches_fixture(path):
    from sonicsolver_amd import Grid, make_desc
    from tests.parity import assert_close

    meta, q, ref, dts = load(path)
    print(f"golden {path.stem}", flush=True)
    with Grid(make_desc(**desc_kw(meta))) as g:
        g.upload(q)
        gdts = g.run(meta["steps"], meta["cfl"])
        out = g.download()
    np.testing.assert_allclose(gdts, dts, rtol=1e-12, atol=0)
    assert_close(out, ref, what=path.stem)
