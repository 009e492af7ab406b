"""The C ABI: the library loads and exports every function include/sonic_amd.h
declares; host-side argument checks run without touching a GPU."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "sonic_amd.h"


def declared():
    text = "".join(h.read_text() for h in sorted((ROOT / "include").glob("*.h")))
    return sorted(set(re.findall(r"^int\s+(sf_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_abi():
    names = declared()
    for n in ("sf_grid_create", "sf_grid_destroy", "sf_state_upload", "sf_state_download",
              "sf_step_prepare", "sf_compute_dt", "sf_step_begin", "sf_stage", "sf_step_commit",
              "sf_run", "sf_residual", "sf_error_flags", "sf_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from sonicsolver_amd import _native
    lib = _native.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sf_\w+)$", out, flags=re.M))
    for n in declared():
        assert n in exported, n
        assert hasattr(lib, n)


def test_library_is_gfx950_code():
    from sonicsolver_amd import _native
    blob = _native.LIB_PATH.read_bytes()
    assert b"gfx950" in blob


def test_invalid_descriptions_rejected_on_host():
    from sonicsolver_amd import _native
    lib = _native.load()
    bad = [
        dict(n=(3, 1, 1)),                                   # too few nodes
        dict(n=(16, 2, 1), active=(1, 0, 0)),                # inactive axis with 2 nodes
        dict(n=(16, 16, 16), active=(1, 0, 1)),              # z active without y
        dict(n=(16, 1, 1), gamma=1.0),                       # gamma <= 1
        dict(n=(16, 1, 1), d=(0.0, 1.0, 1.0)),               # zero spacing
        dict(n=(16, 1, 1), scheme=7),
        dict(n=(16, 1, 1), flux=6),
        dict(n=(16, 1, 1), time=9),
        dict(n=(16, 1, 1), bc=(5, 0, 0)),
        dict(n=(16, 1, 1), scheme=4),
        dict(n=(16, 1, 1), bc=(-1, 0, 0)),
    ]
    for kw in bad:
        desc = _native.make_desc(**kw)
        h = ctypes.c_void_p()
        rc = lib.sf_grid_create(ctypes.byref(desc), ctypes.byref(h))
        assert rc == -1, kw
        assert not h.value
        buf = ctypes.create_string_buffer(256)
        lib.sf_last_error(buf, 256)
        assert buf.value, kw


def test_null_handles_rejected():
    from sonicsolver_amd import _native
    lib = _native.load()
    assert lib.sf_grid_create(None, None) == -1
    assert lib.sf_run(None, 1, 0.5, 1.0, None) == -1
    assert lib.sf_stage(None, 0) == -1
    assert lib.sf_grid_destroy(None) == 0


def test_oracle_exports_every_declared_symbol():
    text = (ROOT / "oracle" / "sfcpu.h").read_text()
    names = set(re.findall(r"\b(sfo_\w+)\s*\(", text))
    from oracle import bind as ob
    lib = ob.oracle()
    for n in names:
        assert hasattr(lib, n), n


def test_batch_items_checked_on_host():
    from sonicsolver_amd.batch import make_item, item_nodes
    from sonicsolver_amd import SonicError
    assert item_nodes(make_item((1001,), (1e-3,))) == 1007
    assert item_nodes(make_item((64, 64), (0.1, 0.1))) == 70 * 70
    for args in [dict(n=(3,), d=(0.1,)), dict(n=(16, 2), d=(0.1, 0.1)),
                 dict(n=(16,), d=(0.0,)), dict(n=(16,), d=(0.1,), scheme=3),
                 dict(n=(16,), d=(0.1,), flux=9), dict(n=(16,), d=(0.1,), cfl=0.0)]:
        with pytest.raises(SonicError) as e:
            item_nodes(make_item(**args))
        assert e.value.code == -1


def test_batch_null_handles_rejected():
    from sonicsolver_amd import batch
    lib = batch._lib()
    assert lib.sf_batch_create(0, 1.4, 1000, 4, None) == -1
    assert lib.sf_batch_step(None, 1, 1) == -1
    assert lib.sf_batch_remove(None, 0) == -1
    assert lib.sf_batch_set_tiles(None, 127, 32) == -1
    assert lib.sf_batch_destroy(None) == 0
