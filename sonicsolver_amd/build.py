"""Build libsonic_amd.so in-tree for gfx950 with hipcc.

    python -m sonicsolver_amd.build

The HIP sources are compiled with packed FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); hipcc's host pass
warns that it does not know that feature, which is harmless.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SONIC_AMD_ARCH", "gfx950")

SOURCES = [PKG / "csrc" / "sf_capi.hip", PKG / "csrc" / "sf_batch.hip", PKG / "csrc" / "sf_diag.hip",
           PKG / "csrc" / "sf_sweep_weno5.hip", PKG / "csrc" / "sf_sweep_teno5.hip",
           PKG / "csrc" / "sf_sweep_weno3.hip", PKG / "csrc" / "sf_sweep_weno7.hip",
           PKG / "host" / "sf_provider.cpp", PKG / "host" / "sf_server.cpp",
           PKG / "host" / "sf_yaml.cpp", PKG / "host" / "sf_case.cpp", PKG / "host" / "sf_vts.cpp"]
HEADERS = list((PKG / "csrc").glob("*.h")) + list((PKG / "host").glob("*.h")) + \
    list((ROOT / "include").glob("*.h"))
OUTPUT = PKG / "libsonic_amd.so"


def hip_flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}", f"-I{PKG / 'host'}"]


def build(verbose: bool = False) -> Path:
    """Compile each source to an object in _build/ (in parallel), then link."""
    objdir = ROOT / "_build" / "sonic_amd"
    objdir.mkdir(parents=True, exist_ok=True)
    srcs = [s for s in SOURCES if s.exists()]
    procs = []
    for src in srcs:
        obj = objdir / (src.name + ".o")
        cmd = [HIPCC, *hip_flags(), "-c", "-o", str(obj), str(src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = False
    for obj, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(OUTPUT) + ".tmp",
           *(str(o) for o, _ in procs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed with status {res.returncode}")
    os.replace(str(OUTPUT) + ".tmp", OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
