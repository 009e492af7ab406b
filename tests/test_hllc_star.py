"""HLLC of the oracle (oracle/sfcpu.c sfo_hllc; no reference counterpart)
pinned by the relations its wave fan must satisfy (tests/hllc_fan.py):
constant outer states, the contact moving with S* and carrying no mass,
one star pressure, pressure work as the energy flux through the contact,
Rankine-Hugoniot at every wave and the integral condition, on random
subsonic pairs whose contact speed differs from both sides' velocities.
The same checks fail on copies of the oracle whose star energy drops the
p / (rho (S_K - u_K)) term or flips its sign, built here on the CPU."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import bind as ob
from tests import hllc_fan as hf

SRC = Path(ob.__file__).resolve().parent / "sfcpu.c"
STAR_ENERGY = "(Ss - un) * (Ss + s->p / mK)"


def _fluxes(fn, qL, qR, n, rays):
    out = []
    for a, b in zip(*hf.shifted_pairs(qL, qR, n, rays)):
        rc, f = fn(a, b, n)
        assert rc == 0
        out.append(f)
    return np.array(out)


def _oracle_hllc(a, b, n):
    return ob.hllc(a, b, n)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_hllc_fan(axis):
    n = hf.NORMALS[axis]
    for qL, qR in hf.subsonic_pairs(6, seed=40 + axis):
        rays = hf.fan_rays(qL, qR, n)
        G = hf.fan_from_fluxes(rays, _fluxes(_oracle_hllc, qL, qR, n, rays), n)
        SL, Ss, SR, ps = hf.check_fan(qL, qR, n, rays, G)
        uL, uR = qL[1:4] @ n / qL[0], qR[1:4] @ n / qR[0]
        assert SL < Ss < SR and min(abs(Ss - uL), abs(Ss - uR)) > 1e-3


def _mutant(tmp_path, replacement):
    src = SRC.read_text()
    assert src.count(STAR_ENERGY) == 1
    c = tmp_path / "sfcpu.c"
    c.write_text(src.replace(STAR_ENERGY, replacement))
    shutil.copy(SRC.with_name("sfcpu.h"), tmp_path / "sfcpu.h")
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o",
                    str(so), str(c), "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    P = ctypes.POINTER(ctypes.c_double)
    lib.sfo_hllc.argtypes = [P, P, P, ctypes.c_double, P]

    def fn(a, b, n):
        a, b, n = (np.ascontiguousarray(x, dtype=float) for x in (a, b, n))
        out = np.zeros(5)
        rc = lib.sfo_hllc(a.ctypes.data_as(P), b.ctypes.data_as(P), n.ctypes.data_as(P), 1.4,
                          out.ctypes.data_as(P))
        return rc, out
    return fn


@pytest.mark.parametrize("replacement", ["(Ss - un) * (Ss)", "(Ss - un) * (Ss - s->p / mK)"])
def test_fan_checks_catch_a_wrong_star_energy(tmp_path, replacement):
    fn = _mutant(tmp_path, replacement)
    n = hf.NORMALS[0]
    caught = 0
    pairs = hf.subsonic_pairs(6, seed=40)
    for qL, qR in pairs:
        rays = hf.fan_rays(qL, qR, n)
        G = hf.fan_from_fluxes(rays, _fluxes(fn, qL, qR, n, rays), n)
        try:
            hf.check_fan(qL, qR, n, rays, G)
        except AssertionError:
            caught += 1
    assert caught == len(pairs)
