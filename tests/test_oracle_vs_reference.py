"""The oracle (oracle/sfcpu.c) against the reference's own code, run through
the harness oracle/_ref/libsfref.so (built from the sonicSolver sources by
`make -C oracle ref`).  Both are compiled with -O2 -ffp-contract=off, so the
comparison is bit-for-bit.  Skipped where the harness has not been built;
the committed golden fixtures (tests/test_golden.py) carry the same pin
there."""
import itertools

import numpy as np
import pytest

from oracle import bind as ob
from sonicsolver_amd import workloads as wl

pytestmark = pytest.mark.skipif(not ob.ref_available(), reason="reference harness not built")


def test_point_kernels_bit_exact():
    rng = np.random.default_rng(17)
    ref = ob.ref()
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    for _ in range(200):
        v = rng.normal(size=5) * rng.choice([1e-3, 1.0, 1e3])
        if rng.random() < 0.3:
            v[rng.integers(3, 5):] += rng.choice([-1, 1]) * 10.0  # a jump
        a = np.ascontiguousarray(v)
        assert ob.weno5(v)[0] == ref.sfref_weno5(a.ctypes.data_as(P))
        assert ob.teno5(v)[0] == ref.sfref_teno5(a.ctypes.data_as(P))
    for _ in range(50):
        q = wl.conserved(np.array([rng.uniform(0.1, 3)]), *rng.uniform(-2, 2, 3),
                         rng.uniform(0.1, 5))[:, 0].copy()
        nv = rng.normal(size=3)
        for n in (np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]),
                  nv / np.linalg.norm(nv)):
            L, R = np.zeros(25), np.zeros(25)
            ref.sfref_char_matrix(q.ctypes.data_as(P), np.ascontiguousarray(n).ctypes.data_as(P),
                                  1.4, L.ctypes.data_as(P), R.ctypes.data_as(P))
            rc, Lo, Ro = ob.char_matrix(q, n)
            assert np.array_equal(Lo.ravel(), L) and np.array_equal(Ro.ravel(), R)
            q2 = wl.conserved(np.array([rng.uniform(0.1, 3)]), *rng.uniform(-2, 2, 3),
                              rng.uniform(0.1, 5))[:, 0].copy()
            out = np.zeros(5)
            ref.sfref_rusanov(q.ctypes.data_as(P), q2.ctypes.data_as(P),
                              np.ascontiguousarray(n).ctypes.data_as(P), 1.4,
                              out.ctypes.data_as(P))
            assert np.array_equal(ob.rusanov(q, q2, n)[1], out)


def test_characteristic_face_flux_bit_exact():
    rng = np.random.default_rng(23)
    ref = ob.ref()
    checked = 0
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    for _ in range(100):
        rho = rng.uniform(0.2, 2, 6)
        if rng.random() < 0.5:
            rho[3:] *= 0.2
        st = wl.conserved(rho, rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 6),
                          rng.uniform(-1, 1, 6), rng.uniform(0.5, 2, 6)).T.copy()
        for axis in range(3):
            cof = np.zeros(3)
            cof[axis] = rng.uniform(0.01, 2)
            for scheme, flux in itertools.product(("weno5", "teno5"), ("rusanov", "hllc")):
                rc, o = ob.char_face_flux(st, cof, scheme, flux)
                if rc != 0:
                    # the reference terminates the process on a non-physical
                    # reconstruction; such inputs are covered by the oracle's
                    # own error-path tests
                    continue
                checked += 1
                out = np.zeros(5)
                assert ref.sfref_char_face_flux(st.ctypes.data_as(P), cof.ctypes.data_as(P),
                                                ob.SCHEMES[scheme], ob.FLUXES[flux], 1.4,
                                                out.ctypes.data_as(P)) == 0
                assert np.array_equal(o, out), (axis, scheme, flux)
    assert checked > 1000


CASES = [
    ((120, 1, 1), (1, 0, 0), ("transmissive",) * 3),
    ((24, 20, 1), (1, 1, 0), ("transmissive",) * 3),
    ((24, 20, 1), (1, 1, 0), ("periodic", "periodic", "transmissive")),
    ((24, 20, 1), (1, 1, 0), ("periodic", "transmissive", "transmissive")),
    ((12, 10, 9), (1, 1, 1), ("periodic",) * 3),
    ((12, 10, 9), (1, 1, 1), ("transmissive", "periodic", "transmissive")),
    ((12, 10, 9), (1, 1, 1), ("transmissive",) * 3),
]


@pytest.mark.parametrize("n,active,bc", CASES)
@pytest.mark.parametrize("scheme,flux,time", [
    ("weno5", "rusanov", "ssprk3"), ("teno5", "hllc", "ssprk3"), ("weno5", "hllc", "rk4"),
    ("teno5", "rusanov", "euler")])
def test_trajectory_bit_exact(n, active, bc, scheme, flux, time):
    q = wl.smooth_field(n, seed=31)
    desc = ob.make_desc(n, active, bc, (0.1, 0.12, 0.09), scheme=scheme, flux=flux, time=time)
    a, da = ob.run(desc, q, 4, 0.5)
    b, db = ob.ref_run(desc, q, 4, 0.5)
    assert np.array_equal(da, db)
    assert np.array_equal(a, b)


def test_sod_full_config_bit_exact():
    """The 1D shock tube configuration at full size (1000 cells, 200 steps)."""
    kw, q = wl.sod_1d(1000)
    desc = ob.make_desc(**kw)
    a, da = ob.run(desc, q, 200, 0.5)
    b, db = ob.ref_run(desc, q, 200, 0.5)
    assert np.array_equal(da, db) and np.array_equal(a, b)


def test_residual_bit_exact():
    for n, active, bc in CASES:
        q = wl.smooth_field(n, seed=5)
        desc = ob.make_desc(n, active, bc, (0.1, 0.12, 0.09), scheme="teno5", flux="hllc")
        assert np.array_equal(ob.residual(desc, q), ob.ref_residual(desc, q))


@pytest.mark.parametrize("end,end_step,by_step,wsteps,wtime", [
    (0.05, 0, 0, 0, 0.02), (0.05, 0, 1, 7, 1e300), (1.0, 13, 0, 0, 0.3)])
def test_time_driver_bit_exact(end, end_step, by_step, wsteps, wtime):
    """The oracle's Time::Driver loop against the reference's own
    (SF_time.cpp:42-93): end time, output instants, step outputs, end step."""
    import ctypes
    ref = ob.ref()
    P = ctypes.POINTER(ctypes.c_double)
    ref.sfref_run_driver.argtypes = [
        ctypes.POINTER(ob.Desc), P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, ctypes.c_int,
        ctypes.POINTER(ctypes.c_int)]
    kw, q = wl.sod_1d(200)
    a = q.copy()
    d = np.zeros(2000)
    n = ctypes.c_int()
    assert ref.sfref_run_driver(ctypes.byref(ob.make_desc(**kw)), a.ctypes.data_as(P), 0.0, end,
                                end_step, by_step, wsteps, wtime, 0.5, 1e300,
                                d.ctypes.data_as(P), 2000, ctypes.byref(n)) == 0
    b, bd = ob.run_driver(ob.make_desc(**kw), q, 0.0, end, 0.5, end_step=end_step,
                          write_by_step=bool(by_step), write_steps=wsteps, write_time=wtime)
    assert n.value == len(bd)
    assert np.array_equal(d[:n.value], bd) and np.array_equal(a, b)


@pytest.mark.skipif(not ob.ref_fma_available(), reason="reference FMA build not built")
def test_teno5_last_bit_sensitivity_of_the_reference():
    """How far the reference moves from itself when only FMA contraction
    changes (its release build contracts; the harness build does not), on
    the batched configuration's Riemann problems.  WENO5 stays at rounding
    level after 60 steps; TENO5 (cutoff 1e-5, epsilon 1e-40) stays within
    1e-12 for 20 steps and then amplifies the last-bit differences far past
    1e-11.  This is the yardstick the device TENO5 long-horizon parity is
    measured against."""
    from tests.parity import rel_errors
    worst = {"weno5": 0.0, "teno5": 0.0}
    for _, kw, q in wl.batch_riemann_1d(8, seed=4096):
        n = kw["n"][0]
        d = ob.make_desc(n=(n, 1, 1), active=(1, 0, 0), d=(kw["d"][0], 1.0, 1.0),
                         scheme=kw["scheme"], flux=kw["flux"], time="ssprk3")
        a, _ = ob.ref_run(d, q, 20, 0.5)
        b, _ = ob.ref_run(d, q, 20, 0.5, fma=True)
        assert rel_errors(b, a)[1].max() <= 1e-12
        a, _ = ob.ref_run(d, q, 60, 0.5)
        b, _ = ob.ref_run(d, q, 60, 0.5, fma=True)
        worst[kw["scheme"]] = max(worst[kw["scheme"]], rel_errors(b, a)[1].max())
    assert worst["weno5"] <= 1e-13
    assert worst["teno5"] > 1e-9


def test_weno3_weno7_cores_bit_exact():
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(41)
    ref = ob.ref()
    for _ in range(300):
        v = rng.normal(size=7) * rng.choice([1e-3, 1.0, 1e3])
        if rng.random() < 0.3:
            v[rng.integers(2, 7):] += rng.choice([-1, 1]) * 10.0
        a = np.ascontiguousarray(v)
        assert ob.weno7(v)[0] == ref.sfref_weno7(a.ctypes.data_as(P))
        b = np.ascontiguousarray(v[:4])
        assert ob.weno3(v[:4])[0] == ref.sfref_weno3(b.ctypes.data_as(P))


def test_splitting_fluxes_bit_exact():
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(43)
    ref = ob.ref()
    for _ in range(200):
        qL = wl.conserved(np.array([rng.uniform(0.1, 3)]), *rng.uniform(-2, 2, 3),
                          rng.uniform(0.1, 5))[:, 0].copy()
        qR = wl.conserved(np.array([rng.uniform(0.1, 3)]), *rng.uniform(-2, 2, 3),
                          rng.uniform(0.1, 5))[:, 0].copy()
        if rng.random() < 0.2:
            qR = qL.copy()
            qR[1:4] *= 0.001        # slow flow: the Steger-Warming smoothing band
        nv = rng.normal(size=3)
        for n in (np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]),
                  nv / np.linalg.norm(nv)):
            n = np.ascontiguousarray(n)
            for flux in ("steger", "laxfriedrichs", "roe"):
                out = np.zeros(5)
                assert ref.sfref_split_flux(ob.FLUXES[flux], qL.ctypes.data_as(P),
                                            qR.ctypes.data_as(P), n.ctypes.data_as(P), 1.4,
                                            out.ctypes.data_as(P)) == 0
                rc, o = ob.face_flux(flux, qL, qR, n)
                assert rc == 0 and np.array_equal(o, out), flux


@pytest.mark.parametrize("scheme", ["weno3", "weno5", "teno5", "weno7"])
def test_every_scheme_and_flux_face_bit_exact(scheme):
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(47)
    ref = ob.ref()
    ns = ob.STENCIL[ob.SCHEMES[scheme]]
    checked = 0
    for _ in range(40):
        rho = rng.uniform(0.2, 2, ns)
        if rng.random() < 0.5:
            rho[ns // 2:] *= 0.2
        st = wl.conserved(rho, rng.uniform(-1, 1, ns), rng.uniform(-1, 1, ns),
                          rng.uniform(-1, 1, ns), rng.uniform(0.5, 2, ns)).T.copy()
        for axis in range(3):
            cof = np.zeros(3)
            cof[axis] = rng.uniform(0.01, 2)
            for flux in ob.FLUXES:
                rc, o = ob.char_face_flux(st, cof, scheme, flux)
                if rc != 0:
                    continue
                out = np.zeros(5)
                assert ref.sfref_char_face_flux_n(st.ctypes.data_as(P), cof.ctypes.data_as(P),
                                                  ob.SCHEMES[scheme], ob.FLUXES[flux], 1.4,
                                                  out.ctypes.data_as(P)) == 0
                assert np.array_equal(o, out), (axis, scheme, flux)
                checked += 1
    assert checked > 400


@pytest.mark.parametrize("n,active,bc", [
    ((40, 1, 1), (1, 0, 0), ("transmissive",) * 3),
    ((16, 14, 1), (1, 1, 0), ("periodic", "transmissive", "transmissive")),
    ((16, 14, 1), (1, 1, 0), ("none_lo", "periodic", "transmissive")),
    ((10, 9, 8), (1, 1, 1), ("transmissive", "periodic", "none"))])
@pytest.mark.parametrize("scheme,flux,time", [
    ("weno3", "steger", "ssprk3"), ("weno7", "roe", "rk4"), ("weno7", "laxfriedrichs", "euler"),
    ("weno3", "hllc", "rk4"), ("weno5", "roe", "ssprk3"), ("teno5", "steger", "ssprk3"),
    ("weno7", "rusanov", "ssprk3")])
def test_new_schemes_and_fluxes_trajectory_bit_exact(n, active, bc, scheme, flux, time):
    q = wl.smooth_field(n, seed=37)
    desc = ob.make_desc(n, active, bc, (0.1, 0.12, 0.09), scheme=scheme, flux=flux, time=time)
    a, da = ob.run(desc, q, 3, 0.4)
    b, db = ob.ref_run(desc, q, 3, 0.4)
    assert np.array_equal(da, db)
    assert np.array_equal(a, b)


def test_weno7_sod_baseline_case_bit_exact():
    """The frozen WENO7 Sod case end to end (454 steps, output-aligned dt)."""
    from tests.test_oracle_schemes import sod_weno7_desc, sod_weno7_state
    q = sod_weno7_state()
    a, da = ob.run_driver(sod_weno7_desc(), q, 0.0, 0.2, 0.5, write_time=0.05)
    b, db = ob.ref_run_driver(sod_weno7_desc(), q, 0.0, 0.2, 0.5, write_time=0.05)
    assert np.array_equal(da, db) and np.array_equal(a, b)


def test_lax_wendroff_bit_exact():
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(53)
    ref = ob.ref()
    for _ in range(100):
        qL = wl.conserved(np.array([rng.uniform(0.1, 3)]), *rng.uniform(-2, 2, 3),
                          rng.uniform(0.1, 5))[:, 0].copy()
        qR = wl.conserved(np.array([rng.uniform(0.1, 3)]), *rng.uniform(-2, 2, 3),
                          rng.uniform(0.1, 5))[:, 0].copy()
        nv = rng.normal(size=3)
        dt, h = rng.uniform(1e-4, 1e-2), rng.uniform(0.01, 0.2)
        for n in (np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), nv / np.linalg.norm(nv)):
            n = np.ascontiguousarray(n)
            rc, o = ob.lax_wendroff(qL, qR, n, dt, h)
            if rc != 0:
                continue
            out = np.zeros(5)
            assert ref.sfref_lax_wendroff(qL.ctypes.data_as(P), qR.ctypes.data_as(P),
                                          n.ctypes.data_as(P), 1.4, dt, h,
                                          out.ctypes.data_as(P)) == 0
            assert np.array_equal(o, out)
    # every scheme's characteristic face with it
    for scheme in ("weno3", "weno5", "teno5", "weno7"):
        ns = ob.STENCIL[ob.SCHEMES[scheme]]
        for _ in range(20):
            st = wl.conserved(rng.uniform(0.5, 2, ns), rng.uniform(-0.5, 0.5, ns),
                              rng.uniform(-0.5, 0.5, ns), rng.uniform(-0.5, 0.5, ns),
                              rng.uniform(0.5, 2, ns)).T.copy()
            cof = np.array([0.0, 0.7, 0.0])
            rc, o = ob.char_face_flux(st, cof, scheme, "laxwendroff", dt=0.003, spacing=0.05)
            assert rc == 0
            out = np.zeros(5)
            assert ref.sfref_char_face_flux_dt(st.ctypes.data_as(P), cof.ctypes.data_as(P),
                                               ob.SCHEMES[scheme], 5, 1.4, 0.003, 0.05,
                                               out.ctypes.data_as(P)) == 0
            assert np.array_equal(o, out)


@pytest.mark.parametrize("scheme,time", [("weno5", "ssprk3"), ("weno7", "euler"),
                                         ("teno5", "rk4"), ("weno3", "ssprk3")])
def test_lax_wendroff_trajectory(scheme, time):
    """The step's dt reaches the flux at every stage.  The harness takes the
    face spacing from its node coordinates (x_i = (i - g) dx), the oracle
    takes dx itself: they differ in the last bit, so this compares to 1e-13."""
    from tests.parity import rel_errors
    n = (18, 14, 1)
    q = wl.smooth_field(n, seed=29)
    desc = ob.make_desc(n, (1, 1, 0), ("none_lo", "periodic", "transmissive"),
                        (0.1, 0.12, 0.09), scheme=scheme, flux="laxwendroff", time=time)
    a, da = ob.run(desc, q, 4, 0.4)
    b, db = ob.ref_run(desc, q, 4, 0.4)
    np.testing.assert_allclose(da, db, rtol=1e-14)
    l1, linf = rel_errors(a, b)
    assert linf.max() < 1e-13


@pytest.mark.parametrize("scheme,flux,time", [
    ("weno5", "hllc", "ssprk3"), ("teno5", "roe", "rk4"), ("weno7", "laxwendroff", "ssprk3"),
    ("weno3", "steger", "euler")])
def test_slab_decomposed_reference_run_is_the_serial_run(scheme, flux, time):
    """The all-cores CPU baseline: the reference's step on z slabs, one
    thread each, with the halo planes exchanged before every ghost fill and
    the dt reduced over slabs, gives the serial run bit for bit (uneven
    slabs included), and the serial run is the oracle's."""
    kw, q = wl.tgv_3d(20)
    kw = dict(kw, n=(20, 18, 20), d=(kw["d"][0], kw["d"][0] * 20 / 18, kw["d"][0]),
              scheme=scheme, flux=flux, time=time)
    q = wl.smooth_field((20, 18, 20), seed=41)
    d = ob.make_desc(**kw)
    a, da = ob.ref_run(d, q, 3, 0.4)
    o, do = ob.run(ob.make_desc(**kw), q, 3, 0.4)
    assert np.array_equal(a, o.reshape(5, -1)) and np.array_equal(da, do)
    halo = {"weno3": 2, "weno5": 3, "teno5": 3, "weno7": 4}[scheme]
    for nslabs in (1, 3, 20 // halo):
        b, db, sec = ob.ref_run_slabs(d, q, 3, 0.4, nslabs)
        assert np.array_equal(a, b) and np.array_equal(da, db), nslabs
        assert sec > 0


def test_slab_run_rejects_what_it_cannot_cut():
    kw, q = wl.tgv_3d(12)
    with pytest.raises(RuntimeError):
        ob.ref_run_slabs(ob.make_desc(**kw), q, 1, 0.4, 5)          # 12 < 5 slabs x 3 nodes
    kw2 = dict(kw, bc=("periodic", "periodic", "transmissive"))
    with pytest.raises(RuntimeError):
        ob.ref_run_slabs(ob.make_desc(**kw2), q, 1, 0.4, 2)         # z not periodic
