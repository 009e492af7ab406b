"""Comparison helpers shared by the parity tests."""
import numpy as np

# Parity tolerance: relative L1 and Linf error on the conserved variables.
REL_TOL = 1e-11


def rel_errors(a, b):
    """Per-variable relative L1 and Linf errors of a against b (5 x N).

    Density and total energy are measured against their own norms.  The
    momentum components are measured against the norm of the momentum
    vector (the largest component norm): a component that is zero up to
    rounding (rhoV in a problem that is uniform along y) carries only
    rounding noise, whose relative difference to itself says nothing."""
    a = np.asarray(a).reshape(5, -1)
    b = np.asarray(b).reshape(5, -1)
    l1 = []
    linf = []
    whole_inf = max(np.abs(b).max(), 1e-300)
    whole_l1 = max(np.abs(b).sum() / 5.0, 1e-300)
    mom_inf = max(np.abs(b[v]).max() for v in (1, 2, 3))
    mom_l1 = max(np.abs(b[v]).sum() for v in (1, 2, 3))
    for v in range(5):
        if v in (1, 2, 3):
            scale_inf, scale_l1 = mom_inf, mom_l1
        else:
            scale_inf, scale_l1 = np.abs(b[v]).max(), np.abs(b[v]).sum()
        # a state at rest is measured against the scale of the whole state
        scale_inf = scale_inf or whole_inf
        scale_l1 = scale_l1 or whole_l1
        d = np.abs(a[v] - b[v])
        linf.append(d.max() / scale_inf)
        l1.append(d.sum() / scale_l1)
    return np.array(l1), np.array(linf)


def assert_close(a, b, tol=REL_TOL, what=""):
    l1, linf = rel_errors(a, b)
    assert np.all(np.isfinite(a)), f"{what}: non-finite values"
    assert l1.max() <= tol and linf.max() <= tol, \
        f"{what}: rel L1 {l1.max():.3e}, rel Linf {linf.max():.3e} exceed {tol:g}"
    return l1.max(), linf.max()
