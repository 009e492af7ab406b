"""Comparison helpers shared by the parity tests."""
import numpy as np

# relative L1 and Linf error on conserved variables.
REL_TOL = 1e-11


def rel_errors(a, b):
    """Per-variable relative L1 and Linf errors of a against b (5 x N)."""
    a = np.asarray(a).reshape(5, -1)
    b = np.asarray(b).reshape(5, -1)
    l1 = []
    linf = []
    whole_inf = max(np.abs(b).max(), 1e-300)
    whole_l1 = max(np.abs(b).sum() / 5.0, 1e-300)
    for v in range(5):
        # a variable that is identically zero in b (rhoW in 2D) is measured
        # against the scale of the whole state
        scale_inf = np.abs(b[v]).max() or whole_inf
        scale_l1 = np.abs(b[v]).sum() or whole_l1
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
