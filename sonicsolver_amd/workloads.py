"""Synthetic initial states of the benchmark configurations.

Each generator returns ``(desc_kwargs, q)`` where ``q`` is the conserved
state, variable-major over the unique nodes, ``q[v, (k*ny + j)*nx + i]``,
v = (rho, rhoU, rhoV, rhoW, rhoE).  Grids are vertex-centred as in the
reference (solution points are mesh nodes).  Random draws use splitmix64
so the inputs are reproducible bit for bit on any host.
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = 1.4
_MASK = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` successive outputs of splitmix64 (Steele, Lea & Flood 2014)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _MASK) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def conserved(rho, u, v, w, p, gamma=GAMMA) -> np.ndarray:
    rho = np.asarray(rho, dtype=np.float64)
    u, v, w, p = (np.broadcast_to(np.asarray(a, dtype=np.float64), rho.shape) for a in (u, v, w, p))
    e = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, e]).reshape(5, -1).copy()


def primitive(q, gamma=GAMMA):
    q = np.asarray(q).reshape(5, -1)
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    p = (gamma - 1.0) * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
    return rho, u, v, w, p


def sod_1d(cells: int = 1000, scheme="weno5", flux="rusanov", time="ssprk3"):
    """1D shock tube on x in [0, 1]: cells+1 nodes, (1, 0, 1) on the nodes
    left of the midpoint (the mesh generator's ``leftHalf`` set holds nodes
    i < nx/2), (0.125, 0, 0.1) on the others."""
    nx = cells + 1
    i = np.arange(nx)
    left = i < nx // 2
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    q = conserved(rho, 0.0, 0.0, 0.0, p)
    desc = dict(n=(nx, 1, 1), active=(1, 0, 0), bc=("transmissive",) * 3,
                d=(1.0 / cells, 1.0, 1.0), scheme=scheme, flux=flux, time=time)
    return desc, q


def vortex_state(x, y, t, beta=5.0, length=10.0, u0=1.0, v0=1.0, gamma=GAMMA):
    """Exact isentropic vortex (Shu 1998) of strength beta advected by (u0, v0),
    on the periodic square [0, length)^2, at time t."""
    xc = (length / 2 + u0 * t) % length
    yc = (length / 2 + v0 * t) % length
    dx = (x - xc + length / 2) % length - length / 2
    dy = (y - yc + length / 2) % length - length / 2
    r2 = dx * dx + dy * dy
    f = beta / (2.0 * math.pi) * np.exp(0.5 * (1.0 - r2))
    u = u0 - f * dy
    v = v0 + f * dx
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * math.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return rho, u, v, p


def vortex_2d(n: int = 2048, noise: float = 1e-3, seed: int = 2026, scheme="teno5",
              flux="hllc", time="ssprk3", length: float = 10.0):
    """Isentropic vortex on the periodic square [0, 10)^2 with n x n unique
    nodes, plus rho *= 1 + noise*U[-1, 1] (pressure unchanged)."""
    h = length / n
    y, x = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="ij")
    rho, u, v, p = vortex_state(x, y, 0.0, length=length)
    if noise:
        rho = rho * (1.0 + noise * (2.0 * uniform01(seed, n * n).reshape(n, n) - 1.0))
    q = conserved(rho, u, v, 0.0, p)
    desc = dict(n=(n, n, 1), active=(1, 1, 0), bc=("periodic", "periodic", "transmissive"),
                d=(h, h, 1.0), scheme=scheme, flux=flux, time=time)
    return desc, q


def tgv_3d(n: int = 512, seed: int = 2026, scheme="weno5", flux="hllc", time="ssprk3"):
    """Taylor-Green-type field on the periodic cube [0, 2 pi)^3:
    u = sin(x+p1)cos(y+p2)cos(z+p3), v = -cos(x+p1)sin(y+p2)cos(z+p3), w = 0,
    rho = 1, p = 100 + (rho/16)(cos 2x + cos 2y)(cos 2z + 2); phases p_i
    uniform in [0, 2 pi) from splitmix64(seed)."""
    h = 2.0 * math.pi / n
    ph = 2.0 * math.pi * uniform01(seed, 3)
    c = np.arange(n) * h
    # separable factors broadcast over (z, y, x); the products are formed in
    # the same order as the pointwise formula
    x, y, z = c[None, None, :], c[None, :, None], c[:, None, None]
    u = (np.sin(x + ph[0]) * np.cos(y + ph[1])) * np.cos(z + ph[2])
    v = (-np.cos(x + ph[0]) * np.sin(y + ph[1])) * np.cos(z + ph[2])
    p = 100.0 + 1.0 / 16.0 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    q = np.empty((5, n, n, n))
    q[0] = 1.0
    q[1] = u
    q[2] = v
    q[3] = 0.0
    q[4] = p / (GAMMA - 1.0) + 0.5 * 1.0 * (u * u + v * v + 0.0 * 0.0)
    q = q.reshape(5, -1)
    desc = dict(n=(n, n, n), active=(1, 1, 1), bc=("periodic",) * 3, d=(h, h, h),
                scheme=scheme, flux=flux, time=time)
    return desc, q


def smooth_field(n, seed: int = 7, amp: float = 0.2):
    """A smooth positive state with velocity in every active direction (tests)."""
    nx, ny, nz = (list(n) + [1, 1])[:3]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = uniform01(seed, 6)
    tx, ty, tz = 2 * np.pi * x / nx, 2 * np.pi * y / max(ny, 1), 2 * np.pi * z / max(nz, 1)
    rho = 1.0 + amp * np.sin(tx + 6 * r[0]) * np.cos(ty + 6 * r[1]) + 0.5 * amp * np.sin(tz + 6 * r[2])
    u = 0.5 + amp * np.cos(ty + 6 * r[3])
    v = -0.3 + amp * np.sin(tx + 6 * r[4])
    w = 0.2 * np.cos(tz + 6 * r[5])
    p = 1.0 + amp * np.cos(tx + ty)
    return conserved(rho, u, v, w, p)


def _loguniform(u, lo, hi):
    return np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))


def riemann_1d(cells: int, left, right, interface: float):
    """1D Riemann problem on x in [0, 1] with cells+1 nodes; node x_i = i/cells
    takes the left (rho, u, p) where x_i < interface.  Returns (item kwargs
    without scheme/flux, q)."""
    nx = cells + 1
    x = np.arange(nx) / cells
    lft = x < interface
    rho = np.where(lft, left[0], right[0])
    u = np.where(lft, left[1], right[1])
    p = np.where(lft, left[2], right[2])
    return dict(n=(nx, 1), d=(1.0 / cells, 1.0)), conserved(rho, u, 0.0, 0.0, p)


def batch_riemann_1d(count: int = 4096, seed: int = 4096, lo: int = 256, hi: int = 8192,
                     cfl: float = 0.5):
    """Independent 1D Riemann problems of uneven length: L = round(exp(U[ln lo,
    ln hi])) cells; rho log-uniform [0.1, 2], u uniform [-1, 1], p log-uniform
    [0.1, 10] on each side; interface uniform in [0.3, 0.7]; problem i uses
    WENO5 for even i, TENO5 for odd i, and Rusanov / HLLC alternating with
    (i >> 1) & 1.  Yields (index, item kwargs, q)."""
    r = uniform01(seed, 8 * count).reshape(count, 8)
    for i in range(count):
        cells = int(round(float(_loguniform(r[i, 0], lo, hi))))
        left = (float(_loguniform(r[i, 1], 0.1, 2.0)), -1.0 + 2.0 * r[i, 2],
                float(_loguniform(r[i, 3], 0.1, 10.0)))
        right = (float(_loguniform(r[i, 4], 0.1, 2.0)), -1.0 + 2.0 * r[i, 5],
                 float(_loguniform(r[i, 6], 0.1, 10.0)))
        x0 = 0.3 + 0.4 * r[i, 7]
        kw, q = riemann_1d(cells, left, right, x0)
        kw.update(scheme="teno5" if i & 1 else "weno5",
                  flux="hllc" if (i >> 1) & 1 else "rusanov", cfl=cfl)
        yield i, kw, q


def quadrant_2d(n: int, states):
    """Four-quadrant 2D Riemann problem on [0, 1]^2 with n x n nodes; states
    are (rho, u, v, p) of the quadrants (x>=1/2, y>=1/2), (x<1/2, y>=1/2),
    (x<1/2, y<1/2), (x>=1/2, y<1/2), counter-clockwise from the upper right.
    Returns (item kwargs without scheme/flux, q)."""
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x, indexing="xy")
    right, top = X >= 0.5, Y >= 0.5
    quad = np.where(top, np.where(right, 0, 1), np.where(right, 3, 2)).ravel()
    st = np.asarray(states, dtype=np.float64)
    rho, u, v, p = (st[quad, k] for k in range(4))
    return dict(n=(n, n), d=(h, h)), conserved(rho, u, v, 0.0, p)


def quadrant_requests(count: int, seed: int = 5005, lo: int = 64, hi: int = 1024,
                      steps=(50, 500), cfl: float = 0.5):
    """Served 2D request stream: N = round(exp(U[ln lo, ln hi])) nodes per
    side, quadrant (rho, u, v, p) with rho in [0.1, 2], u, v in [-1, 1],
    p in [0.1, 2] (uniform), steps uniform in the closed range `steps`;
    WENO5 / HLLC.  Yields (index, item kwargs, steps, q-factory) -- the
    state is built on demand, the largest is 1024^2 x 5 doubles."""
    r = uniform01(seed, 18 * count).reshape(count, 18)
    for i in range(count):
        n = int(round(float(_loguniform(r[i, 0], lo, hi))))
        states = []
        for k in range(4):
            a = r[i, 1 + 4 * k: 5 + 4 * k]
            states.append((0.1 + 1.9 * a[0], -1.0 + 2.0 * a[1], -1.0 + 2.0 * a[2],
                           0.1 + 1.9 * a[3]))
        nsteps = int(steps[0] + min(int(r[i, 17] * (steps[1] - steps[0] + 1)),
                                    steps[1] - steps[0]))
        kw = dict(n=(n, n), d=(1.0 / (n - 1), 1.0 / (n - 1)), scheme="weno5", flux="hllc",
                  cfl=cfl)
        yield i, kw, nsteps, (lambda n=n, states=states: quadrant_2d(n, states)[1])
