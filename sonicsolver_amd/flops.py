"""Algorithmic operation and byte counts of the explicit Euler hot path.

The per-face and per-node floating-point operation counts were obtained by
running the CPU restatement with a counting double type (+, -, *, /, sqrt
= 1 each, so an FMA is 2; pow(x, 6) = 3).  They are algorithmic counts: the
device code performs the same operations (fused where hipcc contracts).
"""

# characteristic reconstruction + face flux, per face (HLLC: subsonic star
# branch; Steger-Warming and Roe: no eigenvalue in their smoothing bands)
FACE = {("weno5", "rusanov"): 1287, ("weno5", "hllc"): 1328, ("weno5", "steger"): 1428,
        ("weno5", "laxfriedrichs"): 1310, ("weno5", "roe"): 1554,
        ("teno5", "rusanov"): 1497, ("teno5", "hllc"): 1538, ("teno5", "steger"): 1638,
        ("teno5", "laxfriedrichs"): 1520, ("teno5", "roe"): 1764,
        ("weno3", "rusanov"): 807, ("weno3", "hllc"): 848, ("weno3", "steger"): 948,
        ("weno3", "laxfriedrichs"): 830, ("weno3", "roe"): 1074,
        ("weno7", "rusanov"): 3327, ("weno7", "hllc"): 3368, ("weno7", "steger"): 3468,
        ("weno7", "laxfriedrichs"): 3350, ("weno7", "roe"): 3594,
        ("weno3", "laxwendroff"): 826, ("weno5", "laxwendroff"): 1306,
        ("teno5", "laxwendroff"): 1516, ("weno7", "laxwendroff"): 3346}
# flux difference and Jacobian scaling, per node, by number of active axes
RESIDUAL = {1: 20, 2: 30, 3: 40}
# CFL scan, per node
DT = {1: 34, 2: 48, 3: 62}
# stage combinations plus post-stage validation, per node and physical step
UPDATE = {"euler": 40, "ssprk3": 132, "rk4": 247}
STAGES = {"euler": 1, "ssprk3": 3, "rk4": 4}

# compulsory HBM traffic per cell and physical step for SSPRK3 on five FP64
# structure-of-arrays variables: stage 1 reads Q^n and writes Q^1 (80 B),
# stages 2 and 3 each read Q^(k), Q^n and write Q^(k+1) (120 B each)
BYTES_PER_CELL_STEP = {"ssprk3": 320, "euler": 80, "rk4": 80 + 3 * 120 + 40}


def flops_per_cell_step(dim: int, scheme: str, flux: str, time: str) -> int:
    """Operations per node and physical step (faces counted as one per node
    and active axis)."""
    s = STAGES[time]
    return s * (dim * FACE[(scheme, flux)] + RESIDUAL[dim]) + UPDATE[time] + DT[dim]


def sweep_flops(n, axis: int, scheme: str, flux: str, final: bool, dim: int, time: str,
                fold_dt: float = 0.0) -> float:
    """Operations of one sweep launch along `axis` over the unique nodes n:
    (n_axis + 1) faces per grid line, plus, for the sweep that applies the
    stage update, the residual, the stage combination and (a fraction
    fold_dt of launches) the next step's CFL scan."""
    nx, ny, nz = n
    nodes = nx * ny * nz
    faces = nodes // n[axis] * (n[axis] + 1)
    f = faces * FACE[(scheme, flux)]
    if final:
        f += nodes * (RESIDUAL[dim] + UPDATE[time] / STAGES[time] + fold_dt * DT[dim])
    return float(f)
