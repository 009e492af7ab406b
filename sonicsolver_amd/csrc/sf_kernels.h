// sf_kernels.h -- gfx950 kernels of the explicit Euler hot path besides the
// sweeps (included by sf_capi.hip; the sweeps are in sf_sweep_kernels.h).
//
// Data layout in HBM: conserved state as structure-of-arrays FP64,
// q[v*P + (k*m1 + j)*m0 + i], with a ghost layer as deep as the scheme's
// halo (2 for WENO3, 3 for WENO5 / TENO5, 4 for WENO7) on every active axis
// (m_d = n_d + 2 g) and none on an inactive one.  Work per RK stage:
//
//   ghost fill (O(surface))      k_bc_copy4 / k_bc_energy / k_bc_periodic
//   x sweep                      k_sweep_x:     one lane per x face, the
//                                               flux difference by a
//                                               cross-lane shift; on grids
//                                               with many x lines
//                                               k_sweep_xl walks whole
//                                               lines in 64-node tiles
//   y / z sweeps                 k_sweep_march: one lane per (x, column
//                                               segment), marching along the
//                                               axis with the stencil window
//                                               in LDS, so each node is
//                                               loaded once and each face
//                                               evaluated once
//   the last sweep (x) also applies the stage update, validates the new
//   state and (on the last stage) folds the CFL scan of the next step into
//   a wave-level minimum (DPP moves) plus one 64-bit atomic per wave.
//
// Sweeps accumulate S = sum_d (F_d(i) - F_d(i-1)) and the update multiplies
// by Jac once, as Math::localSpatialResidual does
// (SF_residualAssembly.h:12-32).
#pragma once

#include "sf_sweep_kernels.h"

namespace sfamd {

// ---------------------------------------------------------------------------
// Standalone CFL scan (first step, or sf_compute_dt).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cfl(GridDev G, const double* __restrict__ q, double cfl,
                                             unsigned long long* dtmin, Fault* fault,
                                             long long tag) {
    if (fault->flags) return;
    const long long N = (long long)G.n[0] * G.n[1] * G.n[2];
    unsigned flags = 0;
    double best = INFINITY;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < N;
         u += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(u % G.n[0]);
        const int j = (int)((u / G.n[0]) % G.n[1]);
        const int k = (int)(u / ((long long)G.n[0] * G.n[1]));
        if (!solved_node(G, i, j, k)) continue;
        const long long id = node_index(G, i, j, k);
        double qq[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) qq[v] = q[v * G.P + id];
        best = fmin(best, local_dt(G, qq, cfl, flags));
    }
    raise_fault(flags, &fault->flags, &fault->tag, tag);
    wave_min_atomic(best, dtmin);
}

__global__ void k_finalize_dt(Clock* c, double* dt_hist, int slot, const unsigned* flags) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || *flags) return;
    double dt = __longlong_as_double((long long)c->dtmin);
    dt = fmin(dt, c->max_dt);
    dt = fmin(dt, c->limit);
    c->dt = dt;
    c->dtmin = 0x7FF0000000000000ull;
    if (dt_hist) dt_hist[slot] = dt;
}

__global__ void k_set_dt(Clock* c, double dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->dt = dt;
}

__global__ void k_set_limits(Clock* c, double max_dt, double limit, int reset_scan) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->max_dt = max_dt;
        c->limit = limit;
        if (reset_scan) c->dtmin = 0x7FF0000000000000ull;
    }
}

__global__ void k_commit_clock(Clock* c, const unsigned* flags) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !*flags) {
        c->time += c->dt;
        c->steps += 1;
    }
}

// ---------------------------------------------------------------------------
// Ghost fill.  Thread per boundary-plane point (2 sides x plane).
// ---------------------------------------------------------------------------
template <int AXIS>
__device__ __forceinline__ bool plane_point(const GridDev& G, long long t, int p[3], int& side) {
    constexpr int e1 = (AXIS + 1) % 3, e2 = (AXIS + 2) % 3;
    const long long plane = (long long)G.n[e1] * G.n[e2];
    if (t >= 2 * plane) return false;
    side = (int)(t / plane);
    const long long r = t % plane;
    p[e1] = (int)(r % G.n[e1]);
    p[e2] = (int)(r / G.n[e1]);
    p[AXIS] = side ? G.n[AXIS] - 1 : 0;
    return true;
}

// Threads per boundary point of a fill along AXIS: along y and z one per
// point (a point's neighbours on the plane are adjacent in memory); along
// x, where the boundary node and its ghosts are the adjacent doubles of one
// row and the plane's points lie a row apart, one per (variable, node) for
// the copy and one per node for the energy, so that neighbouring lanes
// write neighbouring doubles.
template <int AXIS>
__host__ __device__ __forceinline__ int fill_lanes(const GridDev& G, bool copy) {
    return AXIS == 0 ? (copy ? 4 : 1) * (G.g[0] + 1) : 1;
}

// zeroGradient density + momentum (SF_zeroGradient.cpp:88-130) on the
// zeroGradient ends of the axis; the ghost depth is G.g[AXIS].
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_copy4(GridDev G, double* __restrict__ q,
                                                  const unsigned* flags) {
    if (*flags) return;
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int per = fill_lanes<AXIS>(G, true);
    if (!plane_point<AXIS>(G, t / per, p, side)) return;
    if (!zg_end(G, AXIS, side)) return;
    const long long step = G.stride[AXIS];
    const long long b = node_index(G, p[0], p[1], p[2]);
    const long long s = side ? b - step : b + step;
    const long long dir = side ? step : -step;
    if (AXIS == 0) {
        const int u = (int)(t % per), v = u / (G.g[0] + 1), l = u % (G.g[0] + 1);
        q[v * G.P + b + l * dir] = q[v * G.P + s];   // l = 0: the boundary node
        return;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double val = q[v * G.P + s];
        q[v * G.P + b] = val;
        for (int l = 1; l <= G.g[AXIS]; ++l) q[v * G.P + b + l * dir] = val;
    }
}

// zeroGradient pressure -> energy (SF_thermodynamicClosure.cpp:219-281).
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_energy(GridDev G, double* __restrict__ q,
                                                   Fault* fault, long long tag) {
    if (fault->flags) return;
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const int per = fill_lanes<AXIS>(G, false);
    if (!plane_point<AXIS>(G, t / per, p, side)) return;
    if (!zg_end(G, AXIS, side)) return;
    const long long step = G.stride[AXIS];
    const long long b = node_index(G, p[0], p[1], p[2]);
    const long long s = side ? b - step : b + step;
    const long long dir = side ? step : -step;
    const double ps = pressure_of(q, G.P, s, G.gamma);
    if (AXIS == 0) {
        // each lane of the point forms the boundary node's energy and
        // pressure from the same values; lane 0 stores it, lane l >= 1
        // fills ghost l
        const int u = (int)(t % per);
        const double eb = energy_from_pressure(q, G.P, b, G.gamma, ps);
        double qb[5] = {q[b], q[G.P + b], q[2 * G.P + b], q[3 * G.P + b], eb};
        const double pb = close_state(qb, G.gamma).p;
        if (u == 0) {
            q[4 * G.P + b] = eb;
            if (!(ps > 0.0) || !(pb > 0.0)) raise_fault(kFlagState, &fault->flags, &fault->tag, tag);
        } else {
            q[4 * G.P + b + u * dir] = energy_from_pressure(q, G.P, b + u * dir, G.gamma, pb);
        }
        return;
    }
    q[4 * G.P + b] = energy_from_pressure(q, G.P, b, G.gamma, ps);
    const double pb = pressure_of(q, G.P, b, G.gamma);
    if (!(ps > 0.0) || !(pb > 0.0)) raise_fault(kFlagState, &fault->flags, &fault->tag, tag);
    for (int l = 1; l <= G.g[AXIS]; ++l)
        q[4 * G.P + b + l * dir] = energy_from_pressure(q, G.P, b + l * dir, G.gamma, pb);
}

// An end without a condition: its ghost layers take the end node's state
// of src (the state just uploaded) in dst, once; nothing refills them.
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_hold(GridDev G, const double* __restrict__ src,
                                                 double* __restrict__ dst) {
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!plane_point<AXIS>(G, t, p, side)) return;
    if (!G.active[AXIS] || G.bc[AXIS] < 2 || zg_end(G, AXIS, side)) return;
    const long long step = G.stride[AXIS];
    const long long b = node_index(G, p[0], p[1], p[2]);
    const long long dir = side ? step : -step;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const double val = src[v * G.P + b];
        for (int l = 1; l <= G.g[AXIS]; ++l) dst[v * G.P + b + l * dir] = val;
    }
}

// Periodic: ghost layer l on either side takes the node (index mod n).
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_periodic(GridDev G, double* __restrict__ q,
                                                     const unsigned* flags) {
    if (*flags) return;
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!plane_point<AXIS>(G, t, p, side)) return;
    const long long step = G.stride[AXIS];
    const int n = G.n[AXIS];
    p[AXIS] = 0;
    const long long base = node_index(G, p[0], p[1], p[2]);  // node 0 on the axis
    for (int l = 1; l <= G.g[AXIS]; ++l) {
        // low side: ghost -l <- node n-l ; high side: ghost n-1+l <- node l-1
        const long long dst = side ? base + (long long)(n - 1 + l) * step : base - l * step;
        const long long src = side ? base + (long long)((l - 1) % n) * step
                                   : base + (long long)(((n - l) % n + n) % n) * step;
#pragma unroll
        for (int v = 0; v < 5; ++v) q[v * G.P + dst] = q[v * G.P + src];
    }
}

// Clear a fault record (a kernel rather than a memset, so that the clear
// can also be captured in a graph).
__global__ void k_clear_fault(Fault* fault) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        fault->flags = 0u;
        fault->pad = 0u;
        fault->tag = -1;
    }
}

// Copy between the caller's unique-node SoA and the padded layout.
__global__ __launch_bounds__(256) void k_pack(GridDev G, const double* __restrict__ src,
                                              double* __restrict__ dst, int to_padded) {
    const long long N = (long long)G.n[0] * G.n[1] * G.n[2];
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < N;
         u += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(u % G.n[0]);
        const int j = (int)((u / G.n[0]) % G.n[1]);
        const int k = (int)(u / ((long long)G.n[0] * G.n[1]));
        const long long id = node_index(G, i, j, k);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (to_padded) dst[v * G.P + id] = src[v * N + u];
            else dst[v * N + u] = src[v * G.P + id];
        }
    }
}

}  // namespace sfamd
