// sf_kernels.h -- gfx950 kernels of the explicit Euler hot path (included by sf_capi.hip).
//
// Data layout in HBM: conserved state as structure-of-arrays FP64,
// q[v*P + (k*m1 + j)*m0 + i], with a 3-deep ghost layer on every active
// axis (m_d = n_d + 6) and none on an inactive one.  Work per RK stage:
//
//   ghost fill (O(surface))      k_bc_copy4 / k_bc_energy / k_bc_periodic
//   x sweep                      k_sweep_x:     one lane per x face, the
//                                               flux difference via LDS
//   y / z sweeps                 k_sweep_march: one lane per (x, column
//                                               segment), marching along the
//                                               axis with the 6-node stencil
//                                               in registers, so each node is
//                                               loaded once and each face
//                                               evaluated once
//   the last sweep (x) also applies the stage update, validates the new
//   state and (on the last stage) folds the CFL scan of the next step into
//   a wave-level minimum plus one 64-bit atomic per wave.
//
// Sweeps accumulate S = sum_d (F_d(i) - F_d(i-1)) and the update multiplies
// by Jac once, as Math::localSpatialResidual does
// (SF_residualAssembly.h:12-32).
#pragma once

#include "sf_device.h"

namespace sfamd {

// ---------------------------------------------------------------------------
// x sweep over the faces of all x lines numbered as one sequence: face
// f = line * (n0 + 1) + (a + 1) lies between nodes a and a+1 of its line
// (a = -1 .. n0-1).  Tile T is faces T*127 .. T*127+127: lane t evaluates
// face T*127+t and lanes 1..127 form the difference for the node whose upper
// face is theirs.  Tiles overlap by one face, so every node is updated by
// exactly one lane, and no tile is left mostly idle at the end of a line.
// Blocks loop over tiles (stride gridDim.x), so the CFL fold costs one
// atomic per block.
// ---------------------------------------------------------------------------
constexpr int kXBlock = 128;

// The final HLLC sweeps fit four waves per SIMD (128 VGPRs) without
// spilling when asked to; the others stay at three (the final Rusanov sweeps
// would spill at four).
template <int SCHEME, int FLUX, int MODE>
__global__ __launch_bounds__(kXBlock)
__attribute__((amdgpu_waves_per_eu((MODE == kFinal && FLUX == 1) ? 4 : 3, 8))) void k_sweep_x(GridDev G, const double* __restrict__ q,
                                                     double* __restrict__ S, StageArgs A) {
    __shared__ double sh[5][kXBlock];
    __shared__ unsigned long long red[kXBlock / 64];
    // face numbers fit 32 bits: the host rejects grids with 2^32 - 2^8
    // faces or more along x
    const int t = threadIdx.x;
    const unsigned per_line = G.n[0] + 1;
    const unsigned faces = per_line * G.n[1] * G.n[2];
    const unsigned tiles = (faces - 1 + kXBlock - 2) / (kXBlock - 1);
    unsigned flags = 0;
    double dt_local = INFINITY;
    for (unsigned T = blockIdx.x; T < tiles; T += gridDim.x) {
        const unsigned f = T * (kXBlock - 1) + t;
        const bool face = f < faces;
        const unsigned line = face ? f / per_line : 0;
        const int a = face ? (int)(f - line * per_line) - 1 : 0;  // lower node of the face
        const int j = (int)(line % (unsigned)G.n[1]), k = (int)(line / (unsigned)G.n[1]);
        if (face) {
            const long long base = node_index(G, a, j, k);
            double st[6][5];
#pragma unroll
            for (int s = 0; s < 6; ++s)
#pragma unroll
                for (int v = 0; v < 5; ++v) st[s][v] = q[v * G.P + base + (s - 2)];
            double f5[5];
            face_flux<0, SCHEME, FLUX>(st, G.gamma, G.area[0], f5, flags);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                if (!isfinite(f5[v])) flags |= kFlagFlux;
                sh[v][t] = f5[v];
            }
        }
        __syncthreads();
        if (t >= 1 && face && a >= 0) {
            const long long id = node_index(G, a, j, k);
            double s5[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double d = sh[v][t] - sh[v][t - 1];
                s5[v] = (MODE == kCopyS) ? d : S[v * G.P + id] + d;
            }
            if (MODE == kFinal) {
                double dl = INFINITY;
                if (solved_node(G, a, j, k)) stage_update(G, A, id, q, s5, flags, dl);
                else carry_node(G, A, id, q);
                dt_local = fmin(dt_local, dl);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) S[v * G.P + id] = s5[v];
            }
        }
        __syncthreads();  // sh is rewritten by the next tile
    }
    raise_flags(flags, A.flags);
    if (MODE == kFinal && A.reduce_dt) block_min_atomic(dt_local, A.dtmin, red);
}

// ---------------------------------------------------------------------------
// y / z sweep: lane = x index, blockIdx.y = the remaining transverse index,
// blockIdx.z = segment of SEG nodes along the sweep axis.  Each lane marches
// along its column; the six-node stencil window lives in LDS (one private
// slot set per lane, a ring indexed by `head`), so between faces only the
// previous face flux stays in registers.  Lanes never read each other's
// slots, so no barrier is needed.
// ---------------------------------------------------------------------------
constexpr int kMarchBlock = 64;

template <int AXIS, int SCHEME, int FLUX, int MODE>
__global__ __launch_bounds__(kMarchBlock) void k_sweep_march(GridDev G,
                                                             const double* __restrict__ q,
                                                             double* __restrict__ S, StageArgs A,
                                                             int seg) {
    __shared__ double win[6][5][kMarchBlock];
    const int lane = threadIdx.x;
    const int i = blockIdx.x * kMarchBlock + lane;
    const int o = blockIdx.y;
    const int a0 = blockIdx.z * seg;
    const int a1 = min(a0 + seg, G.n[AXIS]);
    unsigned flags = 0;
    double dt_local = INFINITY;
    if (i < G.n[0]) {
        const int j0 = (AXIS == 1) ? a0 : o;
        const int k0 = (AXIS == 2) ? a0 : o;
        const long long step = G.stride[AXIS];
        // Preload nodes a0-3 .. a0+1 into slots 1..5; the march starts one
        // node early (a = a0-1) so that face a0-1/2 is evaluated by the same
        // instructions as every other face: a face's rounding then does not
        // depend on where a segment starts.
        long long id = node_index(G, i, j0, k0) - 3 * step;
        for (int s = 0; s < 5; ++s)
#pragma unroll
            for (int v = 0; v < 5; ++v) win[s + 1][v][lane] = q[v * G.P + id + s * step];
        id += 2 * step;  // node a0 - 1
        int head = 0;    // slot that receives the next node
        double fprev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma clang loop unroll(disable)
        for (int a = a0 - 1; a < a1; ++a, id += step) {
            // drop the oldest node, append node a+3
#pragma unroll
            for (int v = 0; v < 5; ++v) win[head][v][lane] = q[v * G.P + id + 3 * step];
            head = (head == 5) ? 0 : head + 1;
            double st[6][5];
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const int slot = (head + s >= 6) ? head + s - 6 : head + s;
#pragma unroll
                for (int v = 0; v < 5; ++v) st[s][v] = win[slot][v][lane];
            }
            double f[5];
            face_flux<AXIS, SCHEME, FLUX>(st, G.gamma, G.area[AXIS], f, flags);
            double s5[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                if (!isfinite(f[v])) flags |= kFlagFlux;
                s5[v] = f[v] - fprev[v];
                fprev[v] = f[v];
            }
            if (a < a0) continue;
#pragma unroll
            for (int v = 0; v < 5; ++v)
                if (MODE != kCopyS) s5[v] = S[v * G.P + id] + s5[v];
            if (MODE == kFinal) {
                const int jj = (AXIS == 1) ? a : o;
                const int kk = (AXIS == 2) ? a : o;
                double dl = INFINITY;
                if (solved_node(G, i, jj, kk)) stage_update(G, A, id, q, s5, flags, dl);
                else carry_node(G, A, id, q);
                dt_local = fmin(dt_local, dl);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) S[v * G.P + id] = s5[v];
            }
        }
    }
    raise_flags(flags, A.flags);
    if (MODE == kFinal && A.reduce_dt) wave_min_atomic(dt_local, A.dtmin);
}

// ---------------------------------------------------------------------------
// Standalone CFL scan (first step, or sf_compute_dt).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cfl(GridDev G, const double* __restrict__ q, double cfl,
                                             unsigned long long* dtmin, unsigned* flags_out) {
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
    raise_flags(flags, flags_out);
    wave_min_atomic(best, dtmin);
}

__global__ void k_finalize_dt(Clock* c, double* dt_hist, int slot) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

__global__ void k_commit_clock(Clock* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
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

// zeroGradient density + momentum (SF_zeroGradient.cpp:88-130).
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_copy4(GridDev G, double* __restrict__ q) {
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!plane_point<AXIS>(G, t, p, side)) return;
    const long long step = G.stride[AXIS];
    const long long b = node_index(G, p[0], p[1], p[2]);
    const long long s = side ? b - step : b + step;
    const long long dir = side ? step : -step;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double val = q[v * G.P + s];
        q[v * G.P + b] = val;
#pragma unroll
        for (int l = 1; l <= 3; ++l) q[v * G.P + b + l * dir] = val;
    }
}

// zeroGradient pressure -> energy (SF_thermodynamicClosure.cpp:219-281).
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_energy(GridDev G, double* __restrict__ q,
                                                   unsigned* flags_out) {
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!plane_point<AXIS>(G, t, p, side)) return;
    const long long step = G.stride[AXIS];
    const long long b = node_index(G, p[0], p[1], p[2]);
    const long long s = side ? b - step : b + step;
    const long long dir = side ? step : -step;
    const double ps = pressure_of(q, G.P, s, G.gamma);
    q[4 * G.P + b] = energy_from_pressure(q, G.P, b, G.gamma, ps);
    const double pb = pressure_of(q, G.P, b, G.gamma);
    if (!(ps > 0.0) || !(pb > 0.0)) atomicOr(flags_out, (unsigned)kFlagState);
#pragma unroll
    for (int l = 1; l <= 3; ++l)
        q[4 * G.P + b + l * dir] = energy_from_pressure(q, G.P, b + l * dir, G.gamma, pb);
}

// Periodic: ghost layer l on either side takes the node (index mod n).
template <int AXIS>
__global__ __launch_bounds__(256) void k_bc_periodic(GridDev G, double* __restrict__ q) {
    int p[3], side;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (!plane_point<AXIS>(G, t, p, side)) return;
    const long long step = G.stride[AXIS];
    const int n = G.n[AXIS];
    p[AXIS] = 0;
    const long long base = node_index(G, p[0], p[1], p[2]);  // node 0 on the axis
#pragma unroll
    for (int l = 1; l <= 3; ++l) {
        // low side: ghost -l <- node n-l ; high side: ghost n-1+l <- node l-1
        const long long dst = side ? base + (long long)(n - 1 + l) * step : base - l * step;
        const long long src = side ? base + (long long)((l - 1) % n) * step
                                   : base + (long long)(((n - l) % n + n) % n) * step;
#pragma unroll
        for (int v = 0; v < 5; ++v) q[v * G.P + dst] = q[v * G.P + src];
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
