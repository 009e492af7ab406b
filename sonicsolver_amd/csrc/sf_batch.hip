// sf_batch.hip -- batched solves of independent 1D / 2D problems of uneven
// size.  One device arena holds every admitted problem's padded SoA state;
// persistent kernels walk a tile table, so one launch per sweep advances all
// problems by one RK stage; each problem keeps its own CFL dt and clock, and
// its result does not depend on which other problems share the batch (no
// cross-problem arithmetic; the per-problem dt is a minimum, which does not
// depend on the order of the atomics).  The launch sequence of a step is
// captured in a hipGraph and replayed while the set of problems is
// unchanged.
//
// Boundaries are transmissive (the reference's zeroGradient sets, applied
// axis by axis as Boundary::Applicator::apply does,
// solver/boundary/SF_applicator.h:51-83); the time recipe is SSPRK3
// (SF_builtinProviders.cpp:98-121).  Layout of problem p in the arena:
// q[v*P + off_p + (j+gy)*m0 + i+3] with m0 = nx + 6, gy = 3 for a 2D
// problem and 0 for a 1D one.
#include "sonic_amd.h"
#include "sf_batch_copy.h"
#include "sf_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace sfamd {
void set_last_error(const std::string& message);

struct BatchItemDev {
    long long off;      // start of the padded block in the arena
    int n0, n1, m0, gy;
    int scheme, flux;
    double jac, area0, area1, metric0, metric1, cfl;
};

struct BatchDev {
    const BatchItemDev* items;
    long long P;        // arena capacity in padded nodes (per variable)
    double gamma;
    const int4* xtiles;  // (slot, first face f0, -, -)
    int nxt;
    const int4* ytiles;  // (slot, first (segment, column) index, -, -)
    int nyt;
    int yseg;            // nodes per y tile
    const int* active;   // admitted slots
    int nactive;
    double* dt;
    unsigned long long* dtmin;
    double* time;
    long long* steps;
    const long long* target;  // a problem is live while steps < target ...
    unsigned* flags;          // ... and raised no fault flags (per problem)
    int* live;                // liveness fixed for the current step
};

struct BatchStage {
    const double* q0;
    double* qout;
    double base, incr;
    int stage;
    int reduce_dt;
};

// per-problem view with the fields local_dt / stage_update / carry_node read
struct ItemView {
    long long P;
    double jac, gamma;
    int active[3];
    double metric[3];
};

__device__ __forceinline__ ItemView item_view(const BatchDev& B, const BatchItemDev& it) {
    ItemView v;
    v.P = B.P;
    v.jac = it.jac;
    v.gamma = B.gamma;
    v.active[0] = 1;
    v.active[1] = it.n1 > 1;
    v.active[2] = 0;
    v.metric[0] = it.metric0;
    v.metric[1] = it.metric1;
    v.metric[2] = 1.0;
    return v;
}

__device__ __forceinline__ bool item_live(const BatchDev& B, int slot) {
    return B.steps[slot] < B.target[slot] && B.flags[slot] == 0;
}

// Liveness for the kernels of one step: fixed by k_batch_finalize at the
// start of the step, so every wave of every kernel of the step takes the
// same decision for a problem, even while another block raises its flags
// (a multi-wave x tile must agree, since its waves meet at barriers).  A
// problem that faults during a step finishes that step's stages on its own
// nodes, is not committed, and is frozen from the next step on.
__device__ __forceinline__ bool step_live(const BatchDev& B, int slot) {
    return B.live[slot] != 0;
}

__device__ __forceinline__ long long item_index(const BatchItemDev& it, int i, int j) {
    return it.off + (long long)(j + it.gy) * it.m0 + (i + 3);
}

__device__ __forceinline__ bool item_solved(const BatchItemDev& it, int i, int j) {
    if (i == 0 || i == it.n0 - 1) return false;
    if (it.n1 > 1 && (j == 0 || j == it.n1 - 1)) return false;
    return true;
}

// x faces + SSPRK3 stage update over tiles [t0, t1) of one (scheme, flux)
// variant.  A problem's x faces are numbered line by line, f = j*(nx+1) +
// (i+1) for face i+1/2, i = -1..nx-1, and a tile is BS consecutive faces
// starting at f0 (tiles overlap by one face): lane t evaluates face f0+t,
// and lanes 1..BS-1 whose face is not the first of its line difference it
// with their left neighbour's face through LDS and update that node.  Short
// lines therefore share a block and long ones span several.  Block b0 of nb
// takes tiles t0+b0, t0+b0+nb, ... (persistent).
__device__ __forceinline__ void batch_operands(const BatchDev& B, const BatchItemDev& it,
                                               const BatchStage& st,
                                               const double* __restrict__ q,
                                               const double* __restrict__ S, long long id,
                                               double (&pq)[5], double (&ps)[5],
                                               double (&pq0)[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        pq[v] = q[v * B.P + id];
        ps[v] = it.n1 > 1 ? S[v * B.P + id] : 0.0;
        pq0[v] = st.stage > 0 ? st.q0[v * B.P + id] : 0.0;
    }
}

template <int BS, int SCHEME, int FLUX, bool EARLY>
__device__ __forceinline__ void x_tiles(const BatchDev& B, const double* __restrict__ q,
                                        const double* __restrict__ S, const BatchStage& st,
                                        int t0, int t1, int b0, int nb, double (*sh)[BS],
                                        unsigned long long* red) {
    const int t = threadIdx.x;
    unsigned flags = 0;
    for (int tile = t0 + b0; tile < t1; tile += nb) {
        const int4 T = B.xtiles[tile];
        if (!step_live(B, T.x)) continue;  // uniform over the block
        const BatchItemDev it = B.items[T.x];
        // the problem's dt, fixed during a stage: read before the face
        // arithmetic, whose duration hides the load (after the tile's
        // stores, which might alias it, it would be read again)
        const double dt_item = B.dt[T.x];
        const int per_line = it.n0 + 1;
        const int f = T.y + t;
        const bool face = f < per_line * it.n1;
        const int j = face ? f / per_line : 0;
        const int a = f - j * per_line - 1;  // lower node of the face
        // The stage operands of the node this lane updates.  With EARLY they
        // are loaded before the face arithmetic, so that their latency hides
        // behind it.  For WENO5 that costs the registers of a fourth wave
        // per SIMD, which pays where 2D problems read S and not for 1D-only
        // batches; TENO5 runs at three waves either way and always loads
        // early.
        const bool upd = t >= 1 && face && a >= 0;
        double pq[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ps[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double pq0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (EARLY && upd) batch_operands(B, it, st, q, S, item_index(it, a, j), pq, ps, pq0);
        if (face) {
            const long long base = item_index(it, a, j);
            double stc[6][5];
#pragma unroll
            for (int s = 0; s < 6; ++s)
#pragma unroll
                for (int v = 0; v < 5; ++v) stc[s][v] = q[v * B.P + base + (s - 2)];
            double f5[5];
            face_flux<0, SCHEME, FLUX>(stc, B.gamma, it.area0, f5, flags);
            if (!finite_sum5(f5)) flags |= kFlagFlux;
#pragma unroll
            for (int v = 0; v < 5; ++v) sh[v][t] = f5[v];
        }
        __syncthreads();
        double dl = INFINITY;
        if (upd) {
            const long long id = item_index(it, a, j);
            if (!EARLY) batch_operands(B, it, st, q, S, id, pq, ps, pq0);
            double s5[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double d = sh[v][t] - sh[v][t - 1];
                s5[v] = (it.n1 > 1) ? ps[v] + d : d;
            }
            StageArgs A{};
            A.q0 = st.q0;
            A.qout = st.qout;
            A.dt = &B.dt[T.x];
            A.flags = &B.flags[T.x];
            A.base = st.base;
            A.incr = st.incr;
            A.cfl = it.cfl;
            A.time = SF_TIME_SSPRK3;
            A.stage = st.stage;
            A.reduce_dt = st.reduce_dt;
            const ItemView G = item_view(B, it);
            double qn[5];
            if (item_solved(it, a, j)) {
                stage_update_dt<SF_TIME_SSPRK3>(G, A, id, pq, pq0, s5, dt_item, flags, dl, qn);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    qn[v] = pq[v];   // carry_node
                    A.qout[v * B.P + id] = qn[v];
                }
            }
        }
        if (st.reduce_dt) block_min_atomic(dl, &B.dtmin[T.x], red);
        if (flags) {
            atomicOr(&B.flags[T.x], flags);
            flags = 0;
        }
        __syncthreads();
    }
}

// Tile ranges and block ranges of the four (scheme, flux) variants,
// variant v = 2*scheme + flux.
struct Split {
    int tile[5];
    int block[5];
};

// One launch for every variant: block ranges are assigned to variants.
template <int BS>
__global__ __launch_bounds__(BS) void k_batch_x_fused(BatchDev B, const double* __restrict__ q,
                                                      const double* __restrict__ S,
                                                      BatchStage st, Split sp) {
    __shared__ double sh[5][BS];
    __shared__ unsigned long long red[BS / 64];
    const int bid = blockIdx.x;
    int v = 0;
    while (v < 3 && bid >= sp.block[v + 1]) ++v;
    const int b0 = bid - sp.block[v], nb = sp.block[v + 1] - sp.block[v];
    switch (v) {
        case 0: x_tiles<BS, 0, 0, true>(B, q, S, st, sp.tile[0], sp.tile[1], b0, nb, sh, red); break;
        case 1: x_tiles<BS, 0, 1, true>(B, q, S, st, sp.tile[1], sp.tile[2], b0, nb, sh, red); break;
        case 2: x_tiles<BS, 1, 0, true>(B, q, S, st, sp.tile[2], sp.tile[3], b0, nb, sh, red); break;
        default: x_tiles<BS, 1, 1, true>(B, q, S, st, sp.tile[3], sp.tile[4], b0, nb, sh, red); break;
    }
}

// One launch per variant (fewer registers than the fused form).  Three waves
// per SIMD fit without spilling except for TENO5 + HLLC, which keeps two.
template <int BS, int SCHEME, int FLUX, bool EARLY>
__global__ __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu((SCHEME == 1 && FLUX == 1) ? 2 : 3, 8))) void k_batch_x(BatchDev B, const double* __restrict__ q,
                                                const double* __restrict__ S, BatchStage st,
                                                int t0, int t1) {
    __shared__ double sh[5][BS];
    __shared__ unsigned long long red[BS / 64];
    x_tiles<BS, SCHEME, FLUX, EARLY>(B, q, S, st, t0, t1, blockIdx.x, gridDim.x, sh, red);
}

// y faces of the 2D problems over tiles [t0, t1) of one variant.  A
// problem's y work is numbered (segment, column) = k -> (k / nx, k % nx),
// segments of yseg nodes; a tile is 64 consecutive k, one per lane, so
// narrow problems fill whole waves.  Each lane marches up its column
// segment with the stencil window in LDS and registers (one load per node)
// and writes S = F(j+1/2) - F(j-1/2).
template <int SCHEME, int FLUX>
__device__ __forceinline__ void y_tiles(const BatchDev& B, const double* __restrict__ q,
                                        double* __restrict__ S, int t0, int t1, int b0, int nb,
                                        double (*win)[5][64]) {
    const int lane = threadIdx.x;
    unsigned flags = 0;
    for (int tile = t0 + b0; tile < t1; tile += nb) {
        const int4 T = B.ytiles[tile];
        if (!step_live(B, T.x)) continue;
        const BatchItemDev it = B.items[T.x];
        const int k = T.y + lane;
        const int nseg = (it.n1 + B.yseg - 1) / B.yseg;
        if (k >= nseg * it.n0) continue;  // per lane; no barrier below
        const int seg = k / it.n0;
        const int i = k - seg * it.n0;
        const int a0 = seg * B.yseg;
        const int a1 = min(a0 + B.yseg, it.n1);
        const long long step = it.m0;
        // the face a+1/2 reads nodes a-2 .. a+1 from the lane's 4-slot LDS
        // ring, node a+2 carried in registers, node a+3 loaded for it; the
        // march starts at a0-1 so every face is evaluated by the same
        // instructions (see k_sweep_march)
        long long id = item_index(it, i, a0) - 3 * step;  // node a0-3
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int v = 0; v < 5; ++v) win[s][v][lane] = q[v * B.P + id + s * step];
        double carried[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) carried[v] = q[v * B.P + id + 4 * step];
        id += 2 * step;
        int head = 0;
        double fprev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma clang loop unroll(disable)
        for (int a = a0 - 1; a < a1; ++a, id += step) {
            double stc[6][5];
#pragma unroll
            for (int v = 0; v < 5; ++v) stc[5][v] = q[v * B.P + id + 3 * step];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int slot = (head + s) & 3;
#pragma unroll
                for (int v = 0; v < 5; ++v) stc[s][v] = win[slot][v][lane];
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                stc[4][v] = carried[v];
                win[head][v][lane] = carried[v];
                carried[v] = stc[5][v];
            }
            head = (head + 1) & 3;
            double f[5];
            face_flux<1, SCHEME, FLUX>(stc, B.gamma, it.area1, f, flags);
            double d[5];
            if (!finite_sum5(f)) flags |= kFlagFlux;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                d[v] = f[v] - fprev[v];
                fprev[v] = f[v];
            }
            if (a < a0) continue;
#pragma unroll
            for (int v = 0; v < 5; ++v) S[v * B.P + id] = d[v];
        }
        if (flags) {
            atomicOr(&B.flags[T.x], flags);
            flags = 0;
        }
    }
}

__global__ __launch_bounds__(64) void k_batch_y_fused(BatchDev B, const double* __restrict__ q,
                                                      double* __restrict__ S, Split sp) {
    __shared__ double win[4][5][64];
    const int bid = blockIdx.x;
    int v = 0;
    while (v < 3 && bid >= sp.block[v + 1]) ++v;
    const int b0 = bid - sp.block[v], nb = sp.block[v + 1] - sp.block[v];
    switch (v) {
        case 0: y_tiles<0, 0>(B, q, S, sp.tile[0], sp.tile[1], b0, nb, win); break;
        case 1: y_tiles<0, 1>(B, q, S, sp.tile[1], sp.tile[2], b0, nb, win); break;
        case 2: y_tiles<1, 0>(B, q, S, sp.tile[2], sp.tile[3], b0, nb, win); break;
        default: y_tiles<1, 1>(B, q, S, sp.tile[3], sp.tile[4], b0, nb, win); break;
    }
}

// The WENO5 variants fit four waves per SIMD (128 VGPRs) when asked to.
template <int SCHEME, int FLUX>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu((SCHEME == 0) ? 4 : 3, 8))) void k_batch_y(BatchDev B, const double* __restrict__ q,
                                                double* __restrict__ S, int t0, int t1) {
    __shared__ double win[4][5][64];
    y_tiles<SCHEME, FLUX>(B, q, S, t0, t1, blockIdx.x, gridDim.x, win);
}

// Transmissive fill along AXIS of every live problem (blockIdx.x = the
// problem, blockIdx.y = a share of its boundary points, so a large problem's
// accesses are spread over several blocks).
// PHASE 0 copies density and momentum from the inner neighbour into the
// boundary node and its ghosts (SF_zeroGradient.cpp:88-130); PHASE 1
// rebuilds energy from the neighbour's pressure
// (SF_thermodynamicClosure.cpp:219-281).  The phases run x then y each, as
// the grid fill does.
//
// Along y a boundary row is contiguous, so a thread per row point writes
// coalesced.  Along x the boundary node and its ghosts of one row are four
// adjacent doubles per variable while rows lie a row pitch apart, so there
// the work is split finer: PHASE 0 gives each (row, variable, node) a
// thread, PHASE 1 each (row, node) -- every thread of a row forms the
// boundary node's energy and pressure itself, from the same values, so the
// ghosts take exactly the pressure the boundary node holds -- and
// neighbouring lanes write neighbouring doubles.
template <int AXIS, int PHASE>
__global__ __launch_bounds__(256) void k_batch_fill(BatchDev B, double* __restrict__ q) {
    const int slot = B.active[blockIdx.x];
    if (!step_live(B, slot)) return;
    const BatchItemDev it = B.items[slot];
    if (AXIS == 1 && it.n1 <= 1) return;
    if (AXIS == 0) {
        constexpr int per = PHASE == 0 ? 16 : 4;   // threads per boundary row
        const int nrow = it.n1;
        for (int t = blockIdx.y * blockDim.x + threadIdx.x; t < 2 * nrow * per;
             t += gridDim.y * blockDim.x) {
            const int pr = t / per, u = t % per;
            const int side = pr / nrow, r = pr % nrow;
            const long long b = item_index(it, side ? it.n0 - 1 : 0, r);
            const long long s = side ? b - 1 : b + 1;
            const long long dir = side ? 1 : -1;
            if (PHASE == 0) {
                const int v = u >> 2, l = u & 3;   // l = 0: the boundary node
                q[v * B.P + b + l * dir] = q[v * B.P + s];
            } else {
                // every thread of the row forms the boundary node's energy
                // and pressure from the same values; thread 0 stores the
                // energy, thread l >= 1 fills ghost l
                const double ps = pressure_of(q, B.P, s, B.gamma);
                const double eb = energy_from_pressure(q, B.P, b, B.gamma, ps);
                double qb[5] = {q[b], q[B.P + b], q[2 * B.P + b], q[3 * B.P + b], eb};
                const double pb = close_state(qb, B.gamma).p;
                if (u == 0) {
                    q[4 * B.P + b] = eb;
                    if (!(ps > 0.0) || !(pb > 0.0)) atomicOr(&B.flags[slot], (unsigned)kFlagState);
                } else {
                    q[4 * B.P + b + u * dir] = energy_from_pressure(q, B.P, b + u * dir, B.gamma, pb);
                }
            }
        }
        return;
    }
    const int nother = AXIS == 0 ? it.n1 : it.n0;
    const int nax = AXIS == 0 ? it.n0 : it.n1;
    const long long step = AXIS == 0 ? 1 : it.m0;
    for (int t = blockIdx.y * blockDim.x + threadIdx.x; t < 2 * nother;
         t += gridDim.y * blockDim.x) {
        const int side = t / nother, r = t % nother;
        const int bnode = side ? nax - 1 : 0;
        const long long b = AXIS == 0 ? item_index(it, bnode, r) : item_index(it, r, bnode);
        const long long s = side ? b - step : b + step;
        const long long dir = side ? step : -step;
        if (PHASE == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double val = q[v * B.P + s];
                q[v * B.P + b] = val;
#pragma unroll
                for (int l = 1; l <= 3; ++l) q[v * B.P + b + l * dir] = val;
            }
        } else {
            const double ps = pressure_of(q, B.P, s, B.gamma);
            q[4 * B.P + b] = energy_from_pressure(q, B.P, b, B.gamma, ps);
            const double pb = pressure_of(q, B.P, b, B.gamma);
            if (!(ps > 0.0) || !(pb > 0.0)) atomicOr(&B.flags[slot], (unsigned)kFlagState);
#pragma unroll
            for (int l = 1; l <= 3; ++l)
                q[4 * B.P + b + l * dir] = energy_from_pressure(q, B.P, b + l * dir, B.gamma, pb);
        }
    }
}

// CFL scan of the listed problems (blockIdx.x = list entry, blockIdx.y
// splits the nodes) folded into dtmin[slot] (RusanovEOS::deltaT,
// SF_rusanovEOS.cpp:110-147).
__global__ __launch_bounds__(256) void k_batch_cfl(BatchDev B, const int* __restrict__ list,
                                                   const double* __restrict__ q) {
    __shared__ unsigned long long red[4];
    const int slot = list[blockIdx.x];
    const BatchItemDev it = B.items[slot];
    const ItemView G = item_view(B, it);
    const long long N = (long long)it.n0 * it.n1;
    unsigned flags = 0;
    double best = INFINITY;
    for (long long u = blockIdx.y * (long long)blockDim.x + threadIdx.x; u < N;
         u += (long long)gridDim.y * blockDim.x) {
        const int i = (int)(u % it.n0), j = (int)(u / it.n0);
        if (!item_solved(it, i, j)) continue;
        const long long id = item_index(it, i, j);
        double qq[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) qq[v] = q[v * B.P + id];
        best = fmin(best, local_dt(G, qq, it.cfl, flags));
    }
    if (flags) atomicOr(&B.flags[slot], flags);
    block_min_atomic(best, &B.dtmin[slot], red);
}

// dt of the step from the folded CFL minimum; the accumulator is reset for
// the fold of this step's last stage.
__global__ void k_batch_finalize(BatchDev B) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B.nactive) return;
    const int slot = B.active[k];
    const bool live = item_live(B, slot);
    B.live[slot] = live ? 1 : 0;
    if (!live) return;
    B.dt[slot] = __longlong_as_double((long long)B.dtmin[slot]);
    B.dtmin[slot] = 0x7FF0000000000000ull;
}

__global__ void k_batch_commit(BatchDev B) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B.nactive) return;
    const int slot = B.active[k];
    if (!step_live(B, slot) || B.flags[slot] != 0) return;
    B.time[slot] += B.dt[slot];
    B.steps[slot] += 1;
}

__global__ void k_batch_reset(BatchDev B, const int* __restrict__ list, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int slot = list[k];
    B.dtmin[slot] = 0x7FF0000000000000ull;
    B.time[slot] = 0.0;
    B.steps[slot] = 0;
    B.dt[slot] = 0.0;
    B.flags[slot] = 0;
    B.live[slot] = 1;   // for the fill that prepares the problem
}

// unique-node SoA of one problem <-> its padded block in the arena
__global__ __launch_bounds__(256) void k_batch_pack(BatchDev B, int slot,
                                                    const double* __restrict__ src,
                                                    double* __restrict__ dst, int to_arena) {
    const BatchItemDev it = B.items[slot];
    const long long N = (long long)it.n0 * it.n1;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < N;
         u += (long long)gridDim.x * blockDim.x) {
        const long long id = item_index(it, (int)(u % it.n0), (int)(u / it.n0));
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (to_arena) dst[v * B.P + id] = src[v * N + u];
            else dst[v * N + u] = src[v * B.P + id];
        }
    }
}

}  // namespace sfamd

using namespace sfamd;

namespace {

int bfail(int code, const std::string& msg) {
    set_last_error(msg);
    return code;
}

#define SF_BHIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return bfail(SF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

long long padded_nodes(const sf_batch_item& it) {
    const long long m0 = it.n[0] + 6;
    const long long m1 = it.n[1] > 1 ? it.n[1] + 6 : 1;
    return m0 * m1;
}

}  // namespace

struct sf_batch {
    int device = 0;
    int cus = 256;
    double gamma = 1.4;
    long long capacity = 0;     // padded nodes per variable
    int item_capacity = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // copy-out of finished problems
    hipEvent_t copy_after = nullptr;     // orders copy-outs after the compute stream
    double* buf[3] = {nullptr, nullptr, nullptr};
    double* S = nullptr;
    double* staging = nullptr;
    long long staging_nodes = 0;
    BatchItemDev* d_items = nullptr;
    double* d_dt = nullptr;
    unsigned long long* d_dtmin = nullptr;
    double* d_time = nullptr;
    long long* d_steps = nullptr;
    long long* d_target = nullptr;
    unsigned* d_flags = nullptr;
    int* d_live = nullptr;
    int4* d_xt = nullptr;
    int4* d_yt = nullptr;
    int* d_active = nullptr;
    int* d_scan = nullptr;
    size_t xt_cap = 0, yt_cap = 0;
    // launch shape (tunable)
    int xbs = 128;              // x block: 64, 128 or 256 lanes (tile = xbs-1 nodes)
    int yseg = 32;              // y tile length
    int blocks_per_cu = 8;      // persistent grid size
    int fused = 1;              // 1: one launch per sweep for all variants
    // host side
    std::vector<BatchItemDev> items;
    std::vector<sf_batch_item> desc;
    std::vector<char> used;
    std::map<long long, long long> free_ranges;  // offset -> length
    std::vector<int> pending_scan;               // admitted since the last step
    bool tables_dirty = true;
    int nxt = 0, nyt = 0, nactive = 0;
    int xr[5] = {0, 0, 0, 0, 0}, yr[5] = {0, 0, 0, 0, 0};  // tile ranges per variant
    bool any2d = false;
    int max_n0 = 1, max_n1 = 1;  // largest n0 / n1 of the admitted problems
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace {

BatchDev view(const sf_batch* b) {
    BatchDev B{};
    B.items = b->d_items;
    B.P = b->capacity;
    B.gamma = b->gamma;
    B.xtiles = b->d_xt;
    B.nxt = b->nxt;
    B.ytiles = b->d_yt;
    B.nyt = b->nyt;
    B.yseg = b->yseg;
    B.active = b->d_active;
    B.nactive = b->nactive;
    B.dt = b->d_dt;
    B.dtmin = b->d_dtmin;
    B.time = b->d_time;
    B.steps = b->d_steps;
    B.target = b->d_target;
    B.flags = b->d_flags;
    B.live = b->d_live;
    return B;
}

void drop_graph(sf_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

// Tile tables and the active list after admissions / removals.  Tiles are
// grouped by (scheme, flux) variant and, within a variant, listed largest
// problem first, so the persistent blocks start on the big problems and
// finish on the small ones.
int rebuild_tables(sf_batch* b) {
    std::vector<int4> xt, yt;
    std::vector<int> act;
    b->any2d = false;
    b->max_n0 = b->max_n1 = 1;
    for (int s = 0; s < b->item_capacity; ++s)
        if (b->used[s]) {
            act.push_back(s);
            b->max_n0 = std::max(b->max_n0, b->items[s].n0);
            b->max_n1 = std::max(b->max_n1, b->items[s].n1);
        }
    std::vector<int> order = act;
    auto variant = [&](int s) { return 2 * b->items[s].scheme + b->items[s].flux; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (variant(x) != variant(y)) return variant(x) < variant(y);
        return (long long)b->items[x].n0 * b->items[x].n1 >
               (long long)b->items[y].n0 * b->items[y].n1;
    });
    const int xstep = b->xbs - 1;
    int v = 0;
    b->xr[0] = b->yr[0] = 0;
    for (int s : order) {
        while (v < variant(s)) {
            ++v;
            b->xr[v] = (int)xt.size();
            b->yr[v] = (int)yt.size();
        }
        const BatchItemDev& it = b->items[s];
        // x: tiles of xbs consecutive faces, overlapping by one
        const long long faces = (long long)(it.n0 + 1) * it.n1;
        for (long long f0 = 0; f0 < faces - 1; f0 += xstep) xt.push_back(make_int4(s, (int)f0, 0, 0));
        if (it.n1 > 1) {
            b->any2d = true;
            const long long work = (long long)((it.n1 + b->yseg - 1) / b->yseg) * it.n0;
            for (long long k0 = 0; k0 < work; k0 += 64) yt.push_back(make_int4(s, (int)k0, 0, 0));
        }
    }
    while (v < 4) {
        ++v;
        b->xr[v] = (int)xt.size();
        b->yr[v] = (int)yt.size();
    }
    auto upload = [&](int4** ptr, size_t* cap, const std::vector<int4>& vec) -> int {
        if (vec.size() > *cap) {
            if (*ptr) SF_BHIP(hipFree(*ptr));
            *ptr = nullptr;
            *cap = 0;
            const size_t want = std::max<size_t>(vec.size(), 1024);
            SF_BHIP(hipMalloc(ptr, want * sizeof(int4)));
            *cap = want;
        }
        if (!vec.empty())
            SF_BHIP(hipMemcpy(*ptr, vec.data(), vec.size() * sizeof(int4), hipMemcpyHostToDevice));
        return 0;
    };
    if (int rc = upload(&b->d_xt, &b->xt_cap, xt)) return rc;
    if (int rc = upload(&b->d_yt, &b->yt_cap, yt)) return rc;
    if (!act.empty())
        SF_BHIP(hipMemcpy(b->d_active, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice));
    b->nxt = (int)xt.size();
    b->nyt = (int)yt.size();
    b->nactive = (int)act.size();
    b->tables_dirty = false;
    drop_graph(b);
    return 0;
}

// Relative cost of a tile of each variant (TENO5 evaluates its three
// candidate weights with a sixth power; HLLC adds a star state).
constexpr double kVariantCost[4] = {1.0, 1.03, 1.16, 1.19};

// Block ranges of a fused launch of `total` blocks over the variant tile
// ranges r[5]: proportional to the tiles' cost, at least one block for a
// non-empty variant, never more blocks than tiles.
Split split_blocks(const int r[5], unsigned total) {
    Split sp{};
    double w[4], sum = 0.0;
    for (int v = 0; v < 4; ++v) {
        w[v] = (r[v + 1] - r[v]) * kVariantCost[v];
        sum += w[v];
    }
    sp.tile[0] = r[0];
    sp.block[0] = 0;
    for (int v = 0; v < 4; ++v) {
        const int n = r[v + 1] - r[v];
        int nb = 0;
        if (n > 0) nb = std::min(n, std::max(1, (int)std::lround(total * w[v] / sum)));
        sp.tile[v + 1] = r[v + 1];
        sp.block[v + 1] = sp.block[v] + nb;
    }
    return sp;
}

// Blocks of `kernel` (block size `threads`) a compute unit holds at once,
// from the runtime's occupancy calculator (cached per kernel).
template <class Kernel>
int resident_per_cu(Kernel kernel, int threads) {
    static std::map<const void*, int> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    const void* key = reinterpret_cast<const void*>(kernel);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1)
        n = 4;
    cache[key] = n;
    return n;
}

// Persistent grid of a sweep over `tiles` tiles: blocks_per_cu blocks per
// compute unit, or with blocks_per_cu = 0 as many as the device holds at
// once (`per_cu` = the kernel's resident blocks per unit).
unsigned persistent_blocks(const sf_batch* b, int tiles, int per_cu = 0) {
    const int want = b->blocks_per_cu > 0 ? b->blocks_per_cu : std::max(per_cu, 1);
    return (unsigned)std::max(1, std::min(tiles, b->cus * want));
}

// ghost fill of every live problem: copy x, copy y, energy x, energy y;
// blocks per problem sized so each thread takes about four work items of
// the largest problem (at most 32)
void enqueue_fill(sf_batch* b, double* q) {
    const BatchDev B = view(b);
    auto grid = [&](long long items) {
        return dim3((unsigned)b->nactive,
                    (unsigned)std::min<long long>(32, std::max<long long>(1, (items + 1023) / 1024)));
    };
    k_batch_fill<0, 0><<<grid(32LL * b->max_n1), 256, 0, b->stream>>>(B, q);
    if (b->any2d) k_batch_fill<1, 0><<<grid(2LL * b->max_n0), 256, 0, b->stream>>>(B, q);
    k_batch_fill<0, 1><<<grid(8LL * b->max_n1), 256, 0, b->stream>>>(B, q);
    if (b->any2d) k_batch_fill<1, 1><<<grid(2LL * b->max_n0), 256, 0, b->stream>>>(B, q);
}

void launch_x(sf_batch* b, const double* q, BatchStage st) {
    const BatchDev B = view(b);
    if (b->fused) {
#define SF_XF(BS)                                                                            \
    do {                                                                                     \
        const Split sp = split_blocks(                                                       \
            b->xr, persistent_blocks(b, b->nxt, resident_per_cu(k_batch_x_fused<BS>, BS)));  \
        k_batch_x_fused<BS><<<(unsigned)sp.block[4], BS, 0, b->stream>>>(B, q, b->S, st, sp); \
    } while (0)
        switch (b->xbs) {
            case 64: SF_XF(64); break;
            case 256: SF_XF(256); break;
            default: SF_XF(128); break;
        }
#undef SF_XF
        return;
    }
    for (int v = 0; v < 4; ++v) {
        const int t0 = b->xr[v], t1 = b->xr[v + 1];
        if (t1 == t0) continue;
#define SF_XV1(BS, SC, FL, W)                                                              \
    k_batch_x<BS, SC, FL, W>                                                               \
        <<<persistent_blocks(b, t1 - t0, resident_per_cu(k_batch_x<BS, SC, FL, W>, BS)),  \
           BS, 0, b->stream>>>(B, q, b->S, st, t0, t1)
#define SF_XV(BS, SC, FL)                                   \
    do {                                                   \
        if (b->any2d || SC == 1) SF_XV1(BS, SC, FL, true); \
        else SF_XV1(BS, SC, FL, false);                    \
    } while (0)
#define SF_XBS(SC, FL)                        \
    switch (b->xbs) {                         \
        case 64: SF_XV(64, SC, FL); break;    \
        case 256: SF_XV(256, SC, FL); break;  \
        default: SF_XV(128, SC, FL); break;   \
    }
        switch (v) {
            case 0: SF_XBS(0, 0); break;
            case 1: SF_XBS(0, 1); break;
            case 2: SF_XBS(1, 0); break;
            default: SF_XBS(1, 1); break;
        }
#undef SF_XBS
#undef SF_XV
#undef SF_XV1
    }
}

void launch_y(sf_batch* b, const double* q) {
    const BatchDev B = view(b);
    if (b->nyt == 0) return;
    if (b->fused) {
        const Split sp = split_blocks(
            b->yr, persistent_blocks(b, b->nyt, resident_per_cu(k_batch_y_fused, 64)));
        k_batch_y_fused<<<(unsigned)sp.block[4], 64, 0, b->stream>>>(B, q, b->S, sp);
        return;
    }
    for (int v = 0; v < 4; ++v) {
        const int t0 = b->yr[v], t1 = b->yr[v + 1];
        if (t1 == t0) continue;
#define SF_YV(SC, FL)                                                                         \
    k_batch_y<SC, FL><<<persistent_blocks(b, t1 - t0, resident_per_cu(k_batch_y<SC, FL>, 64)), \
                        64, 0, b->stream>>>(B, q, b->S, t0, t1)
        switch (v) {
            case 0: SF_YV(0, 0); break;
            case 1: SF_YV(0, 1); break;
            case 2: SF_YV(1, 0); break;
            default: SF_YV(1, 1); break;
        }
#undef SF_YV
    }
}

// One SSPRK3 step of every live problem (the captured sequence):
// stage s reads buf[in[s]] and writes buf[out[s]]; the step ends in buf[0].
void enqueue_step(sf_batch* b) {
    const BatchDev B = view(b);
    const unsigned nb = (unsigned)((b->nactive + 255) / 256);
    k_batch_finalize<<<nb, 256, 0, b->stream>>>(B);
    static const double base[3] = {0.0, 0.75, 1.0 / 3.0};
    static const double incr[3] = {1.0, 0.25, 2.0 / 3.0};
    const int in[3] = {0, 1, 2}, out[3] = {1, 2, 0};
    for (int s = 0; s < 3; ++s) {
        if (s > 0) enqueue_fill(b, b->buf[in[s]]);
        launch_y(b, b->buf[in[s]]);
        BatchStage st{b->buf[0], b->buf[out[s]], base[s], incr[s], s, s == 2 ? 1 : 0};
        launch_x(b, b->buf[in[s]], st);
    }
    enqueue_fill(b, b->buf[0]);
    k_batch_commit<<<nb, 256, 0, b->stream>>>(B);
}

int check_item(const sf_batch_item* it) {
    if (!it) return bfail(SF_E_ARG, "null batch item");
    if (it->n[0] < 4) return bfail(SF_E_ARG, "a batch problem needs at least 4 nodes along x");
    if (it->n[1] != 1 && it->n[1] < 4)
        return bfail(SF_E_ARG, "a 2D batch problem needs at least 4 nodes along y");
    if (!(it->d[0] > 0.0) || (it->n[1] > 1 && !(it->d[1] > 0.0)))
        return bfail(SF_E_ARG, "node spacing must be positive");
    if (it->scheme != SF_SCHEME_WENO5 && it->scheme != SF_SCHEME_TENO5)
        return bfail(SF_E_ARG, "the batch serves the WENO5 and TENO5 reconstructions "
                               "(WENO3 and WENO7 run on the grid path)");
    if (it->flux != SF_FLUX_RUSANOV && it->flux != SF_FLUX_HLLC)
        return bfail(SF_E_ARG, "the batch serves the Rusanov and HLLC fluxes (Steger-Warming, "
                               "Lax-Friedrichs and Roe run on the grid path)");
    if (!(it->cfl > 0.0) || !std::isfinite(it->cfl)) return bfail(SF_E_ARG, "invalid CFL number");
    return 0;
}

bool valid_slot(const sf_batch* b, int slot) {
    return b && slot >= 0 && slot < b->item_capacity && b->used[slot];
}

// Fill and CFL-scan the problems admitted since the last step.
int prepare_pending(sf_batch* b) {
    if (b->pending_scan.empty()) return 0;
    const int n = (int)b->pending_scan.size();
    SF_BHIP(hipMemcpyAsync(b->d_scan, b->pending_scan.data(), sizeof(int) * n,
                           hipMemcpyHostToDevice, b->stream));
    const BatchDev B = view(b);
    k_batch_reset<<<(n + 255) / 256, 256, 0, b->stream>>>(B, b->d_scan, n);
    enqueue_fill(b, b->buf[0]);
    const dim3 grid((unsigned)n, 16);
    k_batch_cfl<<<grid, 256, 0, b->stream>>>(B, b->d_scan, b->buf[0]);
    SF_BHIP(hipGetLastError());
    SF_BHIP(hipStreamSynchronize(b->stream));
    b->pending_scan.clear();
    return 0;
}

// Queue nsteps steps (graph replay or direct launches); no wait.
// Steps from which capturing a new graph pays for itself against direct
// launches (a capture and instantiation cost about as much as a dozen
// steps' worth of kernel launches).
constexpr int kGraphCaptureSteps = 16;

int advance(sf_batch* b, int nsteps, int use_graph) {
    if (use_graph == 2) use_graph = (b->graph_exec || nsteps >= kGraphCaptureSteps) ? 1 : 0;
    if (use_graph) {
        if (!b->graph_exec) {
            SF_BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(b);
            hipGraph_t g = nullptr;
            SF_BHIP(hipStreamEndCapture(b->stream, &g));
            b->graph = g;
            SF_BHIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
        }
        for (int s = 0; s < nsteps; ++s) SF_BHIP(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
        for (int s = 0; s < nsteps; ++s) enqueue_step(b);
    }
    SF_BHIP(hipGetLastError());
    return 0;
}

// Per-problem fault flags of the admitted problems (0 for free slots);
// returns the number of faulted problems.
int read_faults(sf_batch* b, unsigned* flags, int* faulted) {
    SF_BHIP(hipMemcpy(flags, b->d_flags, sizeof(unsigned) * b->item_capacity,
                      hipMemcpyDeviceToHost));
    int n = 0;
    for (int s = 0; s < b->item_capacity; ++s) {
        if (!b->used[s]) flags[s] = 0;
        n += flags[s] != 0;
    }
    *faulted = n;
    return 0;
}

}  // namespace

extern "C" {

int sf_batch_item_nodes(const sf_batch_item* item, size_t* nodes) {
    if (int rc = check_item(item)) return rc;
    if (nodes) *nodes = (size_t)padded_nodes(*item);
    return 0;
}

int sf_batch_create(int device, double gamma, size_t node_capacity, int item_capacity,
                    sf_batch** out) {
    if (!out) return bfail(SF_E_ARG, "sf_batch_create: null output");
    *out = nullptr;
    if (!(gamma > 1.0)) return bfail(SF_E_ARG, "gamma must exceed 1");
    if (node_capacity < 64 || item_capacity < 1)
        return bfail(SF_E_ARG, "sf_batch_create: capacities must be positive");
    int ndev = 0;
    SF_BHIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return bfail(SF_E_ARG, "device ordinal out of range");
    SF_BHIP(hipSetDevice(device));
    sf_batch* b = new (std::nothrow) sf_batch;
    if (!b) return bfail(SF_E_MEMORY, "host allocation failed");
    b->device = device;
    b->gamma = gamma;
    b->capacity = (long long)node_capacity;
    b->item_capacity = item_capacity;
    if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess || b->cus <= 0)
        b->cus = 256;   // MI355X's CU count; only the launch shapes depend on it
    const size_t state = sizeof(double) * 5 * node_capacity;
    bool ok = true;
    for (int i = 0; i < 3 && ok; ++i) ok = hipMalloc(&b->buf[i], state) == hipSuccess;
    ok = ok && hipMalloc(&b->S, state) == hipSuccess;
    ok = ok && hipMalloc(&b->d_items, sizeof(BatchItemDev) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_dt, sizeof(double) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_dtmin, sizeof(unsigned long long) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_time, sizeof(double) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_steps, sizeof(long long) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_target, sizeof(long long) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_flags, sizeof(unsigned) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_live, sizeof(int) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_active, sizeof(int) * item_capacity) == hipSuccess;
    ok = ok && hipMalloc(&b->d_scan, sizeof(int) * item_capacity) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&b->copy_stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->copy_after, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipMemset(b->d_flags, 0, sizeof(unsigned) * item_capacity) == hipSuccess;
    ok = ok && hipMemset(b->d_live, 0, sizeof(int) * item_capacity) == hipSuccess;
    ok = ok && hipMemset(b->d_steps, 0, sizeof(long long) * item_capacity) == hipSuccess;
    ok = ok && hipMemset(b->d_target, 0, sizeof(long long) * item_capacity) == hipSuccess;
    // the arena starts zeroed, so padding that is never written stays finite
    for (int i = 0; i < 3 && ok; ++i) ok = hipMemset(b->buf[i], 0, state) == hipSuccess;
    ok = ok && hipMemset(b->S, 0, state) == hipSuccess;
    if (!ok) {
        sf_batch_destroy(b);
        return bfail(SF_E_MEMORY, "sf_batch_create: device allocation of " +
                                      std::to_string(4 * state) + " bytes failed");
    }
    b->items.resize(item_capacity);
    b->desc.resize(item_capacity);
    b->used.assign(item_capacity, 0);
    b->free_ranges[0] = (long long)node_capacity;
    *out = b;
    return 0;
}

int sf_batch_destroy(sf_batch* b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);
    drop_graph(b);
    for (double* p : b->buf)
        if (p) (void)hipFree(p);
    if (b->S) (void)hipFree(b->S);
    if (b->staging) (void)hipFree(b->staging);
    void* ptrs[] = {b->d_items, b->d_dt,   b->d_dtmin, b->d_time,   b->d_steps, b->d_target,
                    b->d_flags, b->d_live, b->d_xt,    b->d_yt,     b->d_active, b->d_scan};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->copy_stream) (void)hipStreamDestroy(b->copy_stream);
    if (b->copy_after) (void)hipEventDestroy(b->copy_after);
    delete b;
    return 0;
}


int sf_batch_add(sf_batch* b, const sf_batch_item* item, const double* soa, long long max_steps,
                 int* slot) {
    if (!b || !soa || !slot) return bfail(SF_E_ARG, "sf_batch_add: null argument");
    if (int rc = check_item(item)) return rc;
    if (max_steps < 0) return bfail(SF_E_ARG, "sf_batch_add: negative step limit");
    SF_BHIP(hipSetDevice(b->device));
    const long long need = padded_nodes(*item);
    auto fit = b->free_ranges.end();
    for (auto it = b->free_ranges.begin(); it != b->free_ranges.end(); ++it)
        if (it->second >= need) {
            fit = it;
            break;
        }
    int s = -1;
    for (int k = 0; k < b->item_capacity; ++k)
        if (!b->used[k]) {
            s = k;
            break;
        }
    if (s < 0) return bfail(SF_E_MEMORY, "sf_batch_add: every problem slot is taken");
    if (fit == b->free_ranges.end())
        return bfail(SF_E_MEMORY, "sf_batch_add: no room in the arena for the problem (" +
                                      std::to_string(need) + " padded nodes)");
    const long long N = (long long)item->n[0] * std::max(1, item->n[1]);
    if (N > b->staging_nodes) {
        if (b->staging) SF_BHIP(hipFree(b->staging));
        b->staging = nullptr;
        b->staging_nodes = 0;
        SF_BHIP(hipMalloc(&b->staging, sizeof(double) * 5 * (size_t)N));
        b->staging_nodes = N;
    }
    const long long off = fit->first, len = fit->second;
    BatchItemDev it{};
    it.off = off;
    it.n0 = item->n[0];
    it.n1 = std::max(1, item->n[1]);
    it.m0 = it.n0 + 6;
    it.gy = it.n1 > 1 ? 3 : 0;
    it.scheme = item->scheme;
    it.flux = item->flux;
    it.cfl = item->cfl;
    // uniform Cartesian metrics formed as for a grid with unit spacing on
    // the absent axes (SF_mesh.cpp:320-338, SF_canonicalFace.h:87-108)
    const double dx = item->d[0], dy = it.n1 > 1 ? item->d[1] : 1.0, dz = 1.0;
    it.jac = 1.0 / (dx * dy * dz);
    it.metric0 = 1.0 / dx;
    it.metric1 = 1.0 / dy;
    const double c0 = it.metric0 / it.jac, c1 = it.metric1 / it.jac;
    const double f0 = 0.5 * (c0 + c0), f1 = 0.5 * (c1 + c1);
    it.area0 = std::sqrt(f0 * f0 + 0.0 * 0.0 + 0.0 * 0.0);
    it.area1 = std::sqrt(f1 * f1 + 0.0 * 0.0 + 0.0 * 0.0);
    const long long target = max_steps > 0 ? max_steps : LLONG_MAX;
    SF_BHIP(hipMemcpy(b->d_items + s, &it, sizeof it, hipMemcpyHostToDevice));
    SF_BHIP(hipMemcpy(b->d_target + s, &target, sizeof target, hipMemcpyHostToDevice));
    SF_BHIP(hipMemcpyAsync(b->staging, soa, sizeof(double) * 5 * (size_t)N, hipMemcpyHostToDevice,
                           b->stream));
    b->items[s] = it;
    const BatchDev B = view(b);
    const unsigned nb = (unsigned)std::min<long long>((N + 255) / 256, 4096);
    k_batch_pack<<<nb, 256, 0, b->stream>>>(B, s, b->staging, b->buf[0], 1);
    SF_BHIP(hipGetLastError());
    SF_BHIP(hipStreamSynchronize(b->stream));
    b->free_ranges.erase(fit);
    if (len > need) b->free_ranges[off + need] = len - need;
    b->desc[s] = *item;
    b->used[s] = 1;
    b->pending_scan.push_back(s);
    b->tables_dirty = true;
    *slot = s;
    return 0;
}

int sf_batch_remove(sf_batch* b, int slot) {
    if (!valid_slot(b, slot)) return bfail(SF_E_ARG, "sf_batch_remove: invalid slot");
    SF_BHIP(hipSetDevice(b->device));
    SF_BHIP(hipStreamSynchronize(b->stream));
    b->used[slot] = 0;
    b->pending_scan.erase(std::remove(b->pending_scan.begin(), b->pending_scan.end(), slot),
                          b->pending_scan.end());
    long long off = b->items[slot].off, len = padded_nodes(b->desc[slot]);
    auto next = b->free_ranges.lower_bound(off);
    if (next != b->free_ranges.end() && next->first == off + len) {
        len += next->second;
        b->free_ranges.erase(next);
    }
    auto prev = b->free_ranges.lower_bound(off);
    if (prev != b->free_ranges.begin()) {
        --prev;
        if (prev->first + prev->second == off) {
            off = prev->first;
            len += prev->second;
            b->free_ranges.erase(prev);
        }
    }
    b->free_ranges[off] = len;
    b->tables_dirty = true;
    return 0;
}

int sf_batch_download(sf_batch* b, int slot, double* soa) {
    if (!soa || !valid_slot(b, slot)) return bfail(SF_E_ARG, "sf_batch_download: invalid slot");
    SF_BHIP(hipSetDevice(b->device));
    const BatchItemDev& it = b->items[slot];
    const long long N = (long long)it.n0 * it.n1;
    if (N > b->staging_nodes) return bfail(SF_E_STATE, "sf_batch_download: staging too small");
    const BatchDev B = view(b);
    const unsigned nb = (unsigned)std::min<long long>((N + 255) / 256, 4096);
    k_batch_pack<<<nb, 256, 0, b->stream>>>(B, slot, b->buf[0], b->staging, 0);
    SF_BHIP(hipGetLastError());
    SF_BHIP(hipMemcpyAsync(soa, b->staging, sizeof(double) * 5 * (size_t)N, hipMemcpyDeviceToHost,
                           b->stream));
    SF_BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

int sf_batch_step(sf_batch* b, int nsteps, int use_graph) {
    if (!b || nsteps < 0 || use_graph < 0 || use_graph > 2)
        return bfail(SF_E_ARG, "sf_batch_step: invalid argument");
    SF_BHIP(hipSetDevice(b->device));
    if (b->tables_dirty)
        if (int rc = rebuild_tables(b)) return rc;
    if (int rc = prepare_pending(b)) return rc;
    if (b->nactive == 0 || nsteps == 0) return 0;
    if (int rc = advance(b, nsteps, use_graph)) return rc;
    SF_BHIP(hipStreamSynchronize(b->stream));
    std::vector<unsigned> flags((size_t)b->item_capacity);
    int faulted = 0;
    if (int rc = read_faults(b, flags.data(), &faulted)) return rc;
    if (faulted) {
        int first = 0;
        while (!flags[first]) ++first;
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "sf_batch_step: %d problem(s) raised numerical fault flags (slot %d: 0x%x); "
                      "they are frozen, the others advanced",
                      faulted, first, flags[first]);
        return bfail(SF_E_NUMERIC, msg);
    }
    return 0;
}

int sf_batch_faults(sf_batch* b, uint32_t* flags) {
    if (!b || !flags) return bfail(SF_E_ARG, "sf_batch_faults: null argument");
    SF_BHIP(hipSetDevice(b->device));
    SF_BHIP(hipStreamSynchronize(b->stream));
    int faulted = 0;
    return read_faults(b, flags, &faulted);
}

int sf_batch_clock(sf_batch* b, int slot, double* time, long long* steps) {
    if (!valid_slot(b, slot)) return bfail(SF_E_ARG, "sf_batch_clock: invalid slot");
    SF_BHIP(hipSetDevice(b->device));
    SF_BHIP(hipStreamSynchronize(b->stream));
    if (time) SF_BHIP(hipMemcpy(time, b->d_time + slot, sizeof(double), hipMemcpyDeviceToHost));
    if (steps)
        SF_BHIP(hipMemcpy(steps, b->d_steps + slot, sizeof(long long), hipMemcpyDeviceToHost));
    return 0;
}

int sf_batch_steps(sf_batch* b, long long* steps) {
    if (!b || !steps) return bfail(SF_E_ARG, "sf_batch_steps: null argument");
    SF_BHIP(hipSetDevice(b->device));
    SF_BHIP(hipStreamSynchronize(b->stream));
    SF_BHIP(hipMemcpy(steps, b->d_steps, sizeof(long long) * b->item_capacity,
                      hipMemcpyDeviceToHost));
    for (int s = 0; s < b->item_capacity; ++s)
        if (!b->used[s]) steps[s] = -1;
    return 0;
}

int sf_batch_memory(const sf_batch* b, size_t* used_nodes, size_t* free_nodes, int* items) {
    if (!b) return bfail(SF_E_ARG, "sf_batch_memory: null batch");
    long long fr = 0;
    for (const auto& kv : b->free_ranges) fr += kv.second;
    if (used_nodes) *used_nodes = (size_t)(b->capacity - fr);
    if (free_nodes) *free_nodes = (size_t)fr;
    if (items) *items = (int)std::count(b->used.begin(), b->used.end(), 1);
    return 0;
}

int sf_batch_set_tiles(sf_batch* b, int xlen, int ylen) {
    if (!b) return bfail(SF_E_ARG, "sf_batch_set_tiles: null batch");
    if (xlen != 63 && xlen != 127 && xlen != 255)
        return bfail(SF_E_ARG, "sf_batch_set_tiles: x tile must be 63, 127 or 255 nodes");
    if (ylen < 4 || ylen > 4096) return bfail(SF_E_ARG, "sf_batch_set_tiles: y tile out of range");
    b->xbs = xlen + 1;
    b->yseg = ylen;
    b->tables_dirty = true;
    return 0;
}

}  // extern "C"

// ---- autotuner ----------------------------------------------------------------
// Times every launch shape (schedule x x-tile x y-tile x persistent grid
// size) on the batch's own problems: each trial replays the captured step
// from the same saved state, which is restored afterwards, so tuning leaves
// the problems exactly as it found them.
namespace {

struct Snapshot {
    double* q = nullptr;
    long long extent = 0;
    std::vector<double> dt, time;
    std::vector<unsigned long long> dtmin;
    std::vector<long long> steps;
    std::vector<unsigned> flags;
};

int snapshot_save(sf_batch* b, Snapshot& sn) {
    sn.extent = 0;
    for (int s = 0; s < b->item_capacity; ++s)
        if (b->used[s])
            sn.extent = std::max(sn.extent, b->items[s].off + padded_nodes(b->desc[s]));
    SF_BHIP(hipMalloc(&sn.q, sizeof(double) * 5 * (size_t)std::max(1LL, sn.extent)));
    for (int v = 0; v < 5; ++v)
        SF_BHIP(hipMemcpyAsync(sn.q + v * sn.extent, b->buf[0] + v * b->capacity,
                               sizeof(double) * sn.extent, hipMemcpyDeviceToDevice, b->stream));
    const size_t n = (size_t)b->item_capacity;
    sn.dt.resize(n);
    sn.time.resize(n);
    sn.dtmin.resize(n);
    sn.steps.resize(n);
    sn.flags.resize(n);
    SF_BHIP(hipStreamSynchronize(b->stream));
    SF_BHIP(hipMemcpy(sn.dt.data(), b->d_dt, sizeof(double) * n, hipMemcpyDeviceToHost));
    SF_BHIP(hipMemcpy(sn.time.data(), b->d_time, sizeof(double) * n, hipMemcpyDeviceToHost));
    SF_BHIP(hipMemcpy(sn.dtmin.data(), b->d_dtmin, 8 * n, hipMemcpyDeviceToHost));
    SF_BHIP(hipMemcpy(sn.steps.data(), b->d_steps, 8 * n, hipMemcpyDeviceToHost));
    SF_BHIP(hipMemcpy(sn.flags.data(), b->d_flags, 4 * n, hipMemcpyDeviceToHost));
    return 0;
}

int snapshot_restore(sf_batch* b, const Snapshot& sn) {
    for (int v = 0; v < 5; ++v)
        SF_BHIP(hipMemcpyAsync(b->buf[0] + v * b->capacity, sn.q + v * sn.extent,
                               sizeof(double) * sn.extent, hipMemcpyDeviceToDevice, b->stream));
    SF_BHIP(hipStreamSynchronize(b->stream));
    const size_t n = (size_t)b->item_capacity;
    SF_BHIP(hipMemcpy(b->d_dt, sn.dt.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    SF_BHIP(hipMemcpy(b->d_time, sn.time.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    SF_BHIP(hipMemcpy(b->d_dtmin, sn.dtmin.data(), 8 * n, hipMemcpyHostToDevice));
    SF_BHIP(hipMemcpy(b->d_steps, sn.steps.data(), 8 * n, hipMemcpyHostToDevice));
    SF_BHIP(hipMemcpy(b->d_flags, sn.flags.data(), 4 * n, hipMemcpyHostToDevice));
    return 0;
}

// Mean time of one step of the current shape over `steps` graph replays
// (after one untimed replay that captures the graph).
int time_shape(sf_batch* b, int steps, double* ms) {
    if (b->tables_dirty)
        if (int rc = rebuild_tables(b)) return rc;
    if (int rc = advance(b, 1, 1)) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SF_BHIP(hipEventCreate(&e0));
    SF_BHIP(hipEventCreate(&e1));
    int rc = 0;
    if (hipEventRecord(e0, b->stream) != hipSuccess) rc = bfail(SF_E_HIP, "event record failed");
    if (!rc) rc = advance(b, steps, 1);
    if (!rc && hipEventRecord(e1, b->stream) != hipSuccess)
        rc = bfail(SF_E_HIP, "event record failed");
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = bfail(SF_E_HIP, "event wait failed");
    float t = 0.0f;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess)
        rc = bfail(SF_E_HIP, "event timing failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / steps;
    return rc;
}

}  // namespace

extern "C" int sf_batch_autotune(sf_batch* b, int trial_steps, sf_batch_tuning* out) {
    if (!b || !out || trial_steps < 1) return bfail(SF_E_ARG, "sf_batch_autotune: invalid argument");
    SF_BHIP(hipSetDevice(b->device));
    if (b->tables_dirty)
        if (int rc = rebuild_tables(b)) return rc;
    if (int rc = prepare_pending(b)) return rc;
    if (b->nactive == 0) return bfail(SF_E_STATE, "sf_batch_autotune: the batch is empty");
    Snapshot sn;
    int rc = snapshot_save(b, sn);
    const int keep[4] = {b->xbs, b->yseg, b->blocks_per_cu, b->fused};
    sf_batch_tuning best{};
    best.ms_per_step = 1e300;
    int tried = 0;
    const int xlens[3] = {63, 127, 255};
    const int occs[5] = {0, 2, 4, 8, 16};   // 0: as many as stay resident
    std::vector<int> ysegs = b->any2d ? std::vector<int>{16, 32, 64} : std::vector<int>{b->yseg};
    for (int fused = 0; fused < 2 && !rc; ++fused)
        for (int xl : xlens)
            for (int ys : ysegs)
                for (int oc : occs) {
                    if (rc) break;
                    b->fused = fused;
                    b->xbs = xl + 1;
                    b->yseg = ys;
                    b->blocks_per_cu = oc;
                    b->tables_dirty = true;
                    double ms = 0.0;
                    rc = time_shape(b, trial_steps, &ms);
                    if (!rc) rc = snapshot_restore(b, sn);
                    if (rc) break;
                    ++tried;
                    if (ms < best.ms_per_step) {
                        best.xlen = xl;
                        best.ylen = ys;
                        best.blocks_per_cu = oc;
                        best.fused = fused;
                        best.ms_per_step = ms;
                    }
                }
    if (sn.q) (void)hipFree(sn.q);
    if (rc) {
        b->xbs = keep[0];
        b->yseg = keep[1];
        b->blocks_per_cu = keep[2];
        b->fused = keep[3];
        b->tables_dirty = true;
        return rc;
    }
    best.candidates = tried;
    b->xbs = best.xlen + 1;
    b->yseg = best.ylen;
    b->blocks_per_cu = best.blocks_per_cu;
    b->fused = best.fused;
    b->tables_dirty = true;
    *out = best;
    return 0;
}

extern "C" int sf_batch_set_schedule(sf_batch* b, int fused) {
    if (!b) return bfail(SF_E_ARG, "sf_batch_set_schedule: null batch");
    b->fused = fused ? 1 : 0;
    drop_graph(b);
    return 0;
}

extern "C" int sf_batch_set_occupancy(sf_batch* b, int blocks_per_cu) {
    if (!b) return bfail(SF_E_ARG, "sf_batch_set_occupancy: null batch");
    if (blocks_per_cu < 0 || blocks_per_cu > 64)
        return bfail(SF_E_ARG,
                     "sf_batch_set_occupancy: 1..64 workgroups per compute unit, or 0 for "
                     "as many as stay resident");
    b->blocks_per_cu = blocks_per_cu;
    drop_graph(b);
    return 0;
}

namespace sfamd {

long long batch_copy_doubles(const sf_batch* b, int slot) {
    const BatchItemDev& it = b->items[(size_t)slot];
    return 5LL * it.m0 * (it.n1 + 2 * it.gy) + 1;
}

int batch_copy_out_async(sf_batch* b, int slot, double* dst, hipEvent_t done) {
    if (!valid_slot(b, slot)) return bfail(SF_E_ARG, "batch copy-out: invalid slot");
    SF_BHIP(hipSetDevice(b->device));
    const BatchItemDev& it = b->items[(size_t)slot];
    const long long extent = (long long)it.m0 * (it.n1 + 2 * it.gy);
    SF_BHIP(hipEventRecord(b->copy_after, b->stream));
    SF_BHIP(hipStreamWaitEvent(b->copy_stream, b->copy_after, 0));
    // the state after a whole step is in buf[0]
    for (int v = 0; v < 5; ++v)
        SF_BHIP(hipMemcpyAsync(dst + v * extent, b->buf[0] + v * b->capacity + it.off,
                               sizeof(double) * (size_t)extent, hipMemcpyDeviceToHost,
                               b->copy_stream));
    SF_BHIP(hipMemcpyAsync(dst + 5 * extent, b->d_time + slot, sizeof(double),
                           hipMemcpyDeviceToHost, b->copy_stream));
    SF_BHIP(hipEventRecord(done, b->copy_stream));
    return 0;
}

void batch_unpack(const sf_batch* b, int slot, const double* src, double* soa, double* time) {
    const BatchItemDev& it = b->items[(size_t)slot];
    const long long extent = (long long)it.m0 * (it.n1 + 2 * it.gy);
    const long long N = (long long)it.n0 * it.n1;
    for (int v = 0; v < 5; ++v)
        for (int j = 0; j < it.n1; ++j) {
            const double* row = src + v * extent + (long long)(j + it.gy) * it.m0 + 3;
            std::copy(row, row + it.n0, soa + v * N + (long long)j * it.n0);
        }
    *time = src[5 * extent];
}

}  // namespace sfamd
