// sf_device.h -- device-side pieces shared by the grid kernels
// (sf_kernels.h) and the batched kernels (sf_batch.hip): grid descriptor,
// stage-update arithmetic, validation flags, CFL folding and the pressure /
// energy conversions of the boundary closure.
#pragma once

#include "sf_face.h"

namespace sfamd {

// Unsigned 32-bit division by a divisor fixed for a whole launch, as a
// multiply-high, an add and a shift (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, section 4): with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// floor(x / d) = (umulhi(x, m) + x) >> l for every x < 2^32 (the sum is
// formed in 64 bits).  The host forms (m, l) with make_fastdiv.
struct FastDiv {
    unsigned m;
    unsigned l;
};

inline FastDiv make_fastdiv(unsigned d) {
    unsigned l = 0;
    while (l < 32 && (1ull << l) < d) ++l;
    FastDiv f;
    f.m = (unsigned)((((1ull << l) - d) << 32) / d + 1);
    f.l = l;
    return f;
}

__device__ __forceinline__ unsigned fast_div(unsigned x, FastDiv f) {
    const unsigned t = __umulhi(x, f.m);
    return (unsigned)(((unsigned long long)t + x) >> f.l);
}

struct GridDev {
    int n[3];
    int g[3];
    int m[3];
    int xo;             // position of node i = 0 in its row (a multiple of 16)
    long long P;        // padded nodes per variable (below 2^31)
    long long stride[3];
    unsigned ustride[3];  // the strides in 32 bits, for node_index
    int active[3];
    int bc[3];          // SF_BC_* kind of each axis
    double jac, area[3], gamma;
    double metric[3];   // 1 / spacing on active axes, 0 on inactive ones
    double d[3];        // node spacing (the Lax-Wendroff flux's face spacing)
    // Bit a set: the final sweep writes every node it updates also into the
    // periodic ghost positions of axis a that hold its copies, so no ghost
    // fill is needed for that axis between stages.  A node c of such an axis
    // has a copy unless g <= c < n - g, i.e. when (unsigned)(c - img_lo[a])
    // >= img_span[a] (img_lo = g, img_span = n - 2g): at c + n (img_off =
    // n * stride) if c < g, at c - n otherwise.  On other axes img_lo = 0
    // and img_span = UINT_MAX, so no node has a copy.
    int images;
    int img_lo[3];
    unsigned img_span[3];
    long long img_off[3];
    // Nodes the stage advances along each axis: solved_lo[a] .. solved_hi[a]
    // (a zeroGradient end's node is excluded; see solved_node).
    int solved_lo[3], solved_hi[3];
    // The x sweep's face -> (line, node) and line -> (j, k) divisions:
    // by n0 + 1 faces per line and by n1 lines per z plane.
    FastDiv div_line, div_n1;
};

enum { kCopyS = 0, kAccumS = 1, kFinal = 2 };

// Stage update parameters (Time::Explicit, SF_explicit.cpp:193-238).
struct StageArgs {
    const double* q0;        // step-start state (SSPRK3, RK4)
    double* qout;            // new state
    double* k[3];            // RK4 stage residuals k1..k3
    const double* dt;        // device scalar: dt of this step
    unsigned long long* dtmin;  // CFL accumulator (bits of a positive double)
    unsigned* flags;
    double base, incr;       // SSPRK3 base / increment weights
    double cfl;
    int time;                // 0 Euler, 1 SSPRK3, 2 RK4
    int stage;
    int reduce_dt;           // fold the next step's CFL scan into this update
    long long* fault_tag;    // where the first fault is recorded (may be null)
    long long tag;           // this launch's tag (step sequence * 8 + phase)
};

// Fault record of a grid: the flags raised since they were last cleared,
// and the tag of the launch that raised the first of them: step sequence
// number * 8 + phase (0..3 the stage, its ghost fill included; 6 the ghost
// fill at the start or commit of the step; 7 the CFL scan before it).
// Once flags are raised, every later launch of a stage, fill or clock
// kernel returns at entry, so the state at the start of the faulting step
// stays as it was.
struct Fault {
    unsigned flags;
    unsigned pad;
    long long tag;
};
enum { kPhaseFill = 6, kPhaseScan = 7 };

// Padded index of node (i, j, k); the host keeps P below 2^31, so it is
// formed in 32-bit arithmetic.
__device__ __forceinline__ long long node_index(const GridDev& G, int i, int j, int k) {
    return (long long)((unsigned)(k + G.g[2]) * G.ustride[2] + (unsigned)(j + G.g[1]) * G.ustride[1]
                       + (unsigned)(i + G.xo));
}

// Boundary kind per axis (SF_BC_*): 0 zeroGradient on both ends,
// 1 periodic, 2 / 3 / 4 no condition on the lower / upper / both ends (the
// other end zeroGradient).  zg_end: the end (side 0 lower, 1 upper) carries
// zeroGradient.
__host__ __device__ __forceinline__ bool zg_end(const GridDev& G, int d, int side) {
    const int b = G.bc[d];
    return G.active[d] && (b == 0 || (b == 2 && side == 1) || (b == 3 && side == 0));
}

// A node is advanced unless it lies on a zeroGradient end, which the
// boundary applicator marks as a solver-boundary point
// (SF_applicator.h:162-187); an end without a condition is not marked.
// The host folds that rule into the per-axis ranges solved_lo / solved_hi.
__device__ __forceinline__ bool solved_node(const GridDev& G, int i, int j, int k) {
    return i >= G.solved_lo[0] && i <= G.solved_hi[0] && j >= G.solved_lo[1] &&
           j <= G.solved_hi[1] && k >= G.solved_lo[2] && k <= G.solved_hi[2];
}

// Whether the CFL sum takes axis d: a grid's metric is zero on an inactive
// axis, so its term adds an exact zero and needs no test.
template <class GT>
__device__ __forceinline__ bool cfl_axis(const GT& G, int d) { return G.active[d]; }
__device__ __forceinline__ bool cfl_axis(const GridDev&, int) { return true; }

// Local CFL step of RusanovEOS::deltaT (SF_rusanovEOS.cpp:110-147).  GT is
// any descriptor with gamma, active[3] and metric[3].
template <class GT>
__device__ __forceinline__ double local_dt(const GT& G, const double q[5], double cfl,
                                           unsigned& flags) {
    const Closed s = close_state(q, G.gamma);
    double spectral = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (cfl_axis(G, d)) spectral += fabs(s.u[d] * G.metric[d]) + s.c * G.metric[d];
    if (!(spectral > 0.0) || !isfinite(spectral)) flags |= kFlagDt;
    return cfl / spectral;
}

// A 64-bit value moved across lanes by one DPP control (two 32-bit moves);
// lanes whose row ROWS leaves out keep their own value.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
    const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
    const unsigned l = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
    const unsigned h = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
    return ((unsigned long long)h << 32) | l;
}

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ unsigned long long dpp_min_step(unsigned long long b) {
    const unsigned long long o = dpp_u64<CTRL, ROWS>(b);
    return o < b ? o : b;
}

// Minimum of b over the 64 lanes of a wave, all of them active, by DPP
// moves (no LDS): lane pairs and quads by quad permutations, 8- and
// 16-lane rows by the half-row and row mirrors, then row 0 into row 1 and
// row 2 into row 3 by the row-15 broadcast and rows 0-1 into rows 2-3 by
// the row-31 broadcast, which leaves the minimum in lane 63.
__device__ __forceinline__ unsigned long long wave_min_bits(unsigned long long b) {
    b = dpp_min_step<0xB1>(b);         // quad_perm [1,0,3,2]
    b = dpp_min_step<0x4E>(b);         // quad_perm [2,3,0,1]
    b = dpp_min_step<0x141>(b);        // row_half_mirror
    b = dpp_min_step<0x140>(b);        // row_mirror
    b = dpp_min_step<0x142, 0xa>(b);   // row_bcast:15 into rows 1 and 3
    b = dpp_min_step<0x143, 0xc>(b);   // row_bcast:31 into rows 2 and 3
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// Wave-wide minimum of a positive double into *target (one atomic per wave;
// every lane of the wave calls it).  Positive doubles order like their bit
// patterns; a lane without a value contributes +infinity.
__device__ __forceinline__ void wave_min_atomic(double v, unsigned long long* target) {
    const unsigned long long b = wave_min_bits(
        (v > 0.0 && isfinite(v)) ? (unsigned long long)__double_as_longlong(v)
                                 : 0x7FF0000000000000ull);
    if ((threadIdx.x & 63) == 0 && b != 0x7FF0000000000000ull) atomicMin(target, b);
}

// Whether all five values are finite, from their sum: a NaN or an infinity
// among them makes the sum non-finite, and finite values make it finite
// unless it overflows, which only states already far outside the range of
// this path's data could do.
__device__ __forceinline__ bool finite_sum5(const double f[5]) {
    return isfinite(((f[0] + f[1]) + (f[2] + f[3])) + f[4]);
}

__device__ __forceinline__ void raise_flags(unsigned f, unsigned* flags) {
    if (f) atomicOr(flags, f);
}

// raise_flags, recording `tag` if these are the first flags raised.
__device__ __forceinline__ void raise_fault(unsigned f, unsigned* flags, long long* tag_out,
                                            long long tag) {
    if (f && atomicOr(flags, f) == 0u && tag_out) *tag_out = tag;
}

// Node (i, j, k) at `id` was just written with qn: copy it into the ghost
// positions of the axes in G.images that hold its periodic copies (node c
// of an axis with n nodes and g ghost layers is ghost c + n if c < g and
// ghost c - n if c >= n - g; the host sets an axis's bit only where
// n >= 2g, so there is at most one copy per axis).  The sweeps read ghosts
// only along their own axis, so no corner copies are needed.
__device__ __forceinline__ void write_images(const GridDev& G, double* __restrict__ out, int i,
                                             int j, int k, long long id, const double qn[5]) {
    const int c[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if ((unsigned)(c[a] - G.img_lo[a]) < G.img_span[a]) continue;
        const long long off = c[a] < G.img_lo[a] ? G.img_off[a] : -G.img_off[a];
#pragma unroll
        for (int v = 0; v < 5; ++v) out[v * G.P + id + off] = qn[v];
    }
}

// The reference updates the state in place, so a node that the stage does
// not advance (a solver-boundary node) keeps the value of the last ghost
// fill; with separate input and output buffers it is copied across.
template <int TIME = -1, class GT>
__device__ __forceinline__ void carry_node(const GT& G, const StageArgs& A, long long id,
                                           const double* __restrict__ q, double qn[5]) {
    const int time = TIME >= 0 ? TIME : A.time;
    if (time == 3) return;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        qn[v] = q[v * G.P + id];
        A.qout[v * G.P + id] = qn[v];
    }
}

// Stage update of one node from its current value qc and (SSPRK3 after
// the first stage, RK4) its step-start value q0v; S is the accumulated flux
// difference and dt the step's dt (*A.dt, which no kernel of a stage
// changes, so a sweep reads it once).  GT is any descriptor with P, jac,
// gamma, active[3] and metric[3].  TIME fixes the time recipe (A.time) at
// compile time; -1 reads it from A.
template <int TIME = -1, class GT>
__device__ __forceinline__ void stage_update_dt(const GT& G, const StageArgs& A, long long id,
                                                const double qc[5], const double q0v[5],
                                                const double S[5], double dt, unsigned& flags,
                                                double& dt_local, double qn[5]) {
    const int time = TIME >= 0 ? TIME : A.time;
    if (time == 3) {  // residual only (diagnostics)
#pragma unroll
        for (int v = 0; v < 5; ++v) A.qout[v * G.P + id] = G.jac * S[v];
        return;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const double R = G.jac * S[v];
        if (time == 1) {
            const double euler = qc[v] - dt * R;
            qn[v] = A.base * q0v[v] + A.incr * euler;   // base = 0 in stage 1
        } else if (time == 0) {
            qn[v] = qc[v] - dt * R;
        } else if (A.stage < 3) {
            A.k[A.stage][v * G.P + id] = R;
            qn[v] = q0v[v] - A.incr * dt * R;
        } else {
            qn[v] = q0v[v]
                  - dt / 6.0 * (A.k[0][v * G.P + id] + 2.0 * A.k[1][v * G.P + id]
                                + 2.0 * A.k[2][v * G.P + id] + R);
        }
        A.qout[v * G.P + id] = qn[v];
    }
    // validateEquationState (SF_equationState.h:25-57): rho, p > 0, finite
    // (p has the sign of the internal energy; with the refined reciprocal
    // of the density instead of a division, that sign can differ only where
    // the internal energy is within an ulp of the kinetic energy)
    const double ke = 0.5 * (qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]) * frcp(qn[0]);
    const double eint = qn[4] - ke;
    if (!(qn[0] > 0.0) || !(eint > 0.0) || !isfinite(eint)) flags |= kFlagState;
    if (A.reduce_dt) dt_local = local_dt(G, qn, A.cfl, flags);
}

// stage_update_dt with the step's dt read from *A.dt.
template <int TIME = -1, class GT>
__device__ __forceinline__ void stage_update_values(const GT& G, const StageArgs& A,
                                                    long long id, const double qc[5],
                                                    const double q0v[5], const double S[5],
                                                    unsigned& flags, double& dt_local,
                                                    double qn[5]) {
    const int time = TIME >= 0 ? TIME : A.time;
    const double dt = time == 3 ? 0.0 : *A.dt;
    stage_update_dt<TIME>(G, A, id, qc, q0v, S, dt, flags, dt_local, qn);
}

// Whether a stage of recipe `time` reads the step-start state q0.
__device__ __forceinline__ bool stage_reads_q0(int time, int stage) {
    return (time == 1 && stage > 0) || time == 2;
}

// stage_update_values with qc and q0v read here.
template <int TIME = -1, class GT>
__device__ __forceinline__ void stage_update(const GT& G, const StageArgs& A, long long id,
                                             const double* __restrict__ q, const double S[5],
                                             unsigned& flags, double& dt_local, double qn[5]) {
    const int time = TIME >= 0 ? TIME : A.time;
    double qc[5], q0v[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        qc[v] = time == 3 ? 0.0 : q[v * G.P + id];
        q0v[v] = stage_reads_q0(time, A.stage) ? A.q0[v * G.P + id] : 0.0;
    }
    stage_update_values<TIME>(G, A, id, qc, q0v, S, flags, dt_local, qn);
}

// Block-wide minimum of a positive double into *target (one atomic; every
// thread of the block calls it).
__device__ __forceinline__ void block_min_atomic(double v, unsigned long long* target,
                                                 unsigned long long* lds) {
    unsigned long long b = wave_min_bits(
        (v > 0.0 && isfinite(v)) ? (unsigned long long)__double_as_longlong(v)
                                 : 0x7FF0000000000000ull);
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nwave; ++w) b = lds[w] < b ? lds[w] : b;
        if (b != 0x7FF0000000000000ull) atomicMin(target, b);
    }
}

// Device clock: dt of the step, physical time, step count, CFL accumulator
// (bits of a positive double) and the dt limits.
struct Clock {
    double dt;
    double time;
    long long steps;
    unsigned long long dtmin;
    double max_dt;
    double limit;
};

__device__ __forceinline__ double pressure_of(const double* q, long long P, long long id,
                                              double gamma) {
    double qq[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) qq[v] = q[v * P + id];
    return close_state(qq, gamma).p;
}

// totalEnergyFromPressure (SF_singleFluidStateModel.cpp:73-109).
__device__ __forceinline__ double energy_from_pressure(const double* q, long long P, long long id,
                                                       double gamma, double p) {
    const double r = q[id], m1 = q[P + id], m2 = q[2 * P + id], m3 = q[3 * P + id];
    const double e = 0.0 + p / ((gamma - 1.0) * r);
    const double kinetic = 0.5 * (1.0 / r) * (m1 * m1 + m2 * m2 + m3 * m3);
    return r * e + kinetic;
}

}  // namespace sfamd
