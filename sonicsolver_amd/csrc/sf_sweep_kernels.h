// sf_sweep_kernels.h -- the sweep kernels of the grid path: the x sweep
// over flattened face tiles and the y / z march.  Templates only, so every
// translation unit that instantiates a scheme can include it.
#pragma once

#include "sf_device.h"

namespace sfamd {

// ---------------------------------------------------------------------------
// x sweep over the faces of all x lines numbered as one sequence: face
// f = line * (n0 + 1) + (a + 1) lies between nodes a and a+1 of its line
// (a = -1 .. n0-1).  A block is one wave; tile T is faces T*63 .. T*63+63:
// lane t evaluates face T*63+t and lanes 1..63 form the difference for the
// node whose upper face is theirs, taking the lower face's flux from lane
// t-1 by a cross-lane shift (no LDS, no barrier).  Tiles overlap by one
// face, so every node is updated by exactly one lane, and no tile is left
// mostly idle at the end of a line.  Blocks loop over tiles (stride
// gridDim.x), so the CFL fold costs one atomic per block.
// ---------------------------------------------------------------------------
constexpr int kXBlock = 64;

// The value of lane t-1 (lane 0 receives 0): one DPP wave shift per 32-bit
// half, a VALU move with no LDS round trip.
__device__ __forceinline__ double lane_shift_up1(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The value of lane 63, as a wave-uniform value.
__device__ __forceinline__ double lane63(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The final sweep's operands for node id: its current value, the flux
// difference the y and z sweeps accumulated and its step-start value.  All
// three are loaded for every lane, without a test: every lane's node index
// lies in the padded grid (a lane without a node of its own takes node
// (-1 or 0, j, k)); on a grid without a y axis no sweep writes S, which
// keeps the zeros it was created with, so its sum is the x difference
// alone; where the recipe does not read q0 (Euler, the first SSPRK3
// stage) the values go unused, and there q0 is the stage input itself, so
// the loads hit the lines the stencil loads bring in.
template <int TIME>
__device__ __forceinline__ void stage_operands(const GridDev& G, const StageArgs& A,
                                               const double* __restrict__ q,
                                               const double* __restrict__ S, long long id,
                                               double (&pq)[5], double (&ps)[5],
                                               double (&pq0)[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        pq[v] = q[v * G.P + id];
        ps[v] = S[v * G.P + id];
        pq0[v] = A.q0[v * G.P + id];
    }
}

// Three waves per SIMD: the registers that hold the final sweep's stage
// operands through the face arithmetic pay more than a fourth wave would
// (and four would spill).
// The x sweep is always a stage's last sweep (x is always active and the
// sweeps run z, y, x), so it always applies the stage update.  TIME: the
// time recipe (A.time, fixed at compile time so that the update carries no
// code for the other recipes).
template <int SCHEME, int FLUX, int TIME>
__global__ __launch_bounds__(kXBlock)
__attribute__((amdgpu_waves_per_eu(3, 8))) void k_sweep_x(GridDev G, const double* __restrict__ q,
                                                     const double* __restrict__ S, StageArgs A) {
    if (*A.flags) return;   // a fault was raised: keep the step-start state
    // face numbers fit 32 bits: the host rejects grids with 2^32 - 2^8
    // faces or more along x
    const int t = threadIdx.x;
    const unsigned per_line = G.n[0] + 1;
    const unsigned faces = per_line * G.n[1] * G.n[2];
    const unsigned tiles = (faces - 1 + kXBlock - 2) / (kXBlock - 1);
    unsigned flags = 0;
    double dt_local = INFINITY;
    const double dt_step = TIME == 3 ? 0.0 : *A.dt;   // the step's dt, read once
    const double lw = (FLUX == 5) ? *A.dt / G.d[0] : 0.0;   // Lax-Wendroff dt / spacing
    // Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8),
    // each with its own L2; tiles are numbered so that the blocks of one XCD
    // take consecutive tiles and share the stencil halo between them in that
    // L2.  The grid size is a multiple of 8.
    const unsigned per_xcd = gridDim.x >> 3;
    const unsigned first = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    for (unsigned T = first; T < tiles; T += gridDim.x) {
        const unsigned f = T * (kXBlock - 1) + t;
        const bool face = f < faces;
        const unsigned line = face ? fast_div(f, G.div_line) : 0;
        const int a = face ? (int)(f - line * per_line) - 1 : 0;  // lower node of the face
        const unsigned kz = fast_div(line, G.div_n1);
        const int j = (int)(line - kz * (unsigned)G.n[1]), k = (int)kz;
        double f5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        // The final sweep's stage operands for the node this lane updates
        // (the lower node of its face): loaded before the face arithmetic,
        // so that their latency hides behind it -- except for WENO7, whose
        // eight-node stencil leaves no registers to hold them at three waves
        // per SIMD; it loads them after the face.
        const bool upd = t >= 1 && face && a >= 0;
        const long long id = node_index(G, a, j, k);
        double pq[5], ps[5], pq0[5];
        constexpr bool kEarly = SCHEME != 3;
        if (kEarly) stage_operands<TIME>(G, A, q, S, id, pq, ps, pq0);
        if (face) {
            constexpr int NS = Stencil<SCHEME>::ns;
            double st[NS][5];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int v = 0; v < 5; ++v) st[s][v] = q[v * G.P + id + (s + 1 - NS / 2)];
            face_flux<0, SCHEME, FLUX>(st, G.gamma, G.area[0], f5, flags, lw);
            if (!finite_sum5(f5)) flags |= kFlagFlux;
        }
        double lower[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) lower[v] = lane_shift_up1(f5[v]);
        if (upd) {
            if (!kEarly) stage_operands<TIME>(G, A, q, S, id, pq, ps, pq0);
            double s5[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double d = f5[v] - lower[v];
                s5[v] = ps[v] + d;
            }
            double dl = INFINITY;
            double qn[5];
            if (solved_node(G, a, j, k)) {
                stage_update_dt<TIME>(G, A, id, pq, pq0, s5, dt_step, flags, dl, qn);
            } else if (TIME != 3) {
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    qn[v] = pq[v];   // carry_node
                    A.qout[v * G.P + id] = qn[v];
                }
            }
            if (TIME != 3 && G.images) write_images(G, A.qout, a, j, k, id, qn);
            dt_local = fmin(dt_local, dl);
        }
    }
    raise_fault(flags, A.flags, A.fault_tag, A.tag);
    if (A.reduce_dt) wave_min_atomic(dt_local, A.dtmin);
}

// ---------------------------------------------------------------------------
// x sweep by whole lines, for grids with many x lines: a block (one wave)
// takes groups of kLineGroup lines and walks each line in tiles of 64
// nodes, lane t evaluating the upper face of node 64m + t; node 64m + t
// takes its lower face from lane t-1 by the cross-lane shift, lane 0 from
// the previous tile's lane 63 (kept in scalar registers), and the first
// tile of a line from the line's first face, which the group evaluates for
// its lines at once before walking them.  Rows start on 128-byte lines, so
// a tile's loads and stores are whole lines.  Every node takes the same
// faces and the same arithmetic as in k_sweep_x, so the two give identical
// results.
//
// All loads go through LDS by 16-byte global-to-LDS copies, which the
// memory pipe completes while the wave computes: the conserved variables
// of the next tile with a halo of four nodes on each side (72 per
// variable, two buffers used in turn) are requested before the current
// tile's faces, and the current tile's S and q0 (one 64-node row each,
// lanes 0-31 one variable and 32-63 the next, two nodes per lane) before
// its update.  The stencils are read from LDS, so the face arithmetic
// never waits for memory and no register holds an operand across it.  The
// copies are not counted by the compiler for LDS reads: one explicit wait
// for the vector memory count to reach zero, after the faces, covers both
// the tile's S, q0 and the next tile's variables (all requested before it).
// ---------------------------------------------------------------------------
constexpr int kLineGroup = 8;   // 16 ran 1-2.5% slower on 256^3, the same on 512^3
constexpr int kHalo = 4;                      // nodes each side of a tile
constexpr int kRow = kXBlock + 2 * kHalo;     // 72 doubles per variable
constexpr int kRowLanes = kRow / 2;           // lanes of 16 bytes per row

template <int SCHEME, int FLUX, int TIME>
__global__ __launch_bounds__(kXBlock)
__attribute__((amdgpu_waves_per_eu(3, 8))) void k_sweep_xl(GridDev G, const double* __restrict__ q,
                                                      const double* __restrict__ S, StageArgs A) {
    if (*A.flags) return;   // a fault was raised: keep the step-start state
    constexpr int NS = Stencil<SCHEME>::ns;
    static_assert(NS / 2 <= kHalo, "the halo covers the stencil");
    __shared__ double first[5][kLineGroup];   // each line's face -1/2
    __shared__ __attribute__((aligned(16))) double qb[2][5][kRow];
    __shared__ __attribute__((aligned(16))) double pk[10][kXBlock];   // S 0-4, q0 5-9
    const int t = threadIdx.x;
    const int n0 = G.n[0];
    const unsigned lines = (unsigned)G.n[1] * (unsigned)G.n[2];
    const unsigned groups = (lines + kLineGroup - 1) / kLineGroup;
    const int tiles = (n0 + kXBlock - 1) / kXBlock;
    const int last_needed = n0 - 1 + NS / 2;   // the highest node a stencil reads
    // S / q0 copies: byte offset of this lane's two nodes in the lower
    // variable of a copy, and of the upper variable (P doubles further;
    // the host keeps 8 P + 1024 below 2^32)
    const int pair = t & 31;
    const unsigned up = (t >> 5) ? (unsigned)G.P * 8u : 0u;
    unsigned flags = 0;
    double dt_local = INFINITY;
    const double dt_step = TIME == 3 ? 0.0 : *A.dt;   // the step's dt, read once
    const double lw = (FLUX == 5) ? *A.dt / G.d[0] : 0.0;   // Lax-Wendroff dt / spacing

    // the 72 nodes 64m-4 .. 64m+67 of line (j, k), all five variables, into
    // buffer b; lanes 0-35 take two nodes each, a lane whose first node no
    // stencil reads copies the row's first two again (the second node of a
    // pair is at most one past the last ghost, which every row holds)
    auto fetch_q = [&](int b, int m, int j, int k) {
        if (t < kRowLanes) {
            const int node0 = m * kXBlock - kHalo;
            const long long id = node_index(G, node0, j, k);
            const unsigned off = (node0 + 2 * t <= last_needed ? (unsigned)t : 0u) * 16u;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const char* base = (const char*)(q + v * G.P + id);
                __builtin_amdgcn_global_load_lds(
                    (__attribute__((address_space(1))) void*)(base + off),
                    (__attribute__((address_space(3))) void*)&qb[b][v][0], 16, 0, 0);
            }
        }
    };
    auto wait_copies = [] {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __builtin_amdgcn_sched_barrier(0);
    };

    for (unsigned grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const unsigned line0 = grp * kLineGroup;
        const unsigned nl = min((unsigned)kLineGroup, lines - line0);
        int j, k;
        {
            const unsigned kz = fast_div(line0, G.div_n1);
            j = (int)(line0 - kz * (unsigned)G.n[1]);
            k = (int)kz;
        }
        fetch_q(0, 0, j, k);   // the group's first tile
        // the first face (-1/2) of the group's lines, lane l for line0 + l
        {
            const unsigned line = line0 + (unsigned)t;
            if (t < kLineGroup && line < lines) {
                const unsigned kz = fast_div(line, G.div_n1);
                const int jl = (int)(line - kz * (unsigned)G.n[1]), kl = (int)kz;
                const long long id = node_index(G, -1, jl, kl);
                double st[NS][5];
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int v = 0; v < 5; ++v) st[s][v] = q[v * G.P + id + (s + 1 - NS / 2)];
                double f5[5];
                face_flux<0, SCHEME, FLUX>(st, G.gamma, G.area[0], f5, flags, lw);
                if (!finite_sum5(f5)) flags |= kFlagFlux;
#pragma unroll
                for (int v = 0; v < 5; ++v) first[v][t] = f5[v];
            }
        }
        wait_copies();
        __syncthreads();
        int b = 0;   // the buffer holding the current tile's variables
        for (unsigned l = 0; l < nl; ++l) {
            double carry[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) carry[v] = first[v][l];
            for (int m = 0; m < tiles; ++m) {
                const int a = m * kXBlock + t;   // this lane's node
                const int rem = n0 - m * kXBlock;
                const bool node = t < rem;
                const long long id0 = node_index(G, m * kXBlock, j, k);
                // the next tile's variables (next line's first tile at the
                // end of a line, none after the group's last)
                {
                    int jn = j, kn = k, mn = m + 1;
                    bool more = true;
                    if (mn == tiles) {
                        mn = 0;
                        more = l + 1 < nl;
                        if (++jn == G.n[1]) { jn = 0; ++kn; }
                    }
                    if (more) fetch_q(b ^ 1, mn, jn, kn);
                }
                // this tile's S and q0; a lane whose nodes lie past the
                // line's end copies the tile's first two again
                {
                    const unsigned lo = (unsigned)(2 * pair < rem ? pair : 0) * 16u;
                    const double* const src[2] = {S, A.q0};
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        const int arr = c / 3, v = 2 * (c % 3);
                        if (v < 4 || t < 32) {
                            const char* base = (const char*)(src[arr] + v * G.P + id0);
                            const unsigned off = lo + (v < 4 ? up : 0u);
                            __builtin_amdgcn_global_load_lds(
                                (__attribute__((address_space(1))) void*)(base + off),
                                (__attribute__((address_space(3))) void*)&pk[5 * arr + v][0], 16, 0,
                                0);
                        }
                    }
                }
                // the stencil from LDS (lanes past the line's end take the
                // last node's)
                const int ti = node ? t : rem - 1;
                double st[NS][5];
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int v = 0; v < 5; ++v) st[s][v] = qb[b][v][ti + kHalo + s + 1 - NS / 2];
                double f5[5];
                face_flux<0, SCHEME, FLUX>(st, G.gamma, G.area[0], f5, flags, lw);
                if (!finite_sum5(f5)) flags |= kFlagFlux;
                double lower[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const double sh = lane_shift_up1(f5[v]);
                    lower[v] = t == 0 ? carry[v] : sh;
                    carry[v] = lane63(f5[v]);   // the next tile's first lower face
                }
                wait_copies();
                if (node) {
                    const long long id = id0 + t;
                    double pq[5], ps[5], pq0[5];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        pq[v] = qb[b][v][t + kHalo];
                        ps[v] = pk[v][t];
                        pq0[v] = pk[5 + v][t];
                    }
                    double s5[5];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        const double d = f5[v] - lower[v];
                        s5[v] = ps[v] + d;
                    }
                    double dl = INFINITY;
                    double qn[5];
                    if (solved_node(G, a, j, k)) {
                        stage_update_dt<TIME>(G, A, id, pq, pq0, s5, dt_step, flags, dl, qn);
                    } else if (TIME != 3) {
#pragma unroll
                        for (int v = 0; v < 5; ++v) {
                            qn[v] = pq[v];   // carry_node
                            A.qout[v * G.P + id] = qn[v];
                        }
                    }
                    if (TIME != 3 && G.images) write_images(G, A.qout, a, j, k, id, qn);
                    dt_local = fmin(dt_local, dl);
                }
                b ^= 1;
            }
            if (++j == G.n[1]) { j = 0; ++k; }
        }
        __syncthreads();
    }
    raise_fault(flags, A.flags, A.fault_tag, A.tag);
    if (A.reduce_dt) wave_min_atomic(dt_local, A.dtmin);
}

// ---------------------------------------------------------------------------
// y / z sweep: lane = x index, blockIdx.y = the remaining transverse index,
// blockIdx.z = segment of SEG nodes along the sweep axis.  Each lane marches
// along its column; the NS-node stencil window lives in LDS (one private
// slot set per lane, a ring indexed by `head`), so between faces only the
// previous face flux stays in registers.  Lanes never read each other's
// slots, so no barrier is needed.
// ---------------------------------------------------------------------------
constexpr int kMarchBlock = 64;

// The WENO5 ring (10 KB per wave) leaves room for four waves per SIMD; the
// register count decides how many run: the WENO5 / WENO3 marches with the
// Rusanov or HLLC flux fit 128 registers and run four, the other fluxes'
// need a few more and run three (holding them to 128 spills).
template <int AXIS, int SCHEME, int FLUX, int MODE>
__global__ __launch_bounds__(kMarchBlock)
void k_sweep_march(GridDev G,
                                                             const double* __restrict__ q,
                                                             double* __restrict__ S, StageArgs A,
                                                             int seg) {
    constexpr int NS = Stencil<SCHEME>::ns;
    constexpr int H = NS / 2;    // face a+1/2 reads nodes a+1-H .. a+H
    constexpr int W = NS - 2;    // LDS ring slots: nodes a+1-H .. a+H-2
    __shared__ double win[W][5][kMarchBlock];
    const int lane = threadIdx.x;
    const int i = blockIdx.x * kMarchBlock + lane;
    const int o = blockIdx.y;
    const int a0 = blockIdx.z * seg;
    const int a1 = min(a0 + seg, G.n[AXIS]);
    static_assert(MODE == kCopyS || MODE == kAccumS, "a march writes or accumulates S; the x sweep applies the update");
    if (*A.flags) return;   // a fault was raised: keep the step-start state
    unsigned flags = 0;
    const double lw = (FLUX == 5) ? *A.dt / G.d[AXIS] : 0.0;   // Lax-Wendroff dt / spacing
    if (i < G.n[0]) {
        const int j0 = (AXIS == 1) ? a0 : o;
        const int k0 = (AXIS == 2) ? a0 : o;
        const long long step = G.stride[AXIS];
        // The face a+1/2 reads its NS nodes as W from the lane's LDS ring
        // (nodes a+1-H .. a+H-2), node a+H-1 carried in registers from the
        // previous face and node a+H loaded for this face.  The march starts
        // one node early (a = a0-1) so that face a0-1/2 is evaluated by the
        // same instructions as every other face: a face's rounding then does
        // not depend on where a segment starts.
        long long id = node_index(G, i, j0, k0) - H * step;  // node a0-H
        for (int s = 0; s < W; ++s)
#pragma unroll
            for (int v = 0; v < 5; ++v) win[s][v][lane] = q[v * G.P + id + s * step];
        double carried[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) carried[v] = q[v * G.P + id + W * step];  // node a0+H-2
        id += (H - 1) * step;  // node a0 - 1
        int head = 0;          // slot of the oldest node
        double fprev[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma clang loop unroll(disable)
        for (int a = a0 - 1; a < a1; ++a, id += step) {
            double st[NS][5];
#pragma unroll
            for (int v = 0; v < 5; ++v) st[NS - 1][v] = q[v * G.P + id + H * step];
#pragma unroll
            for (int s = 0; s < W; ++s) {
                const int slot = (head + s >= W) ? head + s - W : head + s;
#pragma unroll
                for (int v = 0; v < 5; ++v) st[s][v] = win[slot][v][lane];
            }
            // the oldest node leaves the ring, the carried one enters it
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                st[NS - 2][v] = carried[v];
                win[head][v][lane] = carried[v];
                carried[v] = st[NS - 1][v];
            }
            head = (head == W - 1) ? 0 : head + 1;
            double f[5];
            face_flux<AXIS, SCHEME, FLUX>(st, G.gamma, G.area[AXIS], f, flags, lw);
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
#pragma unroll
            for (int v = 0; v < 5; ++v) S[v * G.P + id] = s5[v];
        }
    }
    raise_fault(flags, A.flags, A.fault_tag, A.tag);
}

}  // namespace sfamd
