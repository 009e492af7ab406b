// sf_sweep.h -- launch of one sweep (one axis of one stage) of the grid
// path.  Each reconstruction scheme's kernels are compiled in a translation
// unit of their own (sf_sweep_<scheme>.hip instantiates launch_sweep for
// its scheme id), so the twenty scheme x flux combinations build in
// parallel.
#pragma once

#include "sf_sweep_kernels.h"

#include <algorithm>

namespace sfamd {

struct SweepCall {
    GridDev G;
    const double* q;   // stage input state (ghosts filled)
    double* S;         // flux-difference accumulator
    StageArgs A;
    hipStream_t stream;
    int seg;           // base nodes per march segment (y and z sweeps)
    int cus;           // compute units of the device
    int x_walk;        // x sweep by lines (1), by flat tiles (0), automatic (-1)
};

// Grid of the x sweep (a multiple of 8 for the XCD mapping): one block per
// 128 tiles, at least as many as the device holds at once (compute units x
// the kernel's resident blocks per unit) and at most 16384.  Blocks beyond
// the resident ones run in rounds, each paying its start-up again, which
// costs most where a block gets few tiles: the 2048^2 x sweep ran 0.46 ms
// with 16384 blocks (4 tiles each) and 0.32 ms with one resident round;
// 256^3 1.07 and 0.78 ms.  Long runs per block balance worse: 512^3 ran
// 6.06 ms with 16384 blocks (130 tiles each) and 6.19 ms with one round.
template <int SCHEME, int FLUX, int TIME>
unsigned x_blocks(int cus, long long tiles) {
    static int per_cu = 0;   // one per kernel instantiation
    if (per_cu == 0) {
        int n = 0;
        per_cu = (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &n, k_sweep_x<SCHEME, FLUX, TIME>, kXBlock, 0) == hipSuccess && n > 0)
                     ? n : 8;
    }
    const long long resident = (long long)std::max(cus, 8) * per_cu;
    const long long want = std::min<long long>(std::max<long long>(tiles / 128, resident), 16384);
    return (unsigned)((std::min<long long>(tiles, want) + 7) & ~7LL);
}

// The x sweep is always the last sweep of a stage (x is always active and
// the sweeps run z, y, x), so it is compiled in its final mode only, and the
// y / z marches in their S-writing modes only.  Returns -1 for a combination
// that is not compiled.
template <int SCHEME, int FLUX, int MODE>
int launch_sweep_fm(const SweepCall& c, int axis) {
    const GridDev& G = c.G;
    if (axis == 0) {
        if constexpr (MODE != kFinal) {
            return -1;
        } else {
            // Grids with at least 256 x lines per CU (16 lines for every
            // wave slot at four waves per SIMD) walk whole lines
            // (k_sweep_xl): 512^3 ran its x sweep in 5.03 ms against 5.59
            // with the flattened tiles, 256^3 in 0.67 against 0.77 with
            // WENO5/HLLC, 0.67 against 0.77 with WENO5/Steger-Warming, 0.82
            // against 1.03 with TENO5/HLLC, 1.21 against 1.44 with
            // WENO7/HLLC and 0.72 against 0.77 with WENO3/HLLC; 2D grids and
            // smaller 3D ones, with too few lines to fill the device, keep
            // the tiles.
            // The line walk addresses a lane's stage operands by a 32-bit
            // byte offset of up to one variable (8 P bytes) past the tile.
            const long long lines = (long long)G.n[1] * G.n[2];
            const bool fits = (long long)G.P * 8 + 1024 < (1LL << 32);
            const bool by_lines =
                fits && (c.x_walk == 1 ||
                         (c.x_walk < 0 && G.n[0] >= kXBlock &&
                          lines >= 256LL * std::max(c.cus, 8)));
            if (by_lines) {
                const long long groups = (lines + kLineGroup - 1) / kLineGroup;
                const unsigned nb = (unsigned)std::min<long long>(groups, 16384);
#define SF_LAUNCH_XL(T) \
    k_sweep_xl<SCHEME, FLUX, T><<<nb, kXBlock, 0, c.stream>>>(G, c.q, c.S, c.A)
                switch (c.A.time) {
                case 0: SF_LAUNCH_XL(0); break;
                case 1: SF_LAUNCH_XL(1); break;
                case 2: SF_LAUNCH_XL(2); break;
                default: SF_LAUNCH_XL(3); break;
                }
#undef SF_LAUNCH_XL
                return 0;
            }
            // one resident block per slot, each looping over tiles of
            // kXBlock faces
            const long long faces = (long long)(G.n[0] + 1) * G.n[1] * G.n[2];
            const long long tiles = (faces - 1 + kXBlock - 2) / (kXBlock - 1);
#define SF_LAUNCH_X(T)                                                            \
    k_sweep_x<SCHEME, FLUX, T><<<x_blocks<SCHEME, FLUX, T>(c.cus, tiles), kXBlock, 0, \
                                c.stream>>>(G, c.q, c.S, c.A)
            switch (c.A.time) {
            case 0: SF_LAUNCH_X(0); break;
            case 1: SF_LAUNCH_X(1); break;
            case 2: SF_LAUNCH_X(2); break;
            default: SF_LAUNCH_X(3); break;
            }
#undef SF_LAUNCH_X
            return 0;
        }
    }
    if constexpr (MODE == kFinal) {
        return -1;
    } else {
        const unsigned other = (unsigned)(axis == 1 ? G.n[2] : G.n[1]);
        const unsigned cols = (unsigned)((G.n[0] + kMarchBlock - 1) / kMarchBlock);
        // Each segment re-reads the stencil window and evaluates one extra
        // face at its start, so segments of 64 nodes are used wherever that
        // still leaves 8192 one-wave blocks or more to fill the chip
        // (512^3); where 32-node segments leave fewer than 4096 blocks
        // (2048^2 in 2D: 2048), the segment is halved down to 8 nodes until
        // there are 4096, four waves per SIMD.  A face's arithmetic does not
        // depend on the segment length, so neither does the result.
        // c.seg > 0 fixes the length instead.
        auto blocks = [&](int s) {
            return (long long)cols * other * ((G.n[axis] + s - 1) / s);
        };
        int seg = c.seg;
        if (seg <= 0) {
            seg = 32;
            if (blocks(2 * seg) >= 8192) seg *= 2;
            else
                while (seg > 8 && blocks(seg) < 4096) seg /= 2;
        }
        const dim3 grid(cols, other, (unsigned)((G.n[axis] + seg - 1) / seg));
        if (axis == 1)
            k_sweep_march<1, SCHEME, FLUX, MODE><<<grid, kMarchBlock, 0, c.stream>>>(
                G, c.q, c.S, c.A, seg);
        else
            k_sweep_march<2, SCHEME, FLUX, MODE><<<grid, kMarchBlock, 0, c.stream>>>(
                G, c.q, c.S, c.A, seg);
        return 0;
    }
}

template <int SCHEME, int FLUX>
int launch_sweep_f(const SweepCall& c, int axis, int mode) {
    if (mode == kCopyS) return launch_sweep_fm<SCHEME, FLUX, kCopyS>(c, axis);
    if (mode == kAccumS) return launch_sweep_fm<SCHEME, FLUX, kAccumS>(c, axis);
    return launch_sweep_fm<SCHEME, FLUX, kFinal>(c, axis);
}

// One sweep of scheme SCHEME with face flux `flux` (SF_FLUX_*) along
// `axis` in stage mode `mode` (kCopyS, kAccumS, kFinal).
template <int SCHEME>
int launch_sweep(const SweepCall& c, int flux, int axis, int mode) {
    switch (flux) {
    case 0: return launch_sweep_f<SCHEME, 0>(c, axis, mode);
    case 1: return launch_sweep_f<SCHEME, 1>(c, axis, mode);
    case 2: return launch_sweep_f<SCHEME, 2>(c, axis, mode);
    case 3: return launch_sweep_f<SCHEME, 3>(c, axis, mode);
    case 4: return launch_sweep_f<SCHEME, 4>(c, axis, mode);
    default: return launch_sweep_f<SCHEME, 5>(c, axis, mode);
    }
}

extern template int launch_sweep<0>(const SweepCall&, int, int, int);
extern template int launch_sweep<1>(const SweepCall&, int, int, int);
extern template int launch_sweep<2>(const SweepCall&, int, int, int);
extern template int launch_sweep<3>(const SweepCall&, int, int, int);

}  // namespace sfamd
