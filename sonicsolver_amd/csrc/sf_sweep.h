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
};

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
            // up to 16384 blocks (a multiple of 8: the kernel deals tiles to
            // XCDs), each looping over tiles of kXBlock faces
            const long long faces = (long long)(G.n[0] + 1) * G.n[1] * G.n[2];
            const long long tiles = (faces - 1 + kXBlock - 2) / (kXBlock - 1);
            const unsigned bx = (unsigned)((std::min<long long>(tiles, 16384) + 7) & ~7LL);
            switch (c.A.time) {
            case 0: k_sweep_x<SCHEME, FLUX, 0><<<bx, kXBlock, 0, c.stream>>>(G, c.q, c.S, c.A); break;
            case 1: k_sweep_x<SCHEME, FLUX, 1><<<bx, kXBlock, 0, c.stream>>>(G, c.q, c.S, c.A); break;
            case 2: k_sweep_x<SCHEME, FLUX, 2><<<bx, kXBlock, 0, c.stream>>>(G, c.q, c.S, c.A); break;
            default: k_sweep_x<SCHEME, FLUX, 3><<<bx, kXBlock, 0, c.stream>>>(G, c.q, c.S, c.A); break;
            }
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
