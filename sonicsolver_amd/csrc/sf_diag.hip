// sf_diag.hip -- point evaluations of the device face arithmetic for
// checking it against a CPU restatement: the face fluxes of given state
// pairs (riemann_flux) and the characteristic WENO face flux of given
// stencils (face_flux), with exactly the device functions the sweeps use.
#include "sonic_amd.h"
#include "sf_device.h"

#include <hip/hip_runtime.h>

#include <string>

namespace sfamd {
void set_last_error(const std::string& message);

template <int AXIS, int FLUX>
__global__ __launch_bounds__(64) void k_riemann(int count, const double* __restrict__ qL,
                                                const double* __restrict__ qR, double gamma,
                                                double lw, double* __restrict__ out,
                                                unsigned* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double a[5], b[5], f[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        a[v] = qL[5 * i + v];
        b[v] = qR[5 * i + v];
    }
    unsigned fl = 0;
    riemann_flux<AXIS, FLUX>(a, b, gamma, lw, f, fl);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[5 * i + v] = f[v];
    flags[i] = fl;
}

template <int AXIS, int SCHEME, int FLUX>
__global__ __launch_bounds__(64) void k_face(int count, const double* __restrict__ st,
                                             double gamma, double area,
                                             double* __restrict__ out,
                                             unsigned* __restrict__ flags) {
    constexpr int NS = Stencil<SCHEME>::ns;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double s[NS][5], f[5];
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int v = 0; v < 5; ++v) s[k][v] = st[(i * NS + k) * 5 + v];
    unsigned fl = 0;
    face_flux<AXIS, SCHEME, FLUX>(s, gamma, area, f, fl);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[5 * i + v] = f[v];
    flags[i] = fl;
}

}  // namespace sfamd

using namespace sfamd;

namespace {

int dfail(int code, const std::string& msg) {
    set_last_error(msg);
    return code;
}

// Device scratch of one diagnostic call: inputs, outputs and flags.
struct Scratch {
    double* in = nullptr;
    double* out = nullptr;
    unsigned* flags = nullptr;
    ~Scratch() {
        if (in) (void)hipFree(in);
        if (out) (void)hipFree(out);
        if (flags) (void)hipFree(flags);
    }
};

template <class Launch>
int run_points(int device, int count, size_t in_doubles, const double* const* srcs,
               const size_t* src_doubles, int nsrc, double* out, uint32_t* flags,
               Launch launch) {
    if (hipSetDevice(device) != hipSuccess) return dfail(SF_E_HIP, "hipSetDevice failed");
    Scratch s;
    if (hipMalloc(&s.in, sizeof(double) * in_doubles) != hipSuccess ||
        hipMalloc(&s.out, sizeof(double) * 5 * (size_t)count) != hipSuccess ||
        hipMalloc(&s.flags, sizeof(unsigned) * (size_t)count) != hipSuccess)
        return dfail(SF_E_MEMORY, "diagnostic scratch allocation failed");
    size_t off = 0;
    for (int k = 0; k < nsrc; ++k) {
        if (hipMemcpy(s.in + off, srcs[k], sizeof(double) * src_doubles[k],
                      hipMemcpyHostToDevice) != hipSuccess)
            return dfail(SF_E_HIP, "diagnostic upload failed");
        off += src_doubles[k];
    }
    launch(s.in, s.out, s.flags, (unsigned)((count + 63) / 64));
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return dfail(SF_E_HIP, "diagnostic kernel failed");
    if (hipMemcpy(out, s.out, sizeof(double) * 5 * (size_t)count, hipMemcpyDeviceToHost) !=
            hipSuccess ||
        (flags && hipMemcpy(flags, s.flags, sizeof(unsigned) * (size_t)count,
                            hipMemcpyDeviceToHost) != hipSuccess))
        return dfail(SF_E_HIP, "diagnostic download failed");
    return 0;
}

template <int AXIS>
void launch_riemann(int flux, int count, const double* a, const double* b, double gamma,
                    double lw, double* out, unsigned* flags, unsigned blocks) {
    switch (flux) {
    case SF_FLUX_HLLC: k_riemann<AXIS, 1><<<blocks, 64>>>(count, a, b, gamma, lw, out, flags); break;
    case SF_FLUX_STEGER_WARMING: k_riemann<AXIS, 2><<<blocks, 64>>>(count, a, b, gamma, lw, out, flags); break;
    case SF_FLUX_LAX_FRIEDRICHS: k_riemann<AXIS, 3><<<blocks, 64>>>(count, a, b, gamma, lw, out, flags); break;
    case SF_FLUX_ROE: k_riemann<AXIS, 4><<<blocks, 64>>>(count, a, b, gamma, lw, out, flags); break;
    case SF_FLUX_LAX_WENDROFF: k_riemann<AXIS, 5><<<blocks, 64>>>(count, a, b, gamma, lw, out, flags); break;
    default: k_riemann<AXIS, 0><<<blocks, 64>>>(count, a, b, gamma, lw, out, flags); break;
    }
}

template <int AXIS>
void launch_face(int scheme, int flux, int count, const double* st, double gamma, double area,
                 double* out, unsigned* flags, unsigned blocks) {
    const int v = 2 * scheme + flux;
    switch (v) {
    case 1: k_face<AXIS, 0, 1><<<blocks, 64>>>(count, st, gamma, area, out, flags); break;
    case 2: k_face<AXIS, 1, 0><<<blocks, 64>>>(count, st, gamma, area, out, flags); break;
    case 3: k_face<AXIS, 1, 1><<<blocks, 64>>>(count, st, gamma, area, out, flags); break;
    default: k_face<AXIS, 0, 0><<<blocks, 64>>>(count, st, gamma, area, out, flags); break;
    }
}

}  // namespace

extern "C" {

int sf_riemann_flux(int device, int flux, int axis, double gamma, double lw, int count,
                    const double* qL, const double* qR, double* out, uint32_t* flags) {
    if (!qL || !qR || !out || count < 1 || axis < 0 || axis > 2 || flux < SF_FLUX_RUSANOV ||
        flux > SF_FLUX_LAX_WENDROFF || !(gamma > 1.0))
        return dfail(SF_E_ARG, "sf_riemann_flux: invalid argument");
    const double* srcs[2] = {qL, qR};
    const size_t n[2] = {5 * (size_t)count, 5 * (size_t)count};
    return run_points(device, count, 10 * (size_t)count, srcs, n, 2, out, flags,
                      [&](double* in, double* o, unsigned* f, unsigned blocks) {
                          const double* a = in;
                          const double* b = in + 5 * (size_t)count;
                          if (axis == 0) launch_riemann<0>(flux, count, a, b, gamma, lw, o, f, blocks);
                          if (axis == 1) launch_riemann<1>(flux, count, a, b, gamma, lw, o, f, blocks);
                          if (axis == 2) launch_riemann<2>(flux, count, a, b, gamma, lw, o, f, blocks);
                      });
}

int sf_face_flux(int device, int scheme, int flux, int axis, double gamma, double area,
                 int count, const double* stencils, double* out, uint32_t* flags) {
    if (!stencils || !out || count < 1 || axis < 0 || axis > 2 || !(gamma > 1.0))
        return dfail(SF_E_ARG, "sf_face_flux: invalid argument");
    if ((scheme != SF_SCHEME_WENO5 && scheme != SF_SCHEME_TENO5) ||
        (flux != SF_FLUX_RUSANOV && flux != SF_FLUX_HLLC))
        return dfail(SF_E_ARG, "sf_face_flux: evaluates WENO5 / TENO5 with Rusanov / HLLC");
    const double* srcs[1] = {stencils};
    const size_t n[1] = {30 * (size_t)count};
    return run_points(device, count, 30 * (size_t)count, srcs, n, 1, out, flags,
                      [&](double* in, double* o, unsigned* f, unsigned blocks) {
                          if (axis == 0) launch_face<0>(scheme, flux, count, in, gamma, area, o, f, blocks);
                          if (axis == 1) launch_face<1>(scheme, flux, count, in, gamma, area, o, f, blocks);
                          if (axis == 2) launch_face<2>(scheme, flux, count, in, gamma, area, o, f, blocks);
                      });
}

}  // extern "C"
