// sf_capi.hip -- host side of the C ABI declared in include/sonic_amd.h:
// device buffers, the per-stage launch sequence and the error handling.
#include "sonic_amd.h"
#include "sf_kernels.h"
#include "sf_sweep.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

using namespace sfamd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define SF_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail(SF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace sfamd {
// Used by the provider layer (host/sf_provider.cpp) to report through
// sf_last_error.
void set_last_error(const std::string& message) { g_last_error = message; }
}  // namespace sfamd

namespace {

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

}  // namespace

struct sf_grid {
    sf_grid_desc desc{};
    GridDev G{};
    hipStream_t stream = nullptr;
    double* buf[3] = {nullptr, nullptr, nullptr};
    double* S = nullptr;
    double* k[3] = {nullptr, nullptr, nullptr};
    Clock* clock = nullptr;
    Fault* fault = nullptr;      // device fault record
    unsigned* flags = nullptr;   // &fault->flags
    long long seq = 0;           // sequence number of the next step to begin
    // The step with sequence number base_seq starts from buffer base_cur;
    // each step moves the current state by a fixed rotation of the three
    // buffers (plan_buffers), so the buffer a later step started from
    // follows from its sequence number.
    long long base_seq = 0;
    int base_cur = 0;
    int march_seg = 0;           // nodes per march segment (0: automatic)
    int x_walk = -1;             // x sweep by lines (1), by flat tiles (0), automatic (-1)
    int cus = 256;               // compute units of the grid's device
    double* dt_hist = nullptr;
    int dt_hist_cap = 0;
    double* stage_io = nullptr;  // unique-node staging for transfers
    size_t bytes = 0;
    int cur = 0;                 // buffer holding the current state
    int ghosts_valid = 0;        // ghost layers of buf[cur] are up to date
    int dtmin_valid = 0;         // clock->dtmin holds the scan of buf[cur]
    double dtmin_cfl = 0.0;      // CFL number the folded scan used
    int step_open = 0;
    int next_stage = 0;
    int stage_in[4] = {0, 0, 0, 0}, stage_out[4] = {0, 0, 0, 0};
    double step_cfl = 0.5;       // CFL folded into the last stage's update
    int fold_cfl = 0;
    double max_dt = std::numeric_limits<double>::max();
    double limit = std::numeric_limits<double>::max();
    // kernel timing (sf_profile_*): event pairs recorded around launches
    int profiling = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_used;  // (kernel, index of start event)
};

namespace {

int nstages(int time) { return time == SF_TIME_EULER ? 1 : time == SF_TIME_SSPRK3 ? 3 : 4; }

int check_desc(const sf_grid_desc* d) {
    if (!d) return fail(SF_E_ARG, "sf_grid_create: null description");
    int nact = 0;
    for (int a = 0; a < 3; ++a) {
        if (d->active[a] != 0 && d->active[a] != 1)
            return fail(SF_E_ARG, "active[] entries must be 0 or 1");
        if (d->active[a]) {
            ++nact;
            if (d->n[a] < 4) return fail(SF_E_ARG, "an active axis needs at least 4 nodes");
            if (d->bc[a] < SF_BC_TRANSMISSIVE || d->bc[a] > SF_BC_NONE)
                return fail(SF_E_ARG, "unknown boundary kind");
        } else if (d->n[a] != 1) {
            return fail(SF_E_ARG, "an inactive axis must have exactly one node");
        }
        if (!(d->d[a] > 0.0) || !std::isfinite(d->d[a]))
            return fail(SF_E_ARG, "node spacing must be positive and finite");
    }
    if (nact == 0) return fail(SF_E_ARG, "at least one axis must be active");
    if (!d->active[0]) return fail(SF_E_ARG, "the x axis must be active");
    if (d->active[2] && !d->active[1])
        return fail(SF_E_ARG, "a z-active grid must also be y-active");
    if (!(d->gamma > 1.0) || !std::isfinite(d->gamma)) return fail(SF_E_ARG, "gamma must exceed 1");
    if (d->scheme < SF_SCHEME_WENO5 || d->scheme > SF_SCHEME_WENO7)
        return fail(SF_E_ARG, "unknown reconstruction scheme");
    if (d->flux < SF_FLUX_RUSANOV || d->flux > SF_FLUX_LAX_WENDROFF)
        return fail(SF_E_ARG, "unknown face flux");
    if (d->time != SF_TIME_EULER && d->time != SF_TIME_SSPRK3 && d->time != SF_TIME_RK4)
        return fail(SF_E_ARG, "unknown time recipe");
    const long long N = (long long)d->n[0] * d->n[1] * d->n[2];
    if (N > (1LL << 31)) return fail(SF_E_ARG, "grid exceeds 2^31 nodes");
    // padded nodes (ghost layers of the widest halo) index in 32 bits; no
    // grid that fits HBM with its four state buffers comes near the limit
    long long P = 1;
    for (int a = 0; a < 3; ++a) P *= a == 0 ? d->n[0] + 36 : d->active[a] ? d->n[a] + 8 : 1;
    if (P >= (1LL << 31)) return fail(SF_E_ARG, "grid exceeds 2^31 padded nodes");
    // the x sweep numbers faces in 32 bits
    if ((long long)(d->n[0] + 1) * d->n[1] * d->n[2] >= (1LL << 32) - 256)
        return fail(SF_E_ARG, "grid has 2^32 - 2^8 or more x faces");
    return 0;
}

unsigned blocks_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// Start / stop of a timed section on the grid's stream (no-ops unless
// profiling is enabled).
int prof_start(sf_grid* g, int kernel) {
    if (!g->profiling) return -1;
    const size_t need = 2 * (g->ev_used.size() + 1);
    while (g->ev_pool.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        g->ev_pool.push_back(e);
    }
    const int idx = (int)(2 * g->ev_used.size());
    g->ev_used.push_back({kernel, idx});
    (void)hipEventRecord(g->ev_pool[idx], g->stream);
    return idx;
}

void prof_stop(sf_grid* g, int idx) {
    if (idx >= 0) (void)hipEventRecord(g->ev_pool[idx + 1], g->stream);
}

// Ghost fill in the order of Boundary::Applicator::apply (density and
// momentum for every transmissive set, then energy from pressure, axis by
// axis), then the periodic axes.  `all` = 0 skips the periodic axes whose
// ghosts the final sweep keeps as copies of the interior (G.images): a
// buffer written by a stage needs no fill for them.
int fill_ghosts(sf_grid* g, double* q, int all, long long tag) {
    const GridDev& G = g->G;
    const int ev = prof_start(g, SF_KERNEL_GHOST);
    auto plane = [&](int a) {
        return 2LL * G.n[(a + 1) % 3] * G.n[(a + 2) % 3];
    };
    auto zero_gradient = [&](int a) {
        return G.active[a] && (G.bc[a] == SF_BC_TRANSMISSIVE || G.bc[a] == SF_BC_NONE_LO ||
                               G.bc[a] == SF_BC_NONE_HI);
    };
    for (int a = 0; a < 3; ++a) {
        if (!zero_gradient(a)) continue;
        const int per = a == 0 ? fill_lanes<0>(G, true) : 1;
        const unsigned nb = blocks_for(plane(a) * per, 256);
        if (a == 0) k_bc_copy4<0><<<nb, 256, 0, g->stream>>>(G, q, g->flags);
        if (a == 1) k_bc_copy4<1><<<nb, 256, 0, g->stream>>>(G, q, g->flags);
        if (a == 2) k_bc_copy4<2><<<nb, 256, 0, g->stream>>>(G, q, g->flags);
    }
    for (int a = 0; a < 3; ++a) {
        if (!zero_gradient(a)) continue;
        const int per = a == 0 ? fill_lanes<0>(G, false) : 1;
        const unsigned nb = blocks_for(plane(a) * per, 256);
        if (a == 0) k_bc_energy<0><<<nb, 256, 0, g->stream>>>(G, q, g->fault, tag);
        if (a == 1) k_bc_energy<1><<<nb, 256, 0, g->stream>>>(G, q, g->fault, tag);
        if (a == 2) k_bc_energy<2><<<nb, 256, 0, g->stream>>>(G, q, g->fault, tag);
    }
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a] || G.bc[a] != SF_BC_PERIODIC) continue;
        if (!all && (G.images & (1 << a))) continue;
        const unsigned nb = blocks_for(plane(a), 256);
        if (a == 0) k_bc_periodic<0><<<nb, 256, 0, g->stream>>>(G, q, g->flags);
        if (a == 1) k_bc_periodic<1><<<nb, 256, 0, g->stream>>>(G, q, g->flags);
        if (a == 2) k_bc_periodic<2><<<nb, 256, 0, g->stream>>>(G, q, g->flags);
    }
    prof_stop(g, ev);
    SF_HIP(hipGetLastError());
    return 0;
}

// Sweeps run z, y, then x: the x sweep (one face per lane) applies the
// stage update, because its register footprint is the smallest and the
// update adds to it.  The flux differences are therefore summed as
// (dz + dy) + dx rather than the reference's (dx + dy) + dz, which changes
// the last bit at most.
int run_sweeps(sf_grid* g, const double* q, const StageArgs& A) {
    SweepCall c{g->G, q, g->S, A, g->stream, g->march_seg, g->cus, g->x_walk};
    int axes[3], na = 0;
    for (int a = 2; a >= 0; --a)
        if (g->G.active[a]) axes[na++] = a;
    for (int t = 0; t < na; ++t) {
        const int mode = (t == na - 1) ? kFinal : (t == 0) ? kCopyS : kAccumS;
        const int ev = prof_start(g, SF_KERNEL_SWEEP_X + axes[t]);
        int rc;
        switch (g->desc.scheme) {
        case SF_SCHEME_TENO5: rc = launch_sweep<1>(c, g->desc.flux, axes[t], mode); break;
        case SF_SCHEME_WENO3: rc = launch_sweep<2>(c, g->desc.flux, axes[t], mode); break;
        case SF_SCHEME_WENO7: rc = launch_sweep<3>(c, g->desc.flux, axes[t], mode); break;
        default: rc = launch_sweep<0>(c, g->desc.flux, axes[t], mode); break;
        }
        prof_stop(g, ev);
        if (rc) return fail(SF_E_STATE, "run_sweeps: no kernel for this sweep axis and mode");
    }
    SF_HIP(hipGetLastError());
    return 0;
}

int set_device(const sf_grid* g) {
    SF_HIP(hipSetDevice(g->desc.device));
    return 0;
}

int ensure_hist(sf_grid* g, int n) {
    if (n <= g->dt_hist_cap) return 0;
    if (g->dt_hist) SF_HIP(hipFree(g->dt_hist));
    g->dt_hist = nullptr;
    SF_HIP(hipMalloc(&g->dt_hist, sizeof(double) * (size_t)n));
    g->dt_hist_cap = n;
    return 0;
}

// Buffers the current state moves by in one step (plan_buffers).
int rotation(int time) { return time == SF_TIME_RK4 ? 2 : 1; }

// Tags of the launches of a step (Fault::tag): the open step's stages and
// fills, the fill that prepares or commits a step, the CFL scan before it.
long long stage_tag(const sf_grid* g, int s) { return (g->seq - 1) * 8 + s; }
long long next_tag(const sf_grid* g, int phase) { return g->seq * 8 + phase; }
long long open_tag(const sf_grid* g, int phase) { return (g->seq - 1) * 8 + phase; }

// Wait for the grid's work and read its fault record.  With flags raised,
// every launch after the faulting one returned at entry, so the buffer that
// held the state at the start of the faulting step still holds it: it
// becomes the current state again, and the grid's clock (which no step
// committed after the fault) counts the steps before it.  Reports the
// 1-based number of the faulting step on the clock and its phase.
int check_flags(sf_grid* g, const char* where, uint32_t* flags_out = nullptr,
                long long* step_out = nullptr, int* phase_out = nullptr) {
    Fault h{};
    Clock c{};
    SF_HIP(hipMemcpyAsync(&h, g->fault, sizeof h, hipMemcpyDeviceToHost, g->stream));
    SF_HIP(hipMemcpyAsync(&c, g->clock, sizeof c, hipMemcpyDeviceToHost, g->stream));
    SF_HIP(hipStreamSynchronize(g->stream));
    if (flags_out) *flags_out = h.flags;
    if (step_out) *step_out = h.flags ? c.steps + 1 : 0;
    if (phase_out) *phase_out = h.flags ? (int)(h.tag & 7) : -1;
    if (!h.flags) {
        g->base_seq = g->seq;
        g->base_cur = g->cur;
        return 0;
    }
    const long long seq = h.tag >> 3;
    const int phase = (int)(h.tag & 7);
    if (seq >= g->base_seq)
        g->cur = (int)((g->base_cur + (seq - g->base_seq) * rotation(g->desc.time)) % 3);
    g->base_seq = g->seq;
    g->base_cur = g->cur;
    g->step_open = 0;
    g->ghosts_valid = 0;
    g->dtmin_valid = 0;
    char what[32];
    if (phase < 4) std::snprintf(what, sizeof what, "stage %d", phase + 1);
    else std::snprintf(what, sizeof what, "%s", phase == kPhaseScan ? "CFL scan" : "ghost fill");
    char msg[320];
    std::snprintf(msg, sizeof msg,
                  "%s: device raised numerical fault flags 0x%x "
                  "(1 state, 2 face average, 4 TENO cutoff, 8 flux, 16 dt) in step %lld, %s; "
                  "the grid holds the state at the start of that step",
                  where, h.flags, c.steps + 1, what);
    return fail(SF_E_NUMERIC, msg);
}

// The next step's CFL scan: folded into the last stage's update when
// possible, otherwise a standalone pass.
int scan_dt(sf_grid* g, double cfl) {
    if (!(g->dtmin_valid && g->dtmin_cfl == cfl)) {
        k_set_limits<<<1, 64, 0, g->stream>>>(g->clock, g->max_dt, g->limit, 1);
        const long long N = (long long)g->G.n[0] * g->G.n[1] * g->G.n[2];
        const unsigned nb = std::min<unsigned>(blocks_for(N, 256), 8192u);
        const int ev = prof_start(g, SF_KERNEL_CFL);
        k_cfl<<<nb, 256, 0, g->stream>>>(g->G, g->buf[g->cur], cfl, &g->clock->dtmin, g->fault,
                                         next_tag(g, kPhaseScan));
        prof_stop(g, ev);
        SF_HIP(hipGetLastError());
    }
    g->dtmin_valid = 0;
    return 0;
}

int set_limits(sf_grid* g, double max_dt, double limit) {
    g->max_dt = max_dt;
    g->limit = limit;
    k_set_limits<<<1, 64, 0, g->stream>>>(g->clock, max_dt, limit, 0);
    SF_HIP(hipGetLastError());
    return 0;
}

int do_stage(sf_grid* g, int s, int fold_dt) {
    const int in = g->stage_in[s], out = g->stage_out[s];
    if (s == 0 && !g->ghosts_valid) {
        if (int rc = fill_ghosts(g, g->buf[in], 1, stage_tag(g, s))) return rc;
    }
    if (s > 0) {
        if (int rc = fill_ghosts(g, g->buf[in], 0, stage_tag(g, s))) return rc;
    }
    StageArgs A{};
    A.fault_tag = &g->fault->tag;
    A.tag = stage_tag(g, s);
    A.q0 = g->buf[g->cur];
    A.qout = g->buf[out];
    for (int i = 0; i < 3; ++i) A.k[i] = g->k[i];
    A.dt = &g->clock->dt;
    A.dtmin = &g->clock->dtmin;
    A.flags = g->flags;
    A.time = g->desc.time;
    A.stage = s;
    if (g->desc.time == SF_TIME_SSPRK3) {
        static const double base[3] = {0.0, 0.75, 1.0 / 3.0};
        static const double incr[3] = {1.0, 0.25, 2.0 / 3.0};
        A.base = base[s];
        A.incr = incr[s];
    } else if (g->desc.time == SF_TIME_RK4) {
        static const double incr[4] = {0.5, 0.5, 1.0, 0.0};
        A.incr = incr[s];
    } else {
        A.incr = 1.0;
    }
    const bool last = s == nstages(g->desc.time) - 1;
    A.reduce_dt = (last && fold_dt) ? 1 : 0;
    A.cfl = g->step_cfl;
    return run_sweeps(g, g->buf[in], A);
}

// Stage buffers of a step.  No stage writes the buffer that holds the
// step-start state (a), so that state survives a fault in any stage; the
// step's result lands in another buffer, which becomes the current one.
void plan_buffers(sf_grid* g) {
    const int a = g->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    switch (g->desc.time) {
    case SF_TIME_EULER:
        g->stage_in[0] = a; g->stage_out[0] = b;
        break;
    case SF_TIME_SSPRK3:
        g->stage_in[0] = a; g->stage_out[0] = b;
        g->stage_in[1] = b; g->stage_out[1] = c;
        g->stage_in[2] = c; g->stage_out[2] = b;
        break;
    default:
        g->stage_in[0] = a; g->stage_out[0] = b;
        g->stage_in[1] = b; g->stage_out[1] = c;
        g->stage_in[2] = c; g->stage_out[2] = b;
        g->stage_in[3] = b; g->stage_out[3] = c;
        break;
    }
}

// Grid constants of a checked description (no device work).
void set_geometry(sf_grid* g, const sf_grid_desc* desc) {
    g->desc = *desc;
    GridDev& G = g->G;
    for (int a = 0; a < 3; ++a) {
        G.n[a] = desc->n[a];
        // ghost depth = the scheme's halo (builtInConvectionRecipe,
        // SF_termRecipe.h:217-218)
        G.g[a] = desc->active[a] ? (desc->scheme == SF_SCHEME_WENO3 ? 2
                                    : desc->scheme == SF_SCHEME_WENO7 ? 4 : 3) : 0;
        G.m[a] = G.n[a] + 2 * G.g[a];
        if (a == 0) {
            // rows start on 128-byte lines: node 0 of every row sits 16
            // doubles into it (its ghosts just before), and the row length
            // is a multiple of 16, so the marches' 64-column blocks read and
            // write whole cache lines; the row holds at least one node past
            // the last ghost, so the line-walking x sweep can copy the last
            // ghost as the first of a 16-byte pair
            G.xo = 16;
            G.m[0] = (G.xo + G.n[0] + G.g[0] + 1 + 15) / 16 * 16;
        }
        G.active[a] = desc->active[a];
        G.bc[a] = desc->bc[a];
        G.d[a] = desc->d[a];
        // A node has at most one periodic copy in the ghost layers only
        // where n >= 2g; thinner periodic axes are filled between stages
        // (k_bc_periodic takes the node index modulo n).
        G.images |= (desc->active[a] && desc->bc[a] == SF_BC_PERIODIC && G.n[a] >= 2 * G.g[a])
                        ? (1 << a) : 0;
    }
    G.div_line = make_fastdiv((unsigned)G.n[0] + 1u);
    G.div_n1 = make_fastdiv((unsigned)G.n[1]);
    // A zeroGradient end rewrites the boundary nodes of its axis in the
    // ghost fill after the stage; copies written by the stage would miss
    // that, so the copies are used only where no axis has such an end.
    for (int a = 0; a < 3; ++a)
        if (desc->active[a] && (desc->bc[a] == SF_BC_TRANSMISSIVE ||
                                desc->bc[a] == SF_BC_NONE_LO || desc->bc[a] == SF_BC_NONE_HI))
            G.images = 0;
    for (int a = 0; a < 3; ++a) {
        G.solved_lo[a] = zg_end(G, a, 0) ? 1 : 0;
        G.solved_hi[a] = G.n[a] - 1 - (zg_end(G, a, 1) ? 1 : 0);
    }
    G.stride[0] = 1;
    G.stride[1] = G.m[0];
    G.stride[2] = (long long)G.m[0] * G.m[1];
    G.P = (long long)G.m[0] * G.m[1] * G.m[2];
    for (int a = 0; a < 3; ++a) G.ustride[a] = (unsigned)G.stride[a];
    for (int a = 0; a < 3; ++a) {
        const bool img = (G.images >> a) & 1;
        G.img_lo[a] = img ? G.g[a] : 0;
        G.img_span[a] = img ? (unsigned)(G.n[a] - 2 * G.g[a]) : 0xFFFFFFFFu;
        G.img_off[a] = img ? (long long)G.n[a] * G.stride[a] : 0;
    }
    G.gamma = desc->gamma;
    // Uniform Cartesian metrics exactly as the reference forms them:
    // Jac = 1/J, metric = 1/d, face cofactor = 0.5*(metric/Jac + metric/Jac),
    // area = |cofactor| (SF_mesh.cpp:320-338, SF_canonicalFace.h:87-108).
    // The metric of an inactive axis is stored as 0: its CFL term, which
    // the reference skips, then adds an exact zero.
    G.jac = 1.0 / (desc->d[0] * desc->d[1] * desc->d[2]);
    for (int a = 0; a < 3; ++a) {
        G.metric[a] = desc->active[a] ? 1.0 / desc->d[a] : 0.0;
        const double c = (1.0 / desc->d[a]) / G.jac;
        const double cof = 0.5 * (c + c);
        G.area[a] = std::sqrt(cof * cof + 0.0 * 0.0 + 0.0 * 0.0);
    }
}

int grid_free(sf_grid* g) {
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (double* p : g->buf) if (p) (void)hipFree(p);
    for (double* p : g->k) if (p) (void)hipFree(p);
    if (g->S) (void)hipFree(g->S);
    if (g->clock) (void)hipFree(g->clock);
    if (g->fault) (void)hipFree(g->fault);
    if (g->dt_hist) (void)hipFree(g->dt_hist);
    if (g->stage_io) (void)hipFree(g->stage_io);
    for (hipEvent_t e : g->ev_pool) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

}  // namespace

extern "C" {

int sf_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return SF_E_ARG;
    std::snprintf(buf, len, "%s", g_last_error.c_str());
    return 0;
}

int sf_version(int* major, int* minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return 0;
}

int sf_grid_create(const sf_grid_desc* desc, sf_grid** out) {
    if (!out) return fail(SF_E_ARG, "sf_grid_create: null output handle");
    *out = nullptr;
    if (int rc = check_desc(desc)) return rc;
    int ndev = 0;
    SF_HIP(hipGetDeviceCount(&ndev));
    if (desc->device < 0 || desc->device >= ndev)
        return fail(SF_E_ARG, "sf_grid_create: device ordinal out of range");
    SF_HIP(hipSetDevice(desc->device));
    sf_grid* g = new (std::nothrow) sf_grid;
    if (!g) return fail(SF_E_MEMORY, "sf_grid_create: host allocation failed");
    set_geometry(g, desc);
    const GridDev& G = g->G;
    if (hipDeviceGetAttribute(&g->cus, hipDeviceAttributeMultiprocessorCount, desc->device) !=
            hipSuccess || g->cus <= 0)
        g->cus = 256;   // MI355X's CU count; only the launch shapes depend on it
    const size_t state = sizeof(double) * 5 * (size_t)G.P;
    const size_t N = (size_t)desc->n[0] * desc->n[1] * desc->n[2];
    auto alloc = [&](double** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(SF_E_MEMORY, std::string("sf_grid_create: hipMalloc of ") +
                                         std::to_string(bytes) + " bytes failed: " +
                                         hipGetErrorString(e));
        }
        g->bytes += bytes;
        return 0;
    };
    int rc = 0;
    for (int i = 0; i < 3 && !rc; ++i) rc = alloc(&g->buf[i], state);
    if (!rc) rc = alloc(&g->S, state);
    if (!rc && desc->time == SF_TIME_RK4)
        for (int i = 0; i < 3 && !rc; ++i) rc = alloc(&g->k[i], state);
    if (!rc) rc = alloc(&g->stage_io, sizeof(double) * 5 * N);
    if (!rc && hipMalloc(&g->clock, sizeof(Clock)) != hipSuccess)
        rc = fail(SF_E_MEMORY, "sf_grid_create: clock allocation failed");
    if (!rc && hipMalloc(&g->fault, sizeof(Fault)) != hipSuccess)
        rc = fail(SF_E_MEMORY, "sf_grid_create: fault record allocation failed");
    if (!rc) g->flags = &g->fault->flags;
    if (!rc && hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(SF_E_HIP, "sf_grid_create: stream creation failed");
    if (rc) {
        grid_free(g);
        return rc;
    }
    Clock c{};
    c.dt = 0.0;
    c.time = 0.0;
    c.steps = 0;
    c.dtmin = kInfBits;
    c.max_dt = std::numeric_limits<double>::max();
    c.limit = std::numeric_limits<double>::max();
    Fault f{0u, 0u, -1};
    if (hipMemcpy(g->clock, &c, sizeof c, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->fault, &f, sizeof f, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(g->buf[0], 0, state) != hipSuccess ||
        hipMemset(g->buf[1], 0, state) != hipSuccess ||
        hipMemset(g->buf[2], 0, state) != hipSuccess ||
        hipMemset(g->S, 0, state) != hipSuccess) {
        grid_free(g);
        return fail(SF_E_HIP, "sf_grid_create: initialisation copy failed");
    }
    *out = g;
    return 0;
}

int sf_grid_destroy(sf_grid* g) {
    if (!g) return 0;
    set_device(g);
    return grid_free(g);
}

int sf_grid_device_bytes(const sf_grid* g, size_t* bytes) {
    if (!g || !bytes) return fail(SF_E_ARG, "sf_grid_device_bytes: null argument");
    *bytes = g->bytes;
    return 0;
}

int sf_grid_stream(sf_grid* g, void** stream) {
    if (!g || !stream) return fail(SF_E_ARG, "sf_grid_stream: null argument");
    *stream = (void*)g->stream;
    return 0;
}

// A new state starts with a clear fault record.
static int upload_from(sf_grid* g, const double* dsrc) {
    const GridDev& G = g->G;
    const long long N = (long long)G.n[0] * G.n[1] * G.n[2];
    const unsigned nb = std::min<unsigned>(blocks_for(N, 256), 65535u);
    k_clear_fault<<<1, 64, 0, g->stream>>>(g->fault);
    g->base_seq = g->seq;
    g->base_cur = g->cur;
    k_pack<<<nb, 256, 0, g->stream>>>(G, dsrc, g->buf[g->cur], 1);
    // ends without a condition: every state buffer's ghost layers there hold
    // the uploaded end nodes from now on
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a] || G.bc[a] < SF_BC_NONE_LO) continue;
        const unsigned np = blocks_for(2LL * G.n[(a + 1) % 3] * G.n[(a + 2) % 3], 256);
        for (int b = 0; b < 3; ++b) {
            if (a == 0) k_bc_hold<0><<<np, 256, 0, g->stream>>>(G, g->buf[g->cur], g->buf[b]);
            if (a == 1) k_bc_hold<1><<<np, 256, 0, g->stream>>>(G, g->buf[g->cur], g->buf[b]);
            if (a == 2) k_bc_hold<2><<<np, 256, 0, g->stream>>>(G, g->buf[g->cur], g->buf[b]);
        }
    }
    SF_HIP(hipGetLastError());
    g->ghosts_valid = 0;
    g->dtmin_valid = 0;
    g->step_open = 0;
    return 0;
}

int sf_state_upload(sf_grid* g, const double* soa) {
    if (!g || !soa) return fail(SF_E_ARG, "sf_state_upload: null argument");
    if (int rc = set_device(g)) return rc;
    const size_t N = (size_t)g->G.n[0] * g->G.n[1] * g->G.n[2];
    SF_HIP(hipMemcpyAsync(g->stage_io, soa, sizeof(double) * 5 * N, hipMemcpyHostToDevice,
                          g->stream));
    if (int rc = upload_from(g, g->stage_io)) return rc;
    SF_HIP(hipStreamSynchronize(g->stream));
    return 0;
}

int sf_state_upload_device(sf_grid* g, const double* soa_device) {
    if (!g || !soa_device) return fail(SF_E_ARG, "sf_state_upload_device: null argument");
    if (int rc = set_device(g)) return rc;
    return upload_from(g, soa_device);
}

static int download_to(sf_grid* g, double* ddst) {
    const long long N = (long long)g->G.n[0] * g->G.n[1] * g->G.n[2];
    const unsigned nb = std::min<unsigned>(blocks_for(N, 256), 65535u);
    k_pack<<<nb, 256, 0, g->stream>>>(g->G, g->buf[g->cur], ddst, 0);
    SF_HIP(hipGetLastError());
    return 0;
}

int sf_state_download(sf_grid* g, double* soa) {
    if (!g || !soa) return fail(SF_E_ARG, "sf_state_download: null argument");
    if (int rc = set_device(g)) return rc;
    if (int rc = download_to(g, g->stage_io)) return rc;
    const size_t N = (size_t)g->G.n[0] * g->G.n[1] * g->G.n[2];
    SF_HIP(hipMemcpyAsync(soa, g->stage_io, sizeof(double) * 5 * N, hipMemcpyDeviceToHost,
                          g->stream));
    SF_HIP(hipStreamSynchronize(g->stream));
    return 0;
}

int sf_state_download_device(sf_grid* g, double* soa_device) {
    if (!g || !soa_device) return fail(SF_E_ARG, "sf_state_download_device: null argument");
    if (int rc = set_device(g)) return rc;
    return download_to(g, soa_device);
}

int sf_step_prepare(sf_grid* g) {
    if (!g) return fail(SF_E_ARG, "sf_step_prepare: null grid");
    if (int rc = set_device(g)) return rc;
    if (g->step_open) return fail(SF_E_STATE, "sf_step_prepare: a step is open");
    if (!g->ghosts_valid) {
        if (int rc = fill_ghosts(g, g->buf[g->cur], 1, next_tag(g, kPhaseFill))) return rc;
        g->ghosts_valid = 1;
    }
    return 0;
}

int sf_compute_dt(sf_grid* g, double cfl, double max_dt, double limit, double* dt_out) {
    if (!g) return fail(SF_E_ARG, "sf_compute_dt: null grid");
    if (!(cfl > 0.0) || !std::isfinite(cfl)) return fail(SF_E_ARG, "sf_compute_dt: invalid CFL");
    if (!(max_dt > 0.0) || !(limit > 0.0))
        return fail(SF_E_ARG, "sf_compute_dt: limits must be positive");
    if (int rc = set_device(g)) return rc;
    if (g->step_open) return fail(SF_E_STATE, "sf_compute_dt: a step is open");
    if (int rc = set_limits(g, max_dt, limit)) return rc;
    if (int rc = scan_dt(g, cfl)) return rc;
    k_finalize_dt<<<1, 64, 0, g->stream>>>(g->clock, nullptr, 0, g->flags);
    SF_HIP(hipGetLastError());
    g->step_cfl = cfl;
    if (dt_out) {
        SF_HIP(hipMemcpyAsync(dt_out, &g->clock->dt, sizeof(double), hipMemcpyDeviceToHost,
                              g->stream));
        SF_HIP(hipStreamSynchronize(g->stream));
        if (int rc = check_flags(g, "sf_compute_dt")) return rc;
    }
    return 0;
}

int sf_set_dt(sf_grid* g, double dt) {
    if (!g) return fail(SF_E_ARG, "sf_set_dt: null grid");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(SF_E_ARG, "sf_set_dt: invalid dt");
    if (int rc = set_device(g)) return rc;
    if (g->step_open) return fail(SF_E_STATE, "sf_set_dt: a step is open");
    k_set_dt<<<1, 64, 0, g->stream>>>(g->clock, dt);
    SF_HIP(hipGetLastError());
    return 0;
}

int sf_step_begin(sf_grid* g) {
    if (!g) return fail(SF_E_ARG, "sf_step_begin: null grid");
    if (g->step_open) return fail(SF_E_STATE, "sf_step_begin: a step is already open");
    plan_buffers(g);
    ++g->seq;
    g->step_open = 1;
    g->next_stage = 0;
    return 0;
}

int sf_stage_count(const sf_grid* g, int* count) {
    if (!g || !count) return fail(SF_E_ARG, "sf_stage_count: null argument");
    *count = nstages(g->desc.time);
    return 0;
}

int sf_stage(sf_grid* g, int stage) {
    if (!g) return fail(SF_E_ARG, "sf_stage: null grid");
    if (!g->step_open || stage != g->next_stage || stage >= nstages(g->desc.time))
        return fail(SF_E_STATE, "sf_stage: stage out of order or no open step");
    if (int rc = set_device(g)) return rc;
    if (int rc = do_stage(g, stage, g->fold_cfl)) return rc;
    g->next_stage++;
    return 0;
}

int sf_step_commit(sf_grid* g) {
    if (!g) return fail(SF_E_ARG, "sf_step_commit: null grid");
    if (!g->step_open || g->next_stage != nstages(g->desc.time))
        return fail(SF_E_STATE, "sf_step_commit: stages incomplete");
    if (int rc = set_device(g)) return rc;
    g->cur = g->stage_out[nstages(g->desc.time) - 1];
    if (int rc = fill_ghosts(g, g->buf[g->cur], 0, open_tag(g, kPhaseFill))) return rc;
    k_commit_clock<<<1, 64, 0, g->stream>>>(g->clock, g->flags);
    SF_HIP(hipGetLastError());
    g->ghosts_valid = 1;
    g->dtmin_valid = g->fold_cfl;
    g->dtmin_cfl = g->step_cfl;
    g->step_open = 0;
    return 0;
}

// nsteps whole steps queued on the grid's stream, dt = min(CFL dt, max_dt)
// each; hist (device memory, may be null) receives the dt values.
static int queue_steps(sf_grid* g, int nsteps, double cfl, double max_dt, double* hist) {
    if (int rc = set_limits(g, max_dt, std::numeric_limits<double>::max())) return rc;
    g->fold_cfl = 1;
    g->step_cfl = cfl;
    for (int s = 0; s < nsteps; ++s) {
        if (int rc = sf_step_prepare(g)) return rc;
        if (int rc = scan_dt(g, cfl)) return rc;
        k_finalize_dt<<<1, 64, 0, g->stream>>>(g->clock, hist, s, g->flags);
        SF_HIP(hipGetLastError());
        if (int rc = sf_step_begin(g)) return rc;
        for (int st = 0; st < nstages(g->desc.time); ++st)
            if (int rc = sf_stage(g, st)) return rc;
        if (int rc = sf_step_commit(g)) return rc;
    }
    g->fold_cfl = 0;
    return 0;
}

int sf_run(sf_grid* g, int nsteps, double cfl, double max_dt, double* dt_seq) {
    if (!g) return fail(SF_E_ARG, "sf_run: null grid");
    if (nsteps < 0) return fail(SF_E_ARG, "sf_run: negative step count");
    if (!(cfl > 0.0) || !std::isfinite(cfl)) return fail(SF_E_ARG, "sf_run: invalid CFL");
    if (!(max_dt > 0.0)) return fail(SF_E_ARG, "sf_run: max_dt must be positive");
    if (g->step_open) return fail(SF_E_STATE, "sf_run: a step is open");
    if (int rc = set_device(g)) return rc;
    if (dt_seq && nsteps > 0) {
        if (int rc = ensure_hist(g, nsteps)) return rc;
    }
    if (int rc = queue_steps(g, nsteps, cfl, max_dt, dt_seq ? g->dt_hist : nullptr)) return rc;
    if (dt_seq && nsteps > 0) {
        SF_HIP(hipMemcpyAsync(dt_seq, g->dt_hist, sizeof(double) * (size_t)nsteps,
                              hipMemcpyDeviceToHost, g->stream));
    }
    return check_flags(g, "sf_run");
}

int sf_clock(sf_grid* g, double* time, long long* steps) {
    if (!g) return fail(SF_E_ARG, "sf_clock: null grid");
    if (int rc = set_device(g)) return rc;
    Clock c;
    SF_HIP(hipMemcpyAsync(&c, g->clock, sizeof c, hipMemcpyDeviceToHost, g->stream));
    SF_HIP(hipStreamSynchronize(g->stream));
    if (time) *time = c.time;
    if (steps) *steps = c.steps;
    return 0;
}

int sf_residual(sf_grid* g, double* soa_out) {
    if (!g || !soa_out) return fail(SF_E_ARG, "sf_residual: null argument");
    if (g->step_open) return fail(SF_E_STATE, "sf_residual: a step is open");
    if (int rc = set_device(g)) return rc;
    if (int rc = sf_step_prepare(g)) return rc;
    const int scratch = (g->cur + 1) % 3;
    SF_HIP(hipMemsetAsync(g->buf[scratch], 0, sizeof(double) * 5 * (size_t)g->G.P, g->stream));
    StageArgs A{};
    A.q0 = g->buf[g->cur];   // a valid state, though the residual-only update reads none
    A.qout = g->buf[scratch];
    A.flags = g->flags;
    A.time = 3;  // residual only
    A.dt = &g->clock->dt;
    A.dtmin = &g->clock->dtmin;
    A.fault_tag = &g->fault->tag;
    A.tag = next_tag(g, 0);
    if (int rc = run_sweeps(g, g->buf[g->cur], A)) return rc;
    const long long N = (long long)g->G.n[0] * g->G.n[1] * g->G.n[2];
    const unsigned nb = std::min<unsigned>(blocks_for(N, 256), 65535u);
    k_pack<<<nb, 256, 0, g->stream>>>(g->G, g->buf[scratch], g->stage_io, 0);
    SF_HIP(hipGetLastError());
    SF_HIP(hipMemcpyAsync(soa_out, g->stage_io, sizeof(double) * 5 * (size_t)N,
                          hipMemcpyDeviceToHost, g->stream));
    SF_HIP(hipStreamSynchronize(g->stream));
    return check_flags(g, "sf_residual");
}

int sf_error_flags(sf_grid* g, uint32_t* flags, int clear) {
    if (!g || !flags) return fail(SF_E_ARG, "sf_error_flags: null argument");
    if (int rc = set_device(g)) return rc;
    unsigned h = 0;
    SF_HIP(hipMemcpyAsync(&h, g->flags, sizeof h, hipMemcpyDeviceToHost, g->stream));
    SF_HIP(hipStreamSynchronize(g->stream));
    *flags = h;
    if (clear) {
        k_clear_fault<<<1, 64, 0, g->stream>>>(g->fault);
        SF_HIP(hipGetLastError());
        SF_HIP(hipStreamSynchronize(g->stream));
    }
    return 0;
}

int sf_fault_check(sf_grid* g, uint32_t* flags, long long* step, int* phase) {
    if (!g) return fail(SF_E_ARG, "sf_fault_check: null grid");
    if (int rc = set_device(g)) return rc;
    return check_flags(g, "sf_fault_check", flags, step, phase);
}

int sf_grid_set_x_walk(sf_grid* g, int mode) {
    if (!g) return fail(SF_E_ARG, "sf_grid_set_x_walk: null grid");
    if (mode < -1 || mode > 1)
        return fail(SF_E_ARG, "sf_grid_set_x_walk: -1 (automatic), 0 (flat tiles) or 1 (lines)");
    g->x_walk = mode;
    return 0;
}

int sf_grid_set_march_segment(sf_grid* g, int nodes) {
    if (!g) return fail(SF_E_ARG, "sf_grid_set_march_segment: null grid");
    if (nodes != 0 && (nodes < 1 || nodes > 4096))
        return fail(SF_E_ARG, "sf_grid_set_march_segment: 0 (automatic) or 1..4096 nodes");
    g->march_seg = nodes;
    return 0;
}

int sf_profile_enable(sf_grid* g, int enable) {
    if (!g) return fail(SF_E_ARG, "sf_profile_enable: null grid");
    if (int rc = set_device(g)) return rc;
    SF_HIP(hipStreamSynchronize(g->stream));
    g->ev_used.clear();
    g->profiling = enable ? 1 : 0;
    return 0;
}

int sf_profile_read(sf_grid* g, int kernel, double* total_ms, long long* launches) {
    if (!g || kernel < 0 || kernel >= SF_KERNEL_COUNT)
        return fail(SF_E_ARG, "sf_profile_read: invalid argument");
    if (int rc = set_device(g)) return rc;
    SF_HIP(hipStreamSynchronize(g->stream));
    double sum = 0.0;
    long long n = 0;
    for (const auto& u : g->ev_used) {
        if (u.first != kernel) continue;
        float ms = 0.0f;
        SF_HIP(hipEventElapsedTime(&ms, g->ev_pool[u.second], g->ev_pool[u.second + 1]));
        sum += ms;
        ++n;
    }
    if (total_ms) *total_ms = sum;
    if (launches) *launches = n;
    return 0;
}

int sf_synchronize(sf_grid* g) {
    if (!g) return fail(SF_E_ARG, "sf_synchronize: null grid");
    if (int rc = set_device(g)) return rc;
    SF_HIP(hipStreamSynchronize(g->stream));
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// A whole run that allocates nothing (sf_run_static): the grid's buffers,
// clock and fault record live in one file-scope device array per GPU.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kStaticBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kStaticBytes];

// Device address of `scratch` on each device (its name is only a host-side
// stand-in), taken once per device with that device current.
constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int static_scratch(int device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(SF_E_ARG, "sf_run_static: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SF_E_HIP, std::string("sf_run_static: hipGetSymbolAddress: ") +
                                      hipGetErrorString(e));
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[device];
    return 0;
}

size_t round256(size_t bytes) { return (bytes + 255) & ~size_t(255); }

__global__ __launch_bounds__(256) void k_zero(double* __restrict__ p, long long n) {
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < n;
         u += (long long)gridDim.x * blockDim.x)
        p[u] = 0.0;
}

__global__ void k_init_clock(Clock* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->dt = 0.0;
        c->time = 0.0;
        c->steps = 0;
        c->dtmin = kInfBits;
        c->max_dt = std::numeric_limits<double>::max();
        c->limit = std::numeric_limits<double>::max();
    }
}

}  // namespace

extern "C" int sf_run_static(const sf_grid_desc* desc, double* soa_device, int nsteps,
                             double cfl, double max_dt, double* dt_seq_device, void* stream,
                             size_t* scratch_bytes) {
    if (!soa_device) return fail(SF_E_ARG, "sf_run_static: null state");
    if (nsteps < 0) return fail(SF_E_ARG, "sf_run_static: negative step count");
    if (!(cfl > 0.0) || !std::isfinite(cfl)) return fail(SF_E_ARG, "sf_run_static: invalid CFL");
    if (!(max_dt > 0.0)) return fail(SF_E_ARG, "sf_run_static: max_dt must be positive");
    if (int rc = check_desc(desc)) return rc;
    int ndev = 0;
    SF_HIP(hipGetDeviceCount(&ndev));
    if (desc->device < 0 || desc->device >= ndev)
        return fail(SF_E_ARG, "sf_run_static: device ordinal out of range");
    sf_grid g;
    set_geometry(&g, desc);
    if (hipDeviceGetAttribute(&g.cus, hipDeviceAttributeMultiprocessorCount, desc->device) !=
            hipSuccess || g.cus <= 0)
        g.cus = 256;
    // scratch layout: clock, fault record, then the state buffers, the flux
    // accumulator and (RK4) the three stage residuals
    const size_t state = round256(sizeof(double) * 5 * (size_t)g.G.P);
    const int nbuf = desc->time == SF_TIME_RK4 ? 7 : 4;
    const size_t head = round256(sizeof(Clock)) + round256(sizeof(Fault));
    const size_t need = head + state * nbuf;
    if (scratch_bytes) *scratch_bytes = need;
    if (need > kStaticBytes) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "sf_run_static: the grid needs %zu bytes of device scratch, more than "
                      "the %zu bytes the library holds per device; use sf_grid_create",
                      need, kStaticBytes);
        return fail(SF_E_ARG, msg);
    }
    SF_HIP(hipSetDevice(desc->device));
    unsigned char* base = nullptr;
    if (int rc = static_scratch(desc->device, &base)) return rc;
    g.stream = static_cast<hipStream_t>(stream);
    g.clock = reinterpret_cast<Clock*>(base);
    g.fault = reinterpret_cast<Fault*>(base + round256(sizeof(Clock)));
    g.flags = &g.fault->flags;
    double* bufs[7];
    for (int i = 0; i < nbuf; ++i) bufs[i] = reinterpret_cast<double*>(base + head + state * i);
    for (int i = 0; i < 3; ++i) g.buf[i] = bufs[i];
    g.S = bufs[3];
    if (nbuf == 7)
        for (int i = 0; i < 3; ++i) g.k[i] = bufs[4 + i];
    // Every call on this device shares the array and finds the last call's
    // contents: the clock, the fault record and the buffers start cleared.
    k_init_clock<<<1, 64, 0, g.stream>>>(g.clock);
    k_clear_fault<<<1, 64, 0, g.stream>>>(g.fault);
    const long long words = (long long)(state * nbuf / sizeof(double));
    k_zero<<<std::min<unsigned>(blocks_for(words, 256), 4096u), 256, 0, g.stream>>>(bufs[0], words);
    SF_HIP(hipGetLastError());
    if (int rc = upload_from(&g, soa_device)) return rc;
    if (int rc = queue_steps(&g, nsteps, cfl, max_dt, nsteps > 0 ? dt_seq_device : nullptr))
        return rc;
    const int status = check_flags(&g, "sf_run_static");
    if (status != 0 && status != SF_E_NUMERIC) return status;
    if (int rc = download_to(&g, soa_device)) return rc;
    SF_HIP(hipStreamSynchronize(g.stream));
    return status;
}
