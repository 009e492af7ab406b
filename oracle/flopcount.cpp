// flopcount.cpp -- counts the floating-point operations of the oracle's
// hot-path functions by compiling oracle/sfcpu.c with `double` replaced by
// a counting type.  TEST INFRASTRUCTURE ONLY.
//
// Convention: +, -, *, / and sqrt count 1 each (an FMA is therefore 2);
// negation, fabs, fmin/fmax and comparisons count 0; pow(x, 6) counts 3
// (x^2, x^3, x^6, as the device code forms it).
//
//   make -C oracle flopcount && oracle/_build/flopcount   -> JSON on stdout
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

struct Fl {
    double v;
    static unsigned long long count;
    Fl() = default;
    constexpr Fl(double x) : v(x) {}
    explicit operator bool() const { return v != 0.0; }
};
unsigned long long Fl::count = 0;

inline Fl operator+(Fl a, Fl b) { ++Fl::count; return Fl(a.v + b.v); }
inline Fl operator-(Fl a, Fl b) { ++Fl::count; return Fl(a.v - b.v); }
inline Fl operator*(Fl a, Fl b) { ++Fl::count; return Fl(a.v * b.v); }
inline Fl operator/(Fl a, Fl b) { ++Fl::count; return Fl(a.v / b.v); }
inline Fl operator-(Fl a) { return Fl(-a.v); }
inline Fl& operator+=(Fl& a, Fl b) { a = a + b; return a; }
inline Fl& operator-=(Fl& a, Fl b) { a = a - b; return a; }
inline bool operator<(Fl a, Fl b) { return a.v < b.v; }
inline bool operator>(Fl a, Fl b) { return a.v > b.v; }
inline bool operator<=(Fl a, Fl b) { return a.v <= b.v; }
inline bool operator>=(Fl a, Fl b) { return a.v >= b.v; }
inline bool operator==(Fl a, Fl b) { return a.v == b.v; }
inline bool operator!=(Fl a, Fl b) { return a.v != b.v; }
inline Fl sqrt(Fl a) { ++Fl::count; return Fl(std::sqrt(a.v)); }
inline Fl fabs(Fl a) { return Fl(std::fabs(a.v)); }
inline Fl fmin(Fl a, Fl b) { return Fl(std::fmin(a.v, b.v)); }
inline Fl fmax(Fl a, Fl b) { return Fl(std::fmax(a.v, b.v)); }
inline Fl pow(Fl a, Fl b) {
    Fl::count += (b.v == 6.0) ? 3 : 1;
    return Fl(std::pow(a.v, b.v));
}
inline bool isfinite(Fl a) { return std::isfinite(a.v); }

// system headers first, so the renaming below reaches only the oracle
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define double Fl
#include "sfcpu.c"
#undef double

namespace {

unsigned long long counted(void (*fn)()) {
    Fl::count = 0;
    fn();
    return Fl::count;
}

Fl g_stencil[40];
int g_scheme = 0, g_flux = 0;

void set_stencil() {
    // a smooth subsonic stencil: every face takes the HLLC star branch, no
    // eigenvalue falls in the Steger-Warming or Roe smoothing bands
    for (int s = 0; s < 8; ++s) {
        const double rho = 1.0 + 0.05 * s, u = 0.3 + 0.01 * s, v = -0.2, w = 0.1,
                     p = 1.0 + 0.02 * s;
        const double q[5] = {rho, rho * u, rho * v, rho * w,
                             p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)};
        for (int k = 0; k < 5; ++k) g_stencil[5 * s + k] = Fl(q[k]);
    }
}

void face() {
    Fl cof[3] = {Fl(0.01), Fl(0.0), Fl(0.0)};
    Fl out[5];
    if (sfo_char_face_flux_dt(g_stencil, cof, g_scheme, g_flux, Fl(1.4), Fl(0.01), Fl(0.1), out) != 0)
        std::abort();
}

}  // namespace

int main() {
    set_stencil();
    std::printf("{\n  \"convention\": \"+,-,*,/,sqrt = 1; FMA = 2; pow(x,6) = 3\",\n");
    std::printf("  \"face\": {");
    const char* sname[4] = {"weno5", "teno5", "weno3", "weno7"};
    const char* fname[6] = {"rusanov", "hllc", "steger", "laxfriedrichs", "roe", "laxwendroff"};
    bool first = true;
    for (g_scheme = 0; g_scheme < 4; ++g_scheme)
        for (g_flux = 0; g_flux < 6; ++g_flux) {
            std::printf("%s\"%s_%s\": %llu", first ? "" : ", ", sname[g_scheme], fname[g_flux],
                        counted(face));
            first = false;
        }
    std::printf("},\n");
    // per-node residual, CFL-scan and stage-update costs from whole-grid
    // passes on periodic grids of each dimension
    g_scheme = 0; g_flux = 1;
    const unsigned long long face_cost = counted(face);
    const char* tname[3] = {"euler", "ssprk3", "rk4"};
    const int nst[3] = {1, 3, 4};
    for (int dim = 1; dim <= 3; ++dim) {
        sfo_desc d{};
        const int n[3] = {16, dim >= 2 ? 8 : 1, dim >= 3 ? 8 : 1};
        unsigned long long faces = 0;
        size_t N = (size_t)n[0] * n[1] * n[2];
        for (int a = 0; a < 3; ++a) {
            d.n[a] = n[a];
            d.active[a] = a < dim;
            d.bc[a] = SFO_PERIODIC;
            d.d[a] = Fl(0.1);
            if (a < dim) faces += (unsigned long long)(N / n[a]) * (n[a] + 1);
        }
        d.gamma = Fl(1.4);
        d.scheme = 0;
        d.flux = 1;
        Fl* soa = (Fl*)std::calloc(5 * N, sizeof(Fl));
        Fl* res = (Fl*)std::calloc(5 * N, sizeof(Fl));
        for (size_t u = 0; u < N; ++u) {
            const double x = 0.1 * (u % 16), rho = 1.0 + 0.1 * std::sin(x);
            soa[u] = Fl(rho);
            soa[N + u] = Fl(0.3 * rho);
            soa[2 * N + u] = Fl(-0.2 * rho);
            soa[3 * N + u] = Fl(0.1 * rho);
            soa[4 * N + u] = Fl(1.0 / 0.4 + 0.5 * rho * 0.14);
        }
        std::printf("  \"dim%d\": {", dim);
        for (int t = 0; t < 3; ++t) {
            d.time = t;
            sfo_grid* G = sfo_grid_create(&d);
            sfo_set_state(G, soa);
            Fl::count = 0;
            sfo_residual(G, res);
            const unsigned long long rhs = Fl::count - faces * face_cost;
            Fl::count = 0;
            Fl dt(0.0);
            sfo_compute_dt(G, Fl(0.4), &dt);
            const unsigned long long dtc = Fl::count;
            Fl::count = 0;
            sfo_step(G, Fl(1e-3));
            const unsigned long long update = Fl::count - nst[t] * (faces * face_cost + rhs);
            if (t == 0)
                std::printf("\"residual_per_node\": %.4f, \"dt_per_node\": %.4f, ",
                            (double)rhs / N, (double)dtc / N);
            std::printf("\"%s_update_per_node_step\": %.4f%s", tname[t], (double)update / N,
                        t < 2 ? ", " : "");
            sfo_grid_destroy(G);
        }
        std::printf("}%s\n", dim < 3 ? "," : "");
        std::free(soa);
        std::free(res);
    }
    std::printf("}\n");
    return 0;
}
