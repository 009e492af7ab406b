// sf_face.h -- per-face FP64 arithmetic of the explicit Euler hot path for
// gfx950: characteristic WENO3 / WENO5-JS / TENO5 / WENO7 reconstruction on
// an axis-aligned face followed by a Rusanov, HLLC, Steger-Warming,
// Lax-Friedrichs, Roe or Lax-Wendroff flux.
//
// The operations are those of the reference functions cited at each
// definition (paths relative to the sonicSolver source tree).  On an
// axis-aligned face the unit normal and the tangent basis (l, m) are
// signed unit vectors, so every product with one of their zero components
// is dropped and every product with +-1 becomes a sign; for finite inputs
// this changes nothing but the sign of a zero.  FP64 division is a long
// instruction sequence on CDNA, so where the reference divides several
// times by one quantity the code takes one reciprocal, and the WENO
// weights are normalised with a single division (each definition says
// where).  With hipcc's FMA contraction, results differ from the
// reference in the last few bits only.
#pragma once

#include <hip/hip_runtime.h>

namespace sfamd {

enum : unsigned {
    kFlagState = 1u, kFlagChar = 2u, kFlagTeno = 4u, kFlagFlux = 8u, kFlagDt = 16u
};

// Tangent basis of Riemann::buildCharacteristicMatrix for n = e_AXIS
// (SF_riemann.cpp:88-97):
//   x: l = (-0, 1, 0), m = (0, 0, 1)
//   y: l = (-1, 0, 0), m = (0, -0, 1)
//   z: l = (-1, 0, 0), m = (0, -1, 0)
template <int AXIS> struct Basis;
template <> struct Basis<0> {
    static constexpr int nd = 0, li = 1, mi = 2;
    static constexpr double sl = 1.0, sm = 1.0;
};
template <> struct Basis<1> {
    static constexpr int nd = 1, li = 0, mi = 2;
    static constexpr double sl = -1.0, sm = 1.0;
};
template <> struct Basis<2> {
    static constexpr int nd = 2, li = 0, mi = 1;
    static constexpr double sl = -1.0, sm = -1.0;
};

__device__ __forceinline__ double sq(double x) { return x * x; }

// Reciprocal, quotient and reciprocal square root from the hardware
// estimate (good to about 5e-8 relative) refined by NR Newton steps: with
// two, within about one ulp of the correctly rounded value for the normal,
// finite, nonzero operands of this path; with one, within about 10 ulps
// (reciprocal) and 20 ulps (reciprocal square root).  A zero or non-finite
// operand yields NaN, which the flux and state checks report.  The IEEE
// sequences hipcc emits for '/' and sqrt are two to three times longer.
template <int NR = 2>
__device__ __forceinline__ double frcp(double b) {
    double r = __builtin_amdgcn_rcp(b);
#pragma unroll
    for (int i = 0; i < NR; ++i) r = fma(fma(-b, r, 1.0), r, r);
    return r;
}

template <int NR = 2>
__device__ __forceinline__ double fdiv(double a, double b) {
    const double r = frcp<NR>(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

template <int NR = 2>
__device__ __forceinline__ double frsqrt(double z) {
    double y = __builtin_amdgcn_rsq(z);
    const double h = 0.5 * z;
#pragma unroll
    for (int i = 0; i < NR; ++i) y = y * fma(-h * y, y, 1.5);
    return y;
}

// sqrt(z) as z * rsqrt(z) (z > 0)
template <int NR = 2>
__device__ __forceinline__ double fsqrt(double z) { return z * frsqrt<NR>(z); }

// The reference's cutoff test RN(g / gs) < CT (CT = 1e-5) without the
// division.  RN(x) < CT exactly when x < CT - d with d half the gap from CT
// down to the next double (2^-70 for CT = 1e-5), i.e. when
// g < CT*gs - d*gs.  CT*gs = P + E exactly (FMA), d*gs is exact, and near
// the boundary g - P is exact (Sterbenz), so the comparison below decides
// exactly as the division does (up to a tie at the midpoint).
__device__ __forceinline__ bool teno_cut(double g, double gs) {
#pragma clang fp contract(off)
    const double P = 1.0e-5 * gs;
    const double E = __builtin_fma(1.0e-5, gs, -P);
    return g - P < E - 0x1p-70 * gs;
}

// The left reconstruction at a face reads c0..c4 and the mirrored right one
// c5..c1 (fillScalarWENOStencils<6>, SF_WENO.h:79-98).  Two of the three
// candidate stencils are the same three nodes on both sides, so their
// candidate values (the same parabola at the same face) and the
// second-difference part of their smoothness indicators are formed once;
// the two final divisions share one reciprocal.  Each side's weights are
// those of WENO5::weno5_core (SF_WENO5.cpp:42-62) or TENO5::teno5_core
// (SF_TENO5.cpp:44-92).
struct PairTerms {
    double qL0, qR0, Sa, Sb;     // candidates (x6): left S0, right S0, shared
    double bL[3], bR[3];         // smoothness indicators
};

// The indicators come out as beta + eps: eps enters as the addend of the
// innermost fused multiply-add of each, so the weights' eps + beta costs no
// separate addition (eps = 0 gives beta itself).
__device__ __forceinline__ PairTerms pair_terms(double c0, double c1, double c2, double c3,
                                                double c4, double c5, double eps) {
    PairTerms t;
    t.qL0 = 2.0 * c0 - 7.0 * c1 + 11.0 * c2;
    t.qR0 = 2.0 * c5 - 7.0 * c4 + 11.0 * c3;
    t.Sa = -c1 + 5.0 * c2 + 2.0 * c3;   // left S1 = right S2
    t.Sb = 2.0 * c2 + 5.0 * c3 - c4;    // left S2 = right S1
    constexpr double k = 13.0 / 12.0;
    const double dB = sq(c1 - 2.0 * c2 + c3);
    const double dC = sq(c2 - 2.0 * c3 + c4);
    t.bL[0] = fma(k, sq(c0 - 2.0 * c1 + c2), fma(0.25, sq(c0 - 4.0 * c1 + 3.0 * c2), eps));
    t.bL[1] = fma(0.25, sq(c1 - c3), fma(k, dB, eps));
    t.bL[2] = fma(0.25, sq(3.0 * c2 - 4.0 * c3 + c4), fma(k, dC, eps));
    t.bR[0] = fma(k, sq(c5 - 2.0 * c4 + c3), fma(0.25, sq(c5 - 4.0 * c4 + 3.0 * c3), eps));
    t.bR[1] = fma(0.25, sq(c4 - c2), fma(k, dC, eps));
    t.bR[2] = fma(0.25, sq(3.0 * c3 - 4.0 * c2 + c1), fma(k, dB, eps));
    return t;
}

// a/b and c/d with one reciprocal (both denominators are positive and far
// from the ends of the exponent range here).
template <int NR = 2>
__device__ __forceinline__ void div_pair(double a, double b, double c, double d, double& x,
                                         double& y) {
    const double r = frcp<NR>(b * d);
    x = a * d * r;
    y = c * b * r;
}

template <int NR>
__device__ __forceinline__ void weno5_pair(double c0, double c1, double c2, double c3,
                                           double c4, double c5, double& left,
                                           double& right) {
    const PairTerms t = pair_terms(c0, c1, c2, c3, c4, c5, 1e-6);   // beta + eps
    const double eL0 = sq(t.bL[0]), eL1 = sq(t.bL[1]), eL2 = sq(t.bL[2]);
    const double eR0 = sq(t.bR[0]), eR1 = sq(t.bR[1]), eR2 = sq(t.bR[2]);
    // alpha_k e0 e1 e2 / 10 with d = (0.1, 0.6, 0.3): (1, 6, 3) x the
    // products of the other two e; the candidates carry a factor 6, so the
    // denominator is 6 x (A + 6B + 3C)
    const double AL = eL1 * eL2, BL = eL0 * eL2, CL = eL0 * eL1;
    const double AR = eR1 * eR2, BR = eR0 * eR2, CR = eR0 * eR1;
    const double nL = AL * t.qL0 + (6.0 * BL) * t.Sa + (3.0 * CL) * t.Sb;
    const double nR = AR * t.qR0 + (6.0 * BR) * t.Sb + (3.0 * CR) * t.Sa;
    div_pair<NR>(nL, 6.0 * AL + 36.0 * BL + 18.0 * CL, nR, 6.0 * AR + 36.0 * BR + 18.0 * CR, left,
             right);
}

// TENO5's stencil selection (SF_TENO5.cpp:60-89): gamma_k = (1 + tau5 /
// (beta_k + 1e-40))^6 and the cutoff of gamma_k / sum gamma at 1e-5.  The
// gammas only decide which stencils are kept, so the last bits of the
// quotients change a result only where a ratio lies within rounding of the
// cutoff.  Returns false where the sum of the gammas is not finite, which
// the reference treats as fatal (requireFiniteGammaSum, SF_TENO5.cpp:33-40).
__device__ __forceinline__ bool teno5_weights(const double b[3], double& c0, double& c1,
                                              double& c2) {
    const double tau5 = fabs(b[0] - b[2]);
    // tau5 / e_k for the three k from one reciprocal of e0 e1 e2 (each
    // e_k >= 1e-40, so the product stays far above the subnormal range)
    const double e0 = b[0] + 1.0e-40, e1 = b[1] + 1.0e-40, e2 = b[2] + 1.0e-40;
    const double e01 = e0 * e1;
    const double t = tau5 * frcp(e01 * e2);
    auto p6 = [](double x) {   // three products: within a few ulps of x^6
        const double x3 = x * x * x;
        return x3 * x3;
    };
    const double g0 = p6(1.0 + t * (e1 * e2));
    const double g1 = p6(1.0 + t * (e0 * e2));
    const double g2 = p6(1.0 + t * e01);
    const double gs = g0 + g1 + g2;
    c0 = teno_cut(g0, gs) ? 0.0 : 1.0;
    c1 = teno_cut(g1, gs) ? 0.0 : 1.0;
    c2 = teno_cut(g2, gs) ? 0.0 : 1.0;
    return gs < INFINITY;
}

__device__ __forceinline__ void teno5_pair(double c0, double c1, double c2, double c3,
                                           double c4, double c5, double& left, double& right,
                                           unsigned& flags) {
    const PairTerms t = pair_terms(c0, c1, c2, c3, c4, c5, 0.0);
    double l0, l1, l2, r0, r1, r2;
    const bool fin = teno5_weights(t.bL, l0, l1, l2) & teno5_weights(t.bR, r0, r1, r2);
    const double aL = 0.1 * l0 + 0.6 * l1 + 0.3 * l2;
    const double aR = 0.1 * r0 + 0.6 * r1 + 0.3 * r2;
    if (!fin || !(aL > 0.0) || !(aR > 0.0)) flags |= kFlagTeno;
    const double nL = 0.1 * l0 * t.qL0 + 0.6 * l1 * t.Sa + 0.3 * l2 * t.Sb;
    const double nR = 0.1 * r0 * t.qR0 + 0.6 * r1 * t.Sb + 0.3 * r2 * t.Sa;
    div_pair(nL, 6.0 * aL, nR, 6.0 * aR, left, right);
}

// WENO3::weno3_core (SF_WENO3.cpp:42-57) on both sides of a face: the left
// side reads c0..c3 and the right one c3..c0 (fillScalarWENOStencils<4>,
// SF_WENO.h:79-98).  The right side's two candidates and smoothness
// indicators are the left side's, exchanged, so each is formed once; the
// weights are those of the reference, normalised with one division per
// side (alpha_k = d_k / e_k multiplied through by e_A e_B) and the two
// divisions share one reciprocal.
template <int NR>
__device__ __forceinline__ void weno3_pair(const double c[4], double& left, double& right) {
    const double A = -c[0] + 5.0 * c[1] + 2.0 * c[2];   // 6 x left S0 = right S1
    const double B = 2.0 * c[1] + 5.0 * c[2] - c[3];    // 6 x left S1 = right S0
    const double eps = 1e-6;
    const double eA = sq(fma(c[2] - c[0], c[2] - c[0], eps));
    const double eB = sq(fma(c[3] - c[1], c[3] - c[1], eps));
    const double d0 = 1.0 / 3.0, d1 = 2.0 / 3.0;
    const double nL = d0 * eB * A + d1 * eA * B, nR = d0 * eA * B + d1 * eB * A;
    const double sL = d0 * eB + d1 * eA, sR = d0 * eA + d1 * eB;
    div_pair<NR>(nL, 6.0 * sL, nR, 6.0 * sR, left, right);
}

// Smoothness of one four-node WENO7 candidate over the cell [a, b] of its
// local coordinate, as beta4 (SF_WENO7.cpp:43-68) forms it: the integral
// of the squared first, second and third derivatives of the cubic with
// forward differences d1, 2 d2, 6 d3.  b - a = 1 for every candidate, and
// the second-derivative coefficients D = 2 d2 - 6 d3, E = 6 d3 of the
// reference equal B and 2 C.
struct W7Interval {
    double i2, i3, i4, i5;   // F(b) - F(a) for x^2, x^3/3, x^4/2, x^5/5
};

constexpr W7Interval w7_interval(double a, double b) {
    return {b * b - a * a, b * b * b / 3.0 - a * a * a / 3.0,
            (b * b) * (b * b) / 2.0 - (a * a) * (a * a) / 2.0,
            b * b * b * b * b / 5.0 - a * a * a * a * a / 5.0};
}

// Returns beta + eps, eps added inside the last fused multiply-add.
__device__ __forceinline__ double w7_beta(double v0, double v1, double v2, double v3,
                                          const W7Interval& I, double eps) {
    const double d1 = v1 - v0;
    const double d2 = (v2 - 2.0 * v1 + v0) * 0.5;
    const double d3 = (v3 - 3.0 * v2 + 3.0 * v1 - v0) * (1.0 / 6.0);
    const double A = d1 - d2 + 2.0 * d3;
    const double B = 2.0 * d2 - 6.0 * d3;
    const double C = 3.0 * d3;
    const double E = 2.0 * C;
    const double beta1 = A * A + A * B * I.i2 + (B * B + 2.0 * A * C) * I.i3 + B * C * I.i4 +
                         C * C * I.i5;
    const double beta2 = B * B + B * E * I.i2 + E * E * I.i3;
    const double beta3 = fma(36.0 * d3, d3, eps);
    return beta1 + beta2 + beta3;
}

// WENO7::weno7_core (SF_WENO7.cpp:70-98) on both sides of a face: left
// reads c0..c6, right c7..c1 (fillScalarWENOStencils<8>).  Three of the
// four candidates of each side are candidates of the other side too (the
// same four nodes in reverse, the same polynomial at the same face), so
// five candidate values are formed; the smoothness intervals differ
// between the sides (each side's target cell is its own upwind node), so
// all eight indicators are formed.  Each side's weights are normalised
// with one division (alpha_k = d_k / e_k multiplied through by the
// product of the four e_j) and the two divisions share one reciprocal.
__device__ __forceinline__ void weno7_side_weights(const double b[4], double w[4]) {
    constexpr double d0 = 1.0 / 35.0, d1 = 12.0 / 35.0, d2 = 18.0 / 35.0, d3 = 4.0 / 35.0;
    // b[k] = beta_k + eps (w7_beta)
    const double e0 = sq(b[0]), e1 = sq(b[1]);
    const double e2 = sq(b[2]), e3 = sq(b[3]);
    const double e01 = e0 * e1, e23 = e2 * e3;
    w[0] = d0 * (e1 * e23);
    w[1] = d1 * (e0 * e23);
    w[2] = d2 * (e01 * e3);
    w[3] = d3 * (e01 * e2);
}

__device__ __forceinline__ void weno7_pair(const double c[8], double& left, double& right) {
    constexpr W7Interval I0 = w7_interval(2.5, 3.5), I1 = w7_interval(1.5, 2.5);
    constexpr W7Interval I2 = w7_interval(0.5, 1.5), I3 = w7_interval(-0.5, 0.5);
    // 12 x candidates: left S0, the shared ones (left S1 = right S3, left
    // S2 = right S2, left S3 = right S1), right S0
    const double L0 = -3.0 * c[0] + 13.0 * c[1] - 23.0 * c[2] + 25.0 * c[3];
    const double P1 = c[1] - 5.0 * c[2] + 13.0 * c[3] + 3.0 * c[4];
    const double P2 = -c[2] + 7.0 * c[3] + 7.0 * c[4] - c[5];
    const double P3 = 3.0 * c[3] + 13.0 * c[4] - 5.0 * c[5] + c[6];
    const double R0 = -3.0 * c[7] + 13.0 * c[6] - 23.0 * c[5] + 25.0 * c[4];
    constexpr double eps = 1e-6;
    double bL[4], bR[4];
    bL[0] = w7_beta(c[0], c[1], c[2], c[3], I0, eps);
    bL[1] = w7_beta(c[1], c[2], c[3], c[4], I1, eps);
    bL[2] = w7_beta(c[2], c[3], c[4], c[5], I2, eps);
    bL[3] = w7_beta(c[3], c[4], c[5], c[6], I3, eps);
    bR[0] = w7_beta(c[7], c[6], c[5], c[4], I0, eps);
    bR[1] = w7_beta(c[6], c[5], c[4], c[3], I1, eps);
    bR[2] = w7_beta(c[5], c[4], c[3], c[2], I2, eps);
    bR[3] = w7_beta(c[4], c[3], c[2], c[1], I3, eps);
    double wL[4], wR[4];
    weno7_side_weights(bL, wL);
    weno7_side_weights(bR, wR);
    const double nL = wL[0] * L0 + wL[1] * P1 + wL[2] * P2 + wL[3] * P3;
    const double nR = wR[0] * R0 + wR[1] * P3 + wR[2] * P2 + wR[3] * P1;
    div_pair(nL, 12.0 * (wL[0] + wL[1] + wL[2] + wL[3]), nR,
             12.0 * (wR[0] + wR[1] + wR[2] + wR[3]), left, right);
}

// Scheme ids (SF_SCHEME_*) and their stencils (SF_stencil.h:12-17):
//   0 WENO5 and 1 TENO5 read 6 nodes, 2 WENO3 reads 4, 3 WENO7 reads 8.
// newton: the Newton steps of every reciprocal and reciprocal square root
// in a face of the scheme.  One for WENO5 and WENO3, whose weights are
// smooth functions of their inputs, so a last-bit difference stays a
// last-bit difference; two for TENO5, whose stencil cutoff turns one into
// a different stencil, and for WENO7, which is held to the reference's
// own frozen Sod output.
template <int SCHEME> struct Stencil { static constexpr int ns = 6, newton = 1; };
template <> struct Stencil<1> { static constexpr int ns = 6, newton = 2; };
template <> struct Stencil<2> { static constexpr int ns = 4, newton = 1; };
template <> struct Stencil<3> { static constexpr int ns = 8, newton = 2; };

template <int SCHEME, int NR>
__device__ __forceinline__ void recon_pair(const double (&ch)[Stencil<SCHEME>::ns], double& left,
                                           double& right, unsigned& flags) {
    if constexpr (SCHEME == 1) teno5_pair(ch[0], ch[1], ch[2], ch[3], ch[4], ch[5], left, right, flags);
    else if constexpr (SCHEME == 2) weno3_pair<NR>(ch, left, right);
    else if constexpr (SCHEME == 3) weno7_pair(ch, left, right);
    else weno5_pair<NR>(ch[0], ch[1], ch[2], ch[3], ch[4], ch[5], left, right);
}

// Closure of SingleFluidStateModel::close + PerfectGasEOS
// (SF_singleFluidStateModel.cpp:49-71, SF_perfectGasEOS.h:24-32).
// One reciprocal of the density replaces the five divisions by rho.
struct Closed {
    double u[3], p, c, inv_rho;
};

// close_state with the reciprocal of the density given.
template <int NR = 2>
__device__ __forceinline__ Closed close_state_r(const double q[5], double gamma, double inv_rho) {
    Closed s;
    const double rho = q[0];
    s.inv_rho = inv_rho;
    s.u[0] = q[1] * s.inv_rho;
    s.u[1] = q[2] * s.inv_rho;
    s.u[2] = q[3] * s.inv_rho;
    const double kinetic = 0.5 * rho * (s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]);
    const double e = (q[4] - kinetic) * s.inv_rho;
    s.p = (gamma - 1.0) * rho * e;
    s.c = fsqrt<NR>(gamma * s.p * s.inv_rho);
    return s;
}

template <int NR = 2>
__device__ __forceinline__ Closed close_state(const double q[5], double gamma) {
    return close_state_r<NR>(q, gamma, frcp<NR>(q[0]));
}


// The closure's validity test, as PerfectGasEOS::fromDensityEnergy makes
// it on each face state the Rusanov flux closes (SF_perfectGasEOS.h:26-33
// with requireDensity and checked: rho > 0, and p, T, c > 0 and finite;
// for a perfect gas with rho > 0, T and c are positive and finite exactly
// when p is): the reference stops on a face whose reconstructed state
// fails it.
__device__ __forceinline__ bool closed_ok(const double q[5], const Closed& s) {
    return q[0] > 0.0 && q[0] < INFINITY && s.p > 0.0 && s.p < INFINITY;
}

// physicalFlux (SF_rusanovEOS.cpp:14-32) for n = e_ND.
template <int ND>
__device__ __forceinline__ void physical_flux(const double q[5], const Closed& s,
                                              double f[5]) {
    const double un = s.u[ND];
    f[0] = q[0] * un;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[1 + d] = (d == ND) ? q[1 + d] * un + s.p : q[1 + d] * un;
    f[4] = (q[4] + s.p) * un;
}

// RusanovEOS::faceFlux, SF_rusanovEOS.cpp:47-77.
template <int ND, int NR = 2>
__device__ __forceinline__ void rusanov(const double qL[5], const double qR[5],
                                        double gamma, double out[5], unsigned& flags) {
    const Closed l = close_state<NR>(qL, gamma), r = close_state<NR>(qR, gamma);
    if (!closed_ok(qL, l) || !closed_ok(qR, r)) flags |= kFlagFlux;
    double fL[5], fR[5];
    physical_flux<ND>(qL, l, fL);
    physical_flux<ND>(qR, r, fR);
    const double lambda = fmax(fabs(l.u[ND]) + l.c, fabs(r.u[ND]) + r.c);
#pragma unroll
    for (int v = 0; v < 5; ++v)
        out[v] = 0.5 * (fL[v] + fR[v]) - 0.5 * lambda * (qR[v] - qL[v]);
}

// HLLC (no reference counterpart).  Toro, "Riemann Solvers and Numerical
// Methods for Fluid Dynamics", 3rd ed.: Einfeldt wave speeds from Roe
// averages, contact speed eq. (10.37), star states eq. (10.39) with the
// tangential velocity of each side carried unchanged, flux eq. (10.38).
template <int ND, int NR = 2>
__device__ __forceinline__ void hllc(const double qL[5], const double qR[5],
                                     double gamma, double out[5], unsigned& flags) {
    // 1/rho = (rsqrt rho)^2 and sqrt rho = rho rsqrt rho from one refined
    // reciprocal square root per side
    const double yL = frsqrt<NR>(qL[0]), yR = frsqrt<NR>(qR[0]);
    const Closed l = close_state_r<NR>(qL, gamma, yL * yL), r = close_state_r<NR>(qR, gamma, yR * yR);
    if (!closed_ok(qL, l) || !closed_ok(qR, r)) flags |= kFlagFlux;
    const double unL = l.u[ND], unR = r.u[ND];
    // Roe weights sqrt(rho_K) / (sqrt(rho_L) + sqrt(rho_R))
    const double inv = frcp<NR>(qL[0] * yL + qR[0] * yR);
    const double wL = qL[0] * yL * inv, wR = qR[0] * yR * inv;
    double ut[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ut[d] = wL * l.u[d] + wR * r.u[d];
    const double HL = (qL[4] + l.p) * l.inv_rho, HR = (qR[4] + r.p) * r.inv_rho;
    const double Ht = wL * HL + wR * HR;
    const double ct = fsqrt<NR>((gamma - 1.0) * (Ht - 0.5 * (ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2])));
    const double SL = fmin(unL - l.c, ut[ND] - ct);
    const double SR = fmax(unR + r.c, ut[ND] + ct);
    const double mL = qL[0] * (SL - unL), mR = qR[0] * (SR - unR);
    const double Ss = fdiv<NR>(r.p - l.p + unL * mL - unR * mR, mL - mR);
    // One side's physical flux serves all three regions: the left state's
    // where the fan leaves the face to the right (0 <= S_L) or the contact
    // does (0 <= S*), the right state's otherwise; the star correction
    // applies between the outer waves only.
    const bool supL = 0.0 <= SL, supR = SR <= 0.0;
    const bool left = supL || (!supR && 0.0 <= Ss);
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = left ? qL[v] : qR[v];
    const Closed& s = left ? l : r;
    const double un = left ? unL : unR;
    const double SK = left ? SL : SR;
    const double mK = left ? mL : mR;
    double f[5];
    physical_flux<ND>(q, s, f);
    // scale = mK / (SK - Ss); the pressure term p / mK * scale = p / (SK - Ss)
    const double rs = frcp<NR>(SK - Ss);
    const double scale = mK * rs;
    double us[5];
    us[0] = scale;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        us[1 + d] = scale * ((d == ND) ? s.u[d] + (Ss - un) : s.u[d]);
    us[4] = scale * (q[4] * s.inv_rho + (Ss - un) * Ss) + (Ss - un) * (s.p * rs);
    const bool star = !supL && !supR;
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = star ? f[v] + SK * (us[v] - q[v]) : f[v];
}

// Pressure as Numerics::pressure forms it (SF_utility.h:24-29), the form
// the flux splittings, the Roe flux and the characteristic matrix use.
struct Legacy {
    double inv_rho, u[3], p;
    bool ok;
};

template <int NR = 2>
__device__ __forceinline__ Legacy legacy_state(const double q[5], double gamma) {
    Legacy s;
    s.inv_rho = frcp<NR>(q[0]);
    const double ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * s.inv_rho;
    s.p = (gamma - 1.0) * (q[4] - ke);
    s.u[0] = q[1] * s.inv_rho;
    s.u[1] = q[2] * s.inv_rho;
    s.u[2] = q[3] * s.inv_rho;
    s.ok = q[0] > 0.0 && s.p > 0.0;
    return s;
}

// One half of the Steger-Warming splitting of a state (POS: F+, else F-)
// for n = e_ND, with the reference's smoothed split of eigenvalues below
// 2 eps_sw = 0.02 (Riemann::splitStegerWarming, SF_riemann.cpp:15-67).
// The three distinct eigenvalues u_n, u_n + a, u_n - a are split once each.
template <int ND, bool POS, int NR = 2>
__device__ __forceinline__ void steger_half(const double q[5], double gamma, double f[5],
                                            unsigned& flags) {
    const Legacy s = legacy_state<NR>(q, gamma);
    if (!s.ok) flags |= kFlagFlux;
    const double a = fsqrt<NR>(gamma * s.p * s.inv_rho);
    const double un = s.u[ND];
    const double q2 = s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2];
    auto split = [](double lam) {
        const double eps_sw = 0.01;
        double lp;
        if (fabs(lam) < 2.0 * eps_sw) lp = 0.5 * (lam * lam * (1.0 / (4.0 * eps_sw)) + lam + eps_sw);
        else lp = 0.5 * (lam + fabs(lam));
        return POS ? lp : lam - lp;
    };
    const double l0 = split(un), l3 = split(un + a), l4 = split(un - a);
    const double W1 = 2.0 * (gamma - 1.0) * l0 + l3 + l4;
    const double W2 = l3 - l4;
    const double W3 = l3 + l4;
    const double common = q[0] * frcp<NR>(2.0 * gamma);
    f[0] = common * W1;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[1 + d] = common * ((d == ND) ? s.u[d] * W1 + a * W2 : s.u[d] * W1);
    f[4] = common * (0.5 * q2 * W1 + a * un * W2 + (a * a) * frcp<NR>(gamma - 1.0) * W3);
}

// StegerWarmingFlux::flux (SF_sw.cpp:21-27): F+(qL) + F-(qR).
template <int ND, int NR = 2>
__device__ __forceinline__ void steger_warming(const double qL[5], const double qR[5],
                                               double gamma, double out[5], unsigned& flags) {
    double fp[5], fm[5];
    steger_half<ND, true, NR>(qL, gamma, fp, flags);
    steger_half<ND, false, NR>(qR, gamma, fm, flags);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = fp[v] + fm[v];
}

// One half of the local Lax-Friedrichs splitting, (F(q) +- lambda q) / 2
// with lambda = |u_n| + a of the state (Riemann::splitLaxFriedrichs,
// SF_riemann.cpp:125-154).
template <int ND, bool POS, int NR = 2>
__device__ __forceinline__ void laxf_half(const double q[5], double gamma, double f[5],
                                          unsigned& flags) {
    const Legacy s = legacy_state<NR>(q, gamma);
    if (!s.ok) flags |= kFlagFlux;
    const double a = fsqrt<NR>(gamma * s.p * s.inv_rho);
    const double un = s.u[ND];
    const double lam = POS ? fabs(un) + a : -(fabs(un) + a);
    double F[5];
    F[0] = q[0] * un;
#pragma unroll
    for (int d = 0; d < 3; ++d) F[1 + d] = (d == ND) ? q[1 + d] * un + s.p : q[1 + d] * un;
    F[4] = (q[4] + s.p) * un;
#pragma unroll
    for (int v = 0; v < 5; ++v) f[v] = 0.5 * (F[v] + lam * q[v]);
}

// LaxFriedrichsFlux::flux (SF_lxF.cpp:21-27).
template <int ND, int NR = 2>
__device__ __forceinline__ void lax_friedrichs(const double qL[5], const double qR[5],
                                               double gamma, double out[5], unsigned& flags) {
    double fp[5], fm[5];
    laxf_half<ND, true, NR>(qL, gamma, fp, flags);
    laxf_half<ND, false, NR>(qR, gamma, fm, flags);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = fp[v] + fm[v];
}

// physicalEulerFlux (SF_face.h:112-128) for n = e_ND.
template <int ND>
__device__ __forceinline__ void euler_flux(const double q[5], const Legacy& s, double f[5]) {
    const double un = s.u[ND];
    f[0] = q[0] * un;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[1 + d] = (d == ND) ? q[0] * s.u[d] * un + s.p : q[0] * s.u[d] * un;
    f[4] = (q[4] + s.p) * un;
}

// The eigenvector matrices of Riemann::buildCharacteristicMatrix
// (SF_riemann.cpp:69-123) at a state q for n = e_AXIS, in the factored
// form used by the projections below.
template <int AXIS, int NR = 2>
struct CharBasis {
    double u[3], hq2, c, H, un, ul, um, b2, x1;

    __device__ __forceinline__ CharBasis(const double q[5], double gamma, unsigned& flags) {
        using B = Basis<AXIS>;
        const double invR = frcp<NR>(q[0]);
        const double ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * invR;
        const double p = (gamma - 1.0) * (q[4] - ke);
        if (!(q[0] > 0.0) || !(p > 0.0)) flags |= kFlagChar;
        u[0] = q[1] * invR;
        u[1] = q[2] * invR;
        u[2] = q[3] * invR;
        hq2 = 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double cz = gamma * p * invR;
        const double rc = frsqrt<NR>(cz);
        c = cz * rc;
        H = (q[4] + p) * invR;
        un = u[B::nd];
        ul = B::sl * u[B::li];
        um = B::sm * u[B::mi];
        b2 = 0.5 * ((gamma - 1.0) * (rc * rc));
        x1 = 0.5 * rc;
    }

    // w = L q (gemv5, SF_riemann.h:33-37).  With P = b2 (q2/2 rho - u.m + E)
    // and X = x1 (un rho - m_n), the rows of L give w0 = rho - 2P (the
    // acoustic-free row), w3 = P + X, w4 = P - X, and the shear rows
    // w1 = sl (m_l - u_l rho), w2 = sm (m_m - u_m rho).
    __device__ __forceinline__ void project(const double q[5], double w[5]) const {
        using B = Basis<AXIS>;
        const double r = q[0];
        const double um_dot = u[0] * q[1] + u[1] * q[2] + u[2] * q[3];
        const double P = b2 * (hq2 * r - um_dot + q[4]);
        const double X = x1 * (un * r - q[1 + B::nd]);
        w[0] = r - 2.0 * P;
        w[1] = -ul * r + B::sl * q[1 + B::li];
        w[2] = -um * r + B::sm * q[1 + B::mi];
        w[3] = P + X;
        w[4] = P - X;
    }

    // q = R w.  Every column of R carries u in the momentum rows except the
    // acoustic -+c on the normal and the shear units, so q_m = u q_rho +
    // (the rest).
    __device__ __forceinline__ void back(const double w[5], double q[5]) const {
        using B = Basis<AXIS>;
        const double r = w[0] + w[3] + w[4];
        const double dw = w[4] - w[3];
        q[0] = r;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d == B::nd) q[1 + d] = u[d] * r + c * dw;
            else if (d == B::li) q[1 + d] = u[d] * r + B::sl * w[1];
            else q[1 + d] = u[d] * r + B::sm * w[2];
        }
        q[4] = hq2 * w[0] + ul * w[1] + um * w[2] + H * (w[3] + w[4]) + (c * un) * dw;
    }
};

// RoeFlux::flux (SF_roe.cpp:20-111) for n = e_AXIS: Roe averages,
// Harten's fix with delta = a~/10 on every wave, and the dissipation
// R |Lambda| L (qR - qL) with R, L of the characteristic basis at the
// state (rho~, rho~ u~, rho~ h~) -- the reference's energy slot holds
// rho~ h~, so that basis has sound speed sqrt(gamma) a~ -- and the wave
// speeds taken in the reference's order (u_n, u_n, u_n, u_n + a~,
// u_n - a~) against the basis columns (the fourth is its u_n - c wave).
template <int AXIS, int NR = 2>
__device__ __forceinline__ void roe(const double qL[5], const double qR[5], double gamma,
                                    double out[5], unsigned& flags) {
    constexpr int ND = Basis<AXIS>::nd;
    const double gm1 = gamma - 1.0;
    const Legacy l = legacy_state<NR>(qL, gamma), r = legacy_state<NR>(qR, gamma);
    if (!l.ok || !r.ok) flags |= kFlagFlux;
    const double hL = (qL[4] + l.p) * l.inv_rho, hR = (qR[4] + r.p) * r.inv_rho;
    const double sL = fsqrt<NR>(qL[0]), sR = fsqrt<NR>(qR[0]);
    const double inv = frcp<NR>(sL + sR);
    double ut[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ut[d] = (sL * l.u[d] + sR * r.u[d]) * inv;
    const double ht = (sL * hL + sR * hR) * inv;
    const double q2t = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const double a2 = gm1 * (ht - 0.5 * q2t);
    if (!(a2 > 0.0)) flags |= kFlagFlux;
    const double at = fsqrt<NR>(a2);
    const double unt = ut[ND];
    const double delta = 0.1 * at;
    auto fix = [delta](double lam) {
        const double v = fabs(lam);
        return v < 2.0 * delta ? 0.5 * (v * v * frcp<NR>(2.0 * delta) + delta) : v;
    };
    const double lam0 = fix(unt), lam3 = fix(unt + at), lam4 = fix(unt - at);
    const double rt = sL * sR;
    double qt[5];
    qt[0] = rt;
#pragma unroll
    for (int d = 0; d < 3; ++d) qt[1 + d] = rt * ut[d];
    qt[4] = rt * (0.5 * q2t + a2 * frcp<NR>(gm1));
    const CharBasis<AXIS, NR> cb(qt, gamma, flags);
    double dq[5], al[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) dq[v] = qR[v] - qL[v];
    cb.project(dq, al);
    al[0] *= lam0;
    al[1] *= lam0;
    al[2] *= lam0;
    al[3] *= lam3;
    al[4] *= lam4;
    double dis[5];
    cb.back(al, dis);
    double fL[5], fR[5];
    euler_flux<ND>(qL, l, fL);
    euler_flux<ND>(qR, r, fR);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fL[v] + fR[v] - dis[v]);
}

// LaxWendroffFlux::flux (SF_lw.cpp:17-40): the physical flux of the
// Richtmyer half state q* = (qL + qR)/2 - (lambda/2)(F(qR) - F(qL)),
// lambda = dt / spacing with the step's dt.
template <int ND, int NR = 2>
__device__ __forceinline__ void lax_wendroff(const double qL[5], const double qR[5],
                                             double gamma, double lambda, double out[5],
                                             unsigned& flags) {
    const Legacy l = legacy_state<NR>(qL, gamma), r = legacy_state<NR>(qR, gamma);
    if (!l.ok || !r.ok || !(lambda > 0.0)) flags |= kFlagFlux;
    double fL[5], fR[5];
    euler_flux<ND>(qL, l, fL);
    euler_flux<ND>(qR, r, fR);
    double qh[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) qh[v] = 0.5 * (qL[v] + qR[v]) - 0.5 * lambda * (fR[v] - fL[v]);
    const Legacy h = legacy_state<NR>(qh, gamma);
    if (!h.ok) flags |= kFlagFlux;
    euler_flux<ND>(qh, h, out);
}

// Flux ids (SF_FLUX_*): 0 Rusanov, 1 HLLC, 2 Steger-Warming,
// 3 Lax-Friedrichs, 4 Roe, 5 Lax-Wendroff (which reads lw = dt / spacing).
template <int AXIS, int FLUX, int NR = 2>
__device__ __forceinline__ void riemann_flux(const double qL[5], const double qR[5],
                                             double gamma, double lw, double out[5],
                                             unsigned& flags) {
    constexpr int nd = Basis<AXIS>::nd;
    if constexpr (FLUX == 1) hllc<nd, NR>(qL, qR, gamma, out, flags);
    else if constexpr (FLUX == 2) steger_warming<nd, NR>(qL, qR, gamma, out, flags);
    else if constexpr (FLUX == 3) lax_friedrichs<nd, NR>(qL, qR, gamma, out, flags);
    else if constexpr (FLUX == 4) roe<AXIS, NR>(qL, qR, gamma, out, flags);
    else if constexpr (FLUX == 5) lax_wendroff<nd, NR>(qL, qR, gamma, lw, out, flags);
    else rusanov<nd, NR>(qL, qR, gamma, out, flags);
}

// Flux::characteristicWENOFaceFlux<NS> (SF_WENO.h:185-245) on a face of
// axis AXIS whose cofactor is area * e_AXIS.  st[s][v] holds the stencil
// nodes at offsets 1-NS/2 .. NS/2 around the face.  Returns the
// area-weighted flux.
template <int AXIS, int SCHEME, int FLUX>
__device__ __forceinline__ void face_flux(const double (&st)[Stencil<SCHEME>::ns][5],
                                          double gamma, double area, double out[5],
                                          unsigned& flags, double lw = 0.0) {
    constexpr int NS = Stencil<SCHEME>::ns;
    constexpr int NR = Stencil<SCHEME>::newton;
    // face average (SF_WENO.h:207-211) and its characteristic basis
    double qa[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) qa[v] = 0.5 * (st[NS / 2 - 1][v] + st[NS / 2][v]);
    const CharBasis<AXIS, NR> cb(qa, gamma, flags);
    double ch[5][NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double w[5];
        cb.project(st[s], w);
#pragma unroll
        for (int v = 0; v < 5; ++v) ch[v][s] = w[v];
    }
    double wL[5], wR[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) recon_pair<SCHEME, NR>(ch[v], wL[v], wR[v], flags);
    double qL[5], qR[5];
    cb.back(wL, qL);
    cb.back(wR, qR);
    double f[5];
    riemann_flux<AXIS, FLUX, NR>(qL, qR, gamma, lw, f, flags);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = f[v] * area;
}

}  // namespace sfamd
