// sf_face.h -- per-face FP64 arithmetic of the explicit Euler hot path for
// gfx950: characteristic WENO5-JS / TENO5 reconstruction on an axis-aligned
// face followed by a Rusanov or HLLC flux.
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
// estimate refined by two Newton steps: within about one ulp of the
// correctly rounded value for the normal, finite, nonzero operands of this
// path (a zero or non-finite operand yields NaN, which the flux and state
// checks report).  The IEEE sequences hipcc emits for '/' and sqrt are two
// to three times longer.
__device__ __forceinline__ double frcp(double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    return r;
}

__device__ __forceinline__ double fdiv(double a, double b) {
    const double r = frcp(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

__device__ __forceinline__ double frsqrt(double z) {
    double y = __builtin_amdgcn_rsq(z);
    const double h = 0.5 * z;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// sqrt(z) as z * rsqrt(z) (z > 0)
__device__ __forceinline__ double fsqrt(double z) { return z * frsqrt(z); }

// WENO5::weno5_core, src/methods/numerics/convection/SF_WENO5.cpp:42-62.
// Same weights and candidates, evaluated with one division: with
// e_k = (eps + beta_k)^2 and alpha_k = d_k / e_k,
//   sum_k alpha_k p_k / sum_k alpha_k
//     = sum_k d_k (6 p_k) E_k / (6 sum_k d_k E_k),   E_k = prod_{j != k} e_j,
// and the candidates are kept as 6 p_k.  e_k >= 1e-12, and for the
// characteristic variables (O(rho)) the products stay far from overflow.
// Rounding differs from the seven-division form by a few ulps.
__device__ __forceinline__ double weno5(double v0, double v1, double v2,
                                        double v3, double v4) {
    const double q0 = 2.0 * v0 - 7.0 * v1 + 11.0 * v2;
    const double q1 = -v1 + 5.0 * v2 + 2.0 * v3;
    const double q2 = 2.0 * v2 + 5.0 * v3 - v4;
    const double b0 = 13.0 / 12.0 * sq(v0 - 2.0 * v1 + v2)
                    + 0.25 * sq(v0 - 4.0 * v1 + 3.0 * v2);
    const double b1 = 13.0 / 12.0 * sq(v1 - 2.0 * v2 + v3)
                    + 0.25 * sq(v1 - v3);
    const double b2 = 13.0 / 12.0 * sq(v2 - 2.0 * v3 + v4)
                    + 0.25 * sq(3.0 * v2 - 4.0 * v3 + v4);
    const double eps = 1e-6;
    const double e0 = sq(eps + b0), e1 = sq(eps + b1), e2 = sq(eps + b2);
    const double w0 = 0.1 * (e1 * e2), w1 = 0.6 * (e0 * e2), w2 = 0.3 * (e0 * e1);
    return (w0 * q0 + w1 * q1 + w2 * q2) / (6.0 * (w0 + w1 + w2));
}

// x^6 to within a fraction of an ulp beyond the final rounding: the
// products are carried with their exact rounding errors (FMA two-product),
// so the result is the correctly rounded x^6 except within ~2^-100 of a
// rounding boundary -- the value std::pow(x, 6.0) returns in round-to-nearest
// except in its own rare last-ulp misses.
__device__ __forceinline__ double pow6(double x) {
#pragma clang fp contract(off)
    const double p2 = x * x;
    const double e2 = __builtin_fma(x, x, -p2);
    const double p3 = p2 * x;
    const double e3 = __builtin_fma(p2, x, -p3) + e2 * x;
    const double p6 = p3 * p3;
    const double e6 = __builtin_fma(p3, p3, -p6) + 2.0 * p3 * e3;
    return p6 + e6;
}

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

// TENO5::teno5_core, src/methods/numerics/convection/SF_TENO5.cpp:44-92
// (C = 1, q = 6, CT = 1e-5).  The 1/6 of the candidates is folded into the
// final division.
__device__ __forceinline__ double teno5(double v0, double v1, double v2,
                                        double v3, double v4, unsigned& flags) {
    const double q0 = 2.0 * v0 - 7.0 * v1 + 11.0 * v2;
    const double q1 = -v1 + 5.0 * v2 + 2.0 * v3;
    const double q2 = 2.0 * v2 + 5.0 * v3 - v4;
    const double b0 = 13.0 / 12.0 * sq(v0 - 2.0 * v1 + v2)
                    + 0.25 * sq(v0 - 4.0 * v1 + 3.0 * v2);
    const double b1 = 13.0 / 12.0 * sq(v1 - 2.0 * v2 + v3)
                    + 0.25 * sq(v1 - v3);
    const double b2 = 13.0 / 12.0 * sq(v2 - 2.0 * v3 + v4)
                    + 0.25 * sq(3.0 * v2 - 4.0 * v3 + v4);
    const double tau5 = fabs(b0 - b2);
    const double g0 = pow6(1.0 + tau5 / (b0 + 1.0e-40));
    const double g1 = pow6(1.0 + tau5 / (b1 + 1.0e-40));
    const double g2 = pow6(1.0 + tau5 / (b2 + 1.0e-40));
    const double gs = g0 + g1 + g2;
    const double c0 = teno_cut(g0, gs) ? 0.0 : 1.0;
    const double c1 = teno_cut(g1, gs) ? 0.0 : 1.0;
    const double c2 = teno_cut(g2, gs) ? 0.0 : 1.0;
    const double active = 0.1 * c0 + 0.6 * c1 + 0.3 * c2;
    if (!(active > 0.0)) flags |= kFlagTeno;
    return (0.1 * c0 * q0 + 0.6 * c1 * q1 + 0.3 * c2 * q2) / (6.0 * active);
}

// The left reconstruction at a face reads c0..c4 and the mirrored right one
// c5..c1 (fillScalarWENOStencils<6>, SF_WENO.h:79-98).  Two of the three
// candidate stencils are the same three nodes on both sides, so their
// candidate values (the same parabola at the same face) and the
// second-difference part of their smoothness indicators are formed once;
// the two final divisions share one reciprocal.  Each side's weights are
// those of weno5 / teno5 above.
struct PairTerms {
    double qL0, qR0, Sa, Sb;     // candidates (x6): left S0, right S0, shared
    double bL[3], bR[3];         // smoothness indicators
};

__device__ __forceinline__ PairTerms pair_terms(double c0, double c1, double c2, double c3,
                                                double c4, double c5) {
    PairTerms t;
    t.qL0 = 2.0 * c0 - 7.0 * c1 + 11.0 * c2;
    t.qR0 = 2.0 * c5 - 7.0 * c4 + 11.0 * c3;
    t.Sa = -c1 + 5.0 * c2 + 2.0 * c3;   // left S1 = right S2
    t.Sb = 2.0 * c2 + 5.0 * c3 - c4;    // left S2 = right S1
    const double dB = sq(c1 - 2.0 * c2 + c3);
    const double dC = sq(c2 - 2.0 * c3 + c4);
    t.bL[0] = 13.0 / 12.0 * sq(c0 - 2.0 * c1 + c2) + 0.25 * sq(c0 - 4.0 * c1 + 3.0 * c2);
    t.bL[1] = 13.0 / 12.0 * dB + 0.25 * sq(c1 - c3);
    t.bL[2] = 13.0 / 12.0 * dC + 0.25 * sq(3.0 * c2 - 4.0 * c3 + c4);
    t.bR[0] = 13.0 / 12.0 * sq(c5 - 2.0 * c4 + c3) + 0.25 * sq(c5 - 4.0 * c4 + 3.0 * c3);
    t.bR[1] = 13.0 / 12.0 * dC + 0.25 * sq(c4 - c2);
    t.bR[2] = 13.0 / 12.0 * dB + 0.25 * sq(3.0 * c3 - 4.0 * c2 + c1);
    return t;
}

// a/b and c/d with one reciprocal (both denominators are positive and far
// from the ends of the exponent range here).
__device__ __forceinline__ void div_pair(double a, double b, double c, double d, double& x,
                                         double& y) {
    const double r = frcp(b * d);
    x = a * d * r;
    y = c * b * r;
}

__device__ __forceinline__ void weno5_pair(double c0, double c1, double c2, double c3,
                                           double c4, double c5, double& left,
                                           double& right) {
    const PairTerms t = pair_terms(c0, c1, c2, c3, c4, c5);
    const double eps = 1e-6;
    const double eL0 = sq(eps + t.bL[0]), eL1 = sq(eps + t.bL[1]), eL2 = sq(eps + t.bL[2]);
    const double eR0 = sq(eps + t.bR[0]), eR1 = sq(eps + t.bR[1]), eR2 = sq(eps + t.bR[2]);
    const double wL0 = 0.1 * (eL1 * eL2), wL1 = 0.6 * (eL0 * eL2), wL2 = 0.3 * (eL0 * eL1);
    const double wR0 = 0.1 * (eR1 * eR2), wR1 = 0.6 * (eR0 * eR2), wR2 = 0.3 * (eR0 * eR1);
    const double nL = wL0 * t.qL0 + wL1 * t.Sa + wL2 * t.Sb;
    const double nR = wR0 * t.qR0 + wR1 * t.Sb + wR2 * t.Sa;
    div_pair(nL, 6.0 * (wL0 + wL1 + wL2), nR, 6.0 * (wR0 + wR1 + wR2), left, right);
}

__device__ __forceinline__ void teno5_weights(const double b[3], double& c0, double& c1,
                                              double& c2) {
    const double tau5 = fabs(b[0] - b[2]);
    const double g0 = pow6(1.0 + fdiv(tau5, b[0] + 1.0e-40));
    const double g1 = pow6(1.0 + fdiv(tau5, b[1] + 1.0e-40));
    const double g2 = pow6(1.0 + fdiv(tau5, b[2] + 1.0e-40));
    const double gs = g0 + g1 + g2;
    c0 = teno_cut(g0, gs) ? 0.0 : 1.0;
    c1 = teno_cut(g1, gs) ? 0.0 : 1.0;
    c2 = teno_cut(g2, gs) ? 0.0 : 1.0;
}

__device__ __forceinline__ void teno5_pair(double c0, double c1, double c2, double c3,
                                           double c4, double c5, double& left, double& right,
                                           unsigned& flags) {
    const PairTerms t = pair_terms(c0, c1, c2, c3, c4, c5);
    double l0, l1, l2, r0, r1, r2;
    teno5_weights(t.bL, l0, l1, l2);
    teno5_weights(t.bR, r0, r1, r2);
    const double aL = 0.1 * l0 + 0.6 * l1 + 0.3 * l2;
    const double aR = 0.1 * r0 + 0.6 * r1 + 0.3 * r2;
    if (!(aL > 0.0) || !(aR > 0.0)) flags |= kFlagTeno;
    const double nL = 0.1 * l0 * t.qL0 + 0.6 * l1 * t.Sa + 0.3 * l2 * t.Sb;
    const double nR = 0.1 * r0 * t.qR0 + 0.6 * r1 * t.Sb + 0.3 * r2 * t.Sa;
    div_pair(nL, 6.0 * aL, nR, 6.0 * aR, left, right);
}

template <int SCHEME>
__device__ __forceinline__ void recon_pair(const double ch[6], double& left, double& right,
                                           unsigned& flags) {
    if constexpr (SCHEME == 1) teno5_pair(ch[0], ch[1], ch[2], ch[3], ch[4], ch[5], left, right, flags);
    else weno5_pair(ch[0], ch[1], ch[2], ch[3], ch[4], ch[5], left, right);
}

template <int SCHEME>
__device__ __forceinline__ double recon(double v0, double v1, double v2,
                                        double v3, double v4, unsigned& flags) {
    if constexpr (SCHEME == 1) return teno5(v0, v1, v2, v3, v4, flags);
    else return weno5(v0, v1, v2, v3, v4);
}

// Closure of SingleFluidStateModel::close + PerfectGasEOS
// (SF_singleFluidStateModel.cpp:49-71, SF_perfectGasEOS.h:24-32).
// One reciprocal of the density replaces the five divisions by rho.
struct Closed {
    double u[3], p, c, inv_rho;
};

__device__ __forceinline__ Closed close_state(const double q[5], double gamma) {
    Closed s;
    const double rho = q[0];
    s.inv_rho = frcp(rho);
    s.u[0] = q[1] * s.inv_rho;
    s.u[1] = q[2] * s.inv_rho;
    s.u[2] = q[3] * s.inv_rho;
    const double kinetic = 0.5 * rho * (s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]);
    const double e = (q[4] - kinetic) * s.inv_rho;
    s.p = (gamma - 1.0) * rho * e;
    s.c = fsqrt(gamma * s.p * s.inv_rho);
    return s;
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
template <int ND>
__device__ __forceinline__ void rusanov(const double qL[5], const double qR[5],
                                        double gamma, double out[5]) {
    const Closed l = close_state(qL, gamma), r = close_state(qR, gamma);
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
template <int ND>
__device__ __forceinline__ void hllc(const double qL[5], const double qR[5],
                                     double gamma, double out[5]) {
    const Closed l = close_state(qL, gamma), r = close_state(qR, gamma);
    const double unL = l.u[ND], unR = r.u[ND];
    const double sL = fsqrt(qL[0]), sR = fsqrt(qR[0]);
    const double inv = frcp(sL + sR);
    double ut[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ut[d] = (sL * l.u[d] + sR * r.u[d]) * inv;
    const double HL = (qL[4] + l.p) * l.inv_rho, HR = (qR[4] + r.p) * r.inv_rho;
    const double Ht = (sL * HL + sR * HR) * inv;
    const double ct = fsqrt((gamma - 1.0) * (Ht - 0.5 * (ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2])));
    const double SL = fmin(unL - l.c, ut[ND] - ct);
    const double SR = fmax(unR + r.c, ut[ND] + ct);
    const double mL = qL[0] * (SL - unL), mR = qR[0] * (SR - unR);
    const double Ss = fdiv(r.p - l.p + unL * mL - unR * mR, mL - mR);
    if (0.0 <= SL) {
        physical_flux<ND>(qL, l, out);
        return;
    }
    if (SR <= 0.0) {
        physical_flux<ND>(qR, r, out);
        return;
    }
    const bool left = 0.0 <= Ss;
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
    const double rs = frcp(SK - Ss);
    const double scale = mK * rs;
    double us[5];
    us[0] = scale;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        us[1 + d] = scale * ((d == ND) ? s.u[d] + (Ss - un) : s.u[d]);
    us[4] = scale * (q[4] * s.inv_rho + (Ss - un) * Ss) + (Ss - un) * (s.p * rs);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = f[v] + SK * (us[v] - q[v]);
}

// Flux::characteristicWENOFaceFlux<6> (SF_WENO.h:185-245) on a face of
// axis AXIS whose cofactor is area * e_AXIS.  st[s][v] holds the stencil
// nodes at offsets -2..+3 around the face.  Returns the area-weighted flux.
template <int AXIS, int SCHEME, int FLUX>
__device__ __forceinline__ void face_flux(const double st[6][5], double gamma,
                                          double area, double out[5],
                                          unsigned& flags) {
    using B = Basis<AXIS>;
    constexpr int nd = B::nd, li = B::li, mi = B::mi;
    constexpr double sl = B::sl, sm = B::sm;

    // Face average (SF_WENO.h:207-211) and the eigenvector matrices of
    // Riemann::buildCharacteristicMatrix (SF_riemann.cpp:69-123).
    double qa[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) qa[v] = 0.5 * (st[2][v] + st[3][v]);
    const double invR = frcp(qa[0]);
    const double ke = 0.5 * (qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]) * invR;
    const double p = (gamma - 1.0) * (qa[4] - ke);
    if (!(qa[0] > 0.0) || !(p > 0.0)) flags |= kFlagChar;
    double u[3];
    u[0] = qa[1] * invR;
    u[1] = qa[2] * invR;
    u[2] = qa[3] * invR;
    const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double cz = gamma * p * invR;
    const double rc = frsqrt(cz);
    const double c = cz * rc;
    const double H = (qa[4] + p) * invR;
    const double un = u[nd];
    const double ul = sl * u[li];
    const double um = sm * u[mi];
    const double b1 = (gamma - 1.0) * (rc * rc);
    const double b2 = 0.5 * b1;
    const double x1 = 0.5 * rc;

    // Characteristic projections (gemv5 with L, SF_riemann.h:33-37).  With
    // P = b2 (q2/2 rho - u.m + E) and X = x1 (un rho - m_n), the rows of L
    // give w0 = rho - 2P (acoustic-free row), w3 = P + X, w4 = P - X, and
    // the shear rows w1 = sl (m_l - u_l rho), w2 = sm (m_m - u_m rho).
    const double hq2 = 0.5 * q2;
    double ch[5][6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const double r = st[s][0];
        const double um_dot = u[0] * st[s][1] + u[1] * st[s][2] + u[2] * st[s][3];
        const double P = b2 * (hq2 * r - um_dot + st[s][4]);
        const double X = x1 * (un * r - st[s][1 + nd]);
        ch[0][s] = r - 2.0 * P;
        ch[1][s] = -ul * r + sl * st[s][1 + li];
        ch[2][s] = -um * r + sm * st[s][1 + mi];
        ch[3][s] = P + X;
        ch[4][s] = P - X;
    }
    double wL[5], wR[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) recon_pair<SCHEME>(ch[v], wL[v], wR[v], flags);

    // Back to conservative variables with R (gemv5).  Every column of R
    // carries u in the momentum rows except the acoustic +-c on the normal
    // and the shear units, so q_m = u q_rho + (the rest).
    double qL[5], qR[5];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const double* w = side ? wR : wL;
        double* q = side ? qR : qL;
        const double r = w[0] + w[3] + w[4];
        const double dw = w[4] - w[3];
        q[0] = r;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d == nd) q[1 + d] = u[d] * r + c * dw;
            else if (d == li) q[1 + d] = u[d] * r + sl * w[1];
            else q[1 + d] = u[d] * r + sm * w[2];
        }
        q[4] = hq2 * w[0] + ul * w[1] + um * w[2] + H * (w[3] + w[4]) + (c * un) * dw;
    }

    double f[5];
    if constexpr (FLUX == 1) hllc<nd>(qL, qR, gamma, f);
    else rusanov<nd>(qL, qR, gamma, f);
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = f[v] * area;
}

}  // namespace sfamd
