/*
 * sfcpu.c -- CPU FP64 restatement of sonicSolver's density-based
 * single-fluid explicit hot path.  TEST INFRASTRUCTURE ONLY (the checker);
 * see sfcpu.h.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so the
 * arithmetic is evaluated exactly in source order without FMA contraction).
 *
 * Citations are relative to the sonicSolver source tree.
 */
#include "sfcpu.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double sq(double x) { return x * x; }

/* ---------------------------------------------------------------------- */
/* Scalar reconstructions                                                  */
/* ---------------------------------------------------------------------- */

/* WENO5-JS, Jiang & Shu (1996).  Follows WENO5::weno5_core,
 * src/methods/numerics/convection/SF_WENO5.cpp:42-62 (eps = 1e-6,
 * linear weights 0.1/0.6/0.3, alpha = d/(eps+beta)^2). */
double sfo_weno5_core(const double v[5], unsigned *err)
{
    for (int s = 0; s < 5; ++s)
        if (!isfinite(v[s])) { *err |= SFO_ERR_WENO; return NAN; }
    double p0 = (2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2]) / 6.0;
    double p1 = (-1.0 * v[1] + 5.0 * v[2] + 2.0 * v[3]) / 6.0;
    double p2 = (2.0 * v[2] + 5.0 * v[3] - 1.0 * v[4]) / 6.0;
    const double d0 = 0.1, d1 = 0.6, d2 = 0.3;
    double b0 = 13.0 / 12.0 * sq(v[0] - 2.0 * v[1] + v[2])
              + 0.25 * sq(v[0] - 4.0 * v[1] + 3.0 * v[2]);
    double b1 = 13.0 / 12.0 * sq(v[1] - 2.0 * v[2] + v[3])
              + 0.25 * sq(v[1] - v[3]);
    double b2 = 13.0 / 12.0 * sq(v[2] - 2.0 * v[3] + v[4])
              + 0.25 * sq(3.0 * v[2] - 4.0 * v[3] + v[4]);
    const double eps = 1e-6;
    double a0 = d0 / sq(eps + b0);
    double a1 = d1 / sq(eps + b1);
    double a2 = d2 / sq(eps + b2);
    double sum = a0 + a1 + a2;
    if (!isfinite(sum) || sum <= 0.0) { *err |= SFO_ERR_WENO; return NAN; }
    return (a0 * p0 + a1 * p1 + a2 * p2) / sum;
}

/* TENO5, Fu, Hu & Adams (2016) with C = 1, q = 6, CT = 1e-5.  Follows
 * TENO5::teno5_core, src/methods/numerics/convection/SF_TENO5.cpp:44-92.
 * All-cutoff is an error there (process exit); here it raises a bit. */
double sfo_teno5_core(const double v[5], unsigned *err)
{
    for (int s = 0; s < 5; ++s)
        if (!isfinite(v[s])) { *err |= SFO_ERR_WENO; return NAN; }
    const double p0 = (2.0 * v[0] - 7.0 * v[1] + 11.0 * v[2]) / 6.0;
    const double p1 = (-1.0 * v[1] + 5.0 * v[2] + 2.0 * v[3]) / 6.0;
    const double p2 = (2.0 * v[2] + 5.0 * v[3] - 1.0 * v[4]) / 6.0;
    const double b0 = 13.0 / 12.0 * sq(v[0] - 2.0 * v[1] + v[2])
                    + 0.25 * sq(v[0] - 4.0 * v[1] + 3.0 * v[2]);
    const double b1 = 13.0 / 12.0 * sq(v[1] - 2.0 * v[2] + v[3])
                    + 0.25 * sq(v[1] - v[3]);
    const double b2 = 13.0 / 12.0 * sq(v[2] - 2.0 * v[3] + v[4])
                    + 0.25 * sq(3.0 * v[2] - 4.0 * v[3] + v[4]);
    const double d0 = 0.1, d1 = 0.6, d2 = 0.3;
    const double eps = 1.0e-40, ct = 1.0e-5;
    const double tau5 = fabs(b0 - b2);
    const double g0 = pow(1.0 + tau5 / (b0 + eps), 6.0);
    const double g1 = pow(1.0 + tau5 / (b1 + eps), 6.0);
    const double g2 = pow(1.0 + tau5 / (b2 + eps), 6.0);
    const double gs = g0 + g1 + g2;
    if (!isfinite(gs) || gs <= 0.0) { *err |= SFO_ERR_WENO; return NAN; }
    const double c0 = (g0 / gs < ct) ? 0.0 : 1.0;
    const double c1 = (g1 / gs < ct) ? 0.0 : 1.0;
    const double c2 = (g2 / gs < ct) ? 0.0 : 1.0;
    const double active = d0 * c0 + d1 * c1 + d2 * c2;
    if (active <= 0.0) { *err |= SFO_ERR_TENO; return NAN; }
    return (d0 * c0 * p0 + d1 * c1 * p1 + d2 * c2 * p2) / active;
}

/* Third-order WENO as the reference defines it: two three-point
 * candidates (the WENO5 candidates S1, S2 shifted by one node), linear
 * weights 1/3, 2/3, beta_k the squared central difference over the
 * candidate, eps = 1e-6.  Follows WENO3::weno3_core,
 * src/methods/numerics/convection/SF_WENO3.cpp:42-57 (v = offsets -1..2). */
double sfo_weno3_core(const double v[4], unsigned *err)
{
    for (int s = 0; s < 4; ++s)
        if (!isfinite(v[s])) { *err |= SFO_ERR_WENO; return NAN; }
    double p0 = (-v[0] + 5.0 * v[1] + 2.0 * v[2]) / 6.0;
    double p1 = (2.0 * v[1] + 5.0 * v[2] - v[3]) / 6.0;
    const double d0 = 1.0 / 3.0, d1 = 2.0 / 3.0;
    double b0 = (v[2] - v[0]) * (v[2] - v[0]);
    double b1 = (v[3] - v[1]) * (v[3] - v[1]);
    const double eps = 1e-6;
    double a0 = d0 / ((eps + b0) * (eps + b0));
    double a1 = d1 / ((eps + b1) * (eps + b1));
    double sum = a0 + a1;
    if (!isfinite(sum) || sum <= 0.0) { *err |= SFO_ERR_WENO; return NAN; }
    return (a0 * p0 + a1 * p1) / sum;
}

/* Smoothness of the cubic through four equally spaced values (x = 0..3):
 * the integral over [a, b] of the squared first, second and third
 * derivatives, with the cubic in Newton form (forward differences d1,
 * d2/2, d3/6).  Follows beta4,
 * src/methods/numerics/convection/SF_WENO7.cpp:43-68, term by term; the
 * interval terms F(b) - F(a) for F = x, x^2, x^3/3, x^4/2, x^5/5 are formed
 * by the same expressions, once per interval. */
typedef struct { double i1, i2, i3, i4, i5; } w7_iv;

#define W7_IV(a, b)                                                                    \
    { (b) - (a), (b) * (b) - (a) * (a), (b) * (b) * (b) / 3.0 - (a) * (a) * (a) / 3.0, \
      ((b) * (b)) * ((b) * (b)) / 2.0 - ((a) * (a)) * ((a) * (a)) / 2.0,              \
      (b) * (b) * (b) * (b) * (b) / 5.0 - (a) * (a) * (a) * (a) * (a) / 5.0 }

static const w7_iv W7_INTERVALS[4] = {W7_IV(2.5, 3.5), W7_IV(1.5, 2.5), W7_IV(0.5, 1.5),
                                      W7_IV(-0.5, 0.5)};

static double w7_beta_iv(const double v[4], const w7_iv *I)
{
    double d1 = v[1] - v[0];
    double d2 = (v[2] - 2.0 * v[1] + v[0]) * 0.5;
    double d3 = (v[3] - 3.0 * v[2] + 3.0 * v[1] - v[0]) / 6.0;
    double A = d1 - d2 + 2.0 * d3;
    double B = 2.0 * d2 - 6.0 * d3;
    double C = 3.0 * d3;
    double beta1 = A * A * I->i1 + A * B * I->i2 + (B * B + 2.0 * A * C) * I->i3
                 + B * C * I->i4 + C * C * I->i5;
    double D = 2.0 * d2 - 6.0 * d3;
    double E = 6.0 * d3;
    double beta2 = D * D * I->i1 + D * E * I->i2 + E * E * I->i3;
    double beta3 = 36.0 * d3 * d3 * I->i1;
    return beta1 + beta2 + beta3;
}

static double w7_beta(const double v[4], double a, double b)
{
    const w7_iv I = W7_IV(a, b);
    return w7_beta_iv(v, &I);
}

double sfo_weno7_smoothness(const double v[4], double a, double b)
{
    return w7_beta(v, a, b);
}

/* Seventh-order WENO-JS (Balsara and Shu 2000 candidates and linear
 * weights 1/35, 12/35, 18/35, 4/35; eps = 1e-6; alpha = d/(eps+beta)^2).
 * Follows WENO7::weno7_core,
 * src/methods/numerics/convection/SF_WENO7.cpp:70-98 (v = offsets -3..3;
 * the smoothness interval of candidate k is the cell of node 0 in the
 * local coordinate of that candidate). */
double sfo_weno7_core(const double v[7], unsigned *err)
{
    for (int s = 0; s < 7; ++s)
        if (!isfinite(v[s])) { *err |= SFO_ERR_WENO; return NAN; }
    double p0 = (-3.0 * v[0] + 13.0 * v[1] - 23.0 * v[2] + 25.0 * v[3]) / 12.0;
    double p1 = (v[1] - 5.0 * v[2] + 13.0 * v[3] + 3.0 * v[4]) / 12.0;
    double p2 = (-v[2] + 7.0 * v[3] + 7.0 * v[4] - v[5]) / 12.0;
    double p3 = (3.0 * v[3] + 13.0 * v[4] - 5.0 * v[5] + v[6]) / 12.0;
    const double d0 = 1.0 / 35.0, d1 = 12.0 / 35.0, d2 = 18.0 / 35.0, d3 = 4.0 / 35.0;
    double b0 = w7_beta_iv(v + 0, &W7_INTERVALS[0]);
    double b1 = w7_beta_iv(v + 1, &W7_INTERVALS[1]);
    double b2 = w7_beta_iv(v + 2, &W7_INTERVALS[2]);
    double b3 = w7_beta_iv(v + 3, &W7_INTERVALS[3]);
    const double eps = 1e-6;
    double a0 = d0 / ((eps + b0) * (eps + b0));
    double a1 = d1 / ((eps + b1) * (eps + b1));
    double a2 = d2 / ((eps + b2) * (eps + b2));
    double a3 = d3 / ((eps + b3) * (eps + b3));
    double sum = a0 + a1 + a2 + a3;
    if (!isfinite(sum) || sum <= 0.0) { *err |= SFO_ERR_WENO; return NAN; }
    return (a0 * p0 + a1 * p1 + a2 * p2 + a3 * p3) / sum;
}

/* builtInConvectionRecipe halo widths,
 * src/core/config/types/SF_termRecipe.h:217-218; stencil offsets
 * src/methods/numerics/structured/SF_stencil.h:12-17. */
int sfo_stencil_size(int scheme)
{
    return scheme == SFO_WENO3 ? 4 : scheme == SFO_WENO7 ? 8 : 6;
}

int sfo_halo(int scheme)
{
    return scheme == SFO_WENO3 ? 2 : scheme == SFO_WENO7 ? 4 : 3;
}

/* ---------------------------------------------------------------------- */
/* Thermodynamic closures                                                  */
/* ---------------------------------------------------------------------- */

/* Legacy pressure with positivity check, the form used by the
 * characteristic matrix.  Numerics::pressure + requirePhysicalState,
 * src/methods/numerics/SF_utility.h:24-29,50-66. */
static int legacy_pressure(const double q[5], double gamma, double *p)
{
    int bad = !isfinite(q[0]) || !isfinite(q[1]) || !isfinite(q[2])
           || !isfinite(q[3]) || !isfinite(q[4]) || q[0] <= 0.0;
    if (bad) return -1;
    double invRho = 1.0 / q[0];
    double ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) * invRho;
    double pr = (gamma - 1.0) * (q[4] - ke);
    if (!isfinite(pr) || pr <= 0.0) return -1;
    *p = pr;
    return 0;
}

typedef struct { double rho, u[3], p, c; } closed_t;

/* SingleFluidStateModel::close + PerfectGasEOS::fromDensityEnergy,
 * src/models/physics/fluidStateModel/SF_singleFluidStateModel.cpp:49-71,
 * src/models/physics/EOS/SF_perfectGasEOS.h:24-32 (reference energy 0).
 * Temperature positivity is equivalent to pressure positivity for rho>0. */
static int close_state(const double q[5], double gamma, closed_t *s)
{
    if (!isfinite(q[0]) || q[0] <= 0.0) return -1;
    for (int d = 1; d < 4; ++d) if (!isfinite(q[d])) return -1;
    const double rho = q[0];
    s->rho = rho;
    for (int d = 0; d < 3; ++d) s->u[d] = q[1 + d] / rho;
    const double kinetic = 0.5 * rho
        * (s->u[0] * s->u[0] + s->u[1] * s->u[1] + s->u[2] * s->u[2]);
    const double internal = q[4] - kinetic;
    const double e = internal / rho;
    const double thermal = e - 0.0;
    const double p = (gamma - 1.0) * rho * thermal;
    const double c = sqrt(gamma * p / rho);
    if (!isfinite(e) || !isfinite(p) || p <= 0.0 || !isfinite(c) || c <= 0.0)
        return -1;
    s->p = p;
    s->c = c;
    return 0;
}

/* SingleFluidStateModel::totalEnergyFromPressure ->
 * PerfectGasEOS::fromDensityPressure -> totalEnergyFromThermo,
 * SF_singleFluidStateModel.cpp:73-109, SF_perfectGasEOS.h:44-56. */
static double energy_from_pressure(const double q[5], double gamma, double p)
{
    const double thermal = p / ((gamma - 1.0) * q[0]);
    const double e = 0.0 + thermal;
    const double inverseDensity = 1.0 / q[0];
    const double kinetic = 0.5 * inverseDensity
        * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    return q[0] * e + kinetic;
}

/* ---------------------------------------------------------------------- */
/* Characteristic decomposition                                            */
/* ---------------------------------------------------------------------- */

/* Riemann::buildCharacteristicMatrix,
 * src/methods/numerics/convection/SF_riemann.cpp:69-123.  The tangent
 * basis (l, m) is chosen exactly as there: l from the two components of
 * the normal that exclude the smaller of |n2|, |n3|, and m = n x l. */
int sfo_char_matrix(const double Q[5], const double n[3], double gamma,
                    double L[25], double R[25])
{
    double p;
    if (legacy_pressure(Q, gamma, &p)) return -1;
    double r = Q[0];
    double invR = 1.0 / r;
    double u = Q[1] * invR, v = Q[2] * invR, w = Q[3] * invR;
    double e = Q[4];
    double q2 = u * u + v * v + w * w;
    double c = sqrt(gamma * p * invR);
    double H = (e + p) * invR;
    double n1 = n[0], n2 = n[1], n3 = n[2];
    double l1, l2, l3, m1, m2, m3;
    if (fabs(n3) <= fabs(n2)) {
        double ss = sqrt(n1 * n1 + n2 * n2);
        l1 = -n2 / ss; l2 = n1 / ss; l3 = 0.0;
    } else {
        double ss = sqrt(n1 * n1 + n3 * n3);
        l1 = -n3 / ss; l2 = 0.0; l3 = n1 / ss;
    }
    m1 = n2 * l3 - n3 * l2;
    m2 = n3 * l1 - n1 * l3;
    m3 = n1 * l2 - n2 * l1;
    double un = u * n1 + v * n2 + w * n3;
    double ul = u * l1 + v * l2 + w * l3;
    double um = u * m1 + v * m2 + w * m3;
    double b1 = (gamma - 1.0) / (c * c);
    double b2 = 0.5 * b1;
    double x1 = 1.0 / (2.0 * c);
    double Lm[25] = {
        1.0 - b2 * q2, b1 * u, b1 * v, b1 * w, -b1,
        -ul, l1, l2, l3, 0.0,
        -um, m1, m2, m3, 0.0,
        0.5 * b2 * q2 + x1 * un, -x1 * n1 - b2 * u, -x1 * n2 - b2 * v,
        -x1 * n3 - b2 * w, b2,
        0.5 * b2 * q2 - x1 * un, x1 * n1 - b2 * u, x1 * n2 - b2 * v,
        x1 * n3 - b2 * w, b2};
    double Rm[25] = {
        1.0, 0.0, 0.0, 1.0, 1.0,
        u, l1, m1, u - c * n1, u + c * n1,
        v, l2, m2, v - c * n2, v + c * n2,
        w, l3, m3, w - c * n3, w + c * n3,
        0.5 * q2, ul, um, H - c * un, H + c * un};
    memcpy(L, Lm, sizeof Lm);
    memcpy(R, Rm, sizeof Rm);
    return 0;
}

/* Riemann::gemv5, src/methods/numerics/convection/SF_riemann.h:33-37. */
void sfo_gemv5(const double M[25], const double V[5], double out[5])
{
    for (int i = 0; i < 5; ++i)
        out[i] = M[i * 5 + 0] * V[0] + M[i * 5 + 1] * V[1]
               + M[i * 5 + 2] * V[2] + M[i * 5 + 3] * V[3]
               + M[i * 5 + 4] * V[4];
}

/* ---------------------------------------------------------------------- */
/* Face fluxes                                                             */
/* ---------------------------------------------------------------------- */

/* physicalFlux, src/methods/numerics/Flux/SF_rusanovEOS.cpp:14-32. */
static void physical_flux(const double q[5], const closed_t *s,
                          const double n[3], double f[5])
{
    const double un = s->u[0] * n[0] + s->u[1] * n[1] + s->u[2] * n[2];
    f[0] = q[0] * un;
    for (int d = 0; d < 3; ++d) f[1 + d] = q[1 + d] * un + s->p * n[d];
    f[4] = (q[4] + s->p) * un;
}

int sfo_physical_flux(const double q[5], const double n[3], double gamma,
                      double out[5])
{
    closed_t s;
    if (close_state(q, gamma, &s)) return -1;
    physical_flux(q, &s, n, out);
    return 0;
}

/* RusanovEOS::faceFlux, src/methods/numerics/Flux/SF_rusanovEOS.cpp:47-77. */
int sfo_rusanov(const double qL[5], const double qR[5], const double n[3],
                double gamma, double out[5])
{
    closed_t l, r;
    if (close_state(qL, gamma, &l) || close_state(qR, gamma, &r)) return -1;
    double fL[5], fR[5];
    physical_flux(qL, &l, n, fL);
    physical_flux(qR, &r, n, fR);
    const double unL = l.u[0] * n[0] + l.u[1] * n[1] + l.u[2] * n[2];
    const double unR = r.u[0] * n[0] + r.u[1] * n[1] + r.u[2] * n[2];
    const double a = fabs(unL) + l.c, b = fabs(unR) + r.c;
    const double lambda = a > b ? a : b;   /* std::max(a, b) */
    if (!isfinite(lambda) || lambda <= 0.0) return -1;
    for (int v = 0; v < 5; ++v)
        out[v] = 0.5 * (fL[v] + fR[v]) - 0.5 * lambda * (qR[v] - qL[v]);
    return 0;
}

/* HLLC (no reference counterpart: the flux enum of the reference has no
 * HLLC member).  Toro, 3rd ed., section 10.4-10.6:
 *  - wave speeds: Einfeldt (1988) estimates from Roe averages,
 *      S_L = min(u_nL - c_L, u~_n - c~),  S_R = max(u_nR + c_R, u~_n + c~);
 *  - contact speed S* from Toro eq. (10.37);
 *  - star states from Toro eq. (10.39), written for a unit normal n so the
 *    tangential velocity of each side is carried unchanged;
 *  - flux from Toro eq. (10.38) (the four-region choice by signs of
 *    S_L, S*, S_R). */
int sfo_hllc(const double qL[5], const double qR[5], const double n[3],
             double gamma, double out[5])
{
    closed_t l, r;
    if (close_state(qL, gamma, &l) || close_state(qR, gamma, &r)) return -1;
    const double unL = l.u[0] * n[0] + l.u[1] * n[1] + l.u[2] * n[2];
    const double unR = r.u[0] * n[0] + r.u[1] * n[1] + r.u[2] * n[2];
    const double sL = sqrt(l.rho), sR = sqrt(r.rho);
    const double inv = 1.0 / (sL + sR);
    double ut[3];
    for (int d = 0; d < 3; ++d) ut[d] = (sL * l.u[d] + sR * r.u[d]) * inv;
    const double HL = (qL[4] + l.p) / l.rho, HR = (qR[4] + r.p) / r.rho;
    const double Ht = (sL * HL + sR * HR) * inv;
    const double ct2 = (gamma - 1.0)
        * (Ht - 0.5 * (ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]));
    if (!(ct2 > 0.0) || !isfinite(ct2)) return -1;
    const double ct = sqrt(ct2);
    const double unt = ut[0] * n[0] + ut[1] * n[1] + ut[2] * n[2];
    const double SL = fmin(unL - l.c, unt - ct);
    const double SR = fmax(unR + r.c, unt + ct);
    const double mL = l.rho * (SL - unL), mR = r.rho * (SR - unR);
    const double den = mL - mR;
    if (!(den != 0.0)) return -1;
    const double Ss = (r.p - l.p + unL * mL - unR * mR) / den;
    if (!isfinite(SL) || !isfinite(SR) || !isfinite(Ss)) return -1;
    double f[5];
    if (0.0 <= SL) {
        physical_flux(qL, &l, n, out);
        return 0;
    }
    if (SR <= 0.0) {
        physical_flux(qR, &r, n, out);
        return 0;
    }
    const int left = 0.0 <= Ss;
    const double *q = left ? qL : qR;
    const closed_t *s = left ? &l : &r;
    const double un = left ? unL : unR;
    const double SK = left ? SL : SR;
    const double mK = left ? mL : mR;
    physical_flux(q, s, n, f);
    const double scale = mK / (SK - Ss);
    double us[5];
    us[0] = scale;
    for (int d = 0; d < 3; ++d) us[1 + d] = scale * (s->u[d] + (Ss - un) * n[d]);
    us[4] = scale * (q[4] / s->rho + (Ss - un) * (Ss + s->p / mK));
    for (int v = 0; v < 5; ++v) out[v] = f[v] + SK * (us[v] - q[v]);
    return 0;
}

/* Steger-Warming flux-vector splitting of one state with the reference's
 * smoothed eigenvalue split (|lambda| < 2 eps_sw, eps_sw = 0.01, an
 * absolute threshold in velocity units).  Follows
 * Riemann::splitStegerWarming,
 * src/methods/numerics/convection/SF_riemann.cpp:15-67. */
static int steger_split(const double Q[5], const double n[3], double gamma,
                        double fp[5], double fm[5])
{
    double p;
    if (legacy_pressure(Q, gamma, &p)) return -1;
    double r = Q[0];
    double invR = 1.0 / r;
    double u = Q[1] * invR, v = Q[2] * invR, w = Q[3] * invR;
    double q2 = u * u + v * v + w * w;
    double a = sqrt(gamma * p * invR);
    double un = u * n[0] + v * n[1] + w * n[2];
    double lam[5] = {un, un, un, un + a, un - a};
    double lp[5], lm[5];
    const double eps_sw = 0.01;
    for (int i = 0; i < 5; ++i) {
        if (fabs(lam[i]) < 2.0 * eps_sw) {
            lp[i] = 0.5 * (lam[i] * lam[i] / (4.0 * eps_sw) + lam[i] + eps_sw);
            lm[i] = lam[i] - lp[i];
        } else {
            lp[i] = 0.5 * (lam[i] + fabs(lam[i]));
            lm[i] = 0.5 * (lam[i] - fabs(lam[i]));
        }
    }
    double common = r / (2.0 * gamma);
    const double *ls[2] = {lp, lm};
    double *out[2] = {fp, fm};
    for (int side = 0; side < 2; ++side) {
        const double *l = ls[side];
        double *f = out[side];
        double W1 = 2.0 * (gamma - 1.0) * l[0] + l[3] + l[4];
        double W2 = l[3] - l[4];
        double W3 = l[3] + l[4];
        f[0] = common * W1;
        f[1] = common * (u * W1 + a * n[0] * W2);
        f[2] = common * (v * W1 + a * n[1] * W2);
        f[3] = common * (w * W1 + a * n[2] * W2);
        f[4] = common * (0.5 * q2 * W1 + a * un * W2 + (a * a / (gamma - 1.0)) * W3);
    }
    return 0;
}

/* StegerWarmingFlux::flux, src/methods/numerics/Flux/SF_sw.cpp:21-27:
 * F+ of the left state plus F- of the right state. */
int sfo_steger_warming(const double qL[5], const double qR[5], const double n[3],
                       double gamma, double out[5])
{
    double fp[5], fm[5], unused[5];
    if (steger_split(qL, n, gamma, fp, unused)) return -1;
    if (steger_split(qR, n, gamma, unused, fm)) return -1;
    for (int v = 0; v < 5; ++v) out[v] = fp[v] + fm[v];
    return 0;
}

/* Local Lax-Friedrichs splitting F+- = (F(Q) +- lambda Q)/2 with
 * lambda = |u_n| + a of the same state.  Follows
 * Riemann::splitLaxFriedrichs,
 * src/methods/numerics/convection/SF_riemann.cpp:125-154. */
static int laxf_split(const double Q[5], const double n[3], double gamma,
                      double fp[5], double fm[5])
{
    double p;
    if (legacy_pressure(Q, gamma, &p)) return -1;
    double r = Q[0];
    double invR = 1.0 / r;
    double u = Q[1] * invR, v = Q[2] * invR, w = Q[3] * invR;
    double e = Q[4];
    double a = sqrt(gamma * p * invR);
    double un = u * n[0] + v * n[1] + w * n[2];
    double lam = fabs(un) + a;
    double F[5];
    F[0] = r * un;
    F[1] = Q[1] * un + p * n[0];
    F[2] = Q[2] * un + p * n[1];
    F[3] = Q[3] * un + p * n[2];
    F[4] = (e + p) * un;
    for (int i = 0; i < 5; ++i) {
        fp[i] = 0.5 * (F[i] + lam * Q[i]);
        fm[i] = 0.5 * (F[i] - lam * Q[i]);
    }
    return 0;
}

/* LaxFriedrichsFlux::flux, src/methods/numerics/Flux/SF_lxF.cpp:21-27. */
int sfo_lax_friedrichs(const double qL[5], const double qR[5], const double n[3],
                       double gamma, double out[5])
{
    double fp[5], fm[5], unused[5];
    if (laxf_split(qL, n, gamma, fp, unused)) return -1;
    if (laxf_split(qR, n, gamma, unused, fm)) return -1;
    for (int v = 0; v < 5; ++v) out[v] = fp[v] + fm[v];
    return 0;
}

/* physicalEulerFlux, src/methods/numerics/Flux/SF_face.h:112-128
 * (velocities by division, legacy pressure). */
static int euler_flux(const double q[5], const double n[3], double gamma,
                      double f[5])
{
    double p;
    if (legacy_pressure(q, gamma, &p)) return -1;
    const double rho = q[0];
    const double u = q[1] / rho, v = q[2] / rho, w = q[3] / rho;
    const double un = u * n[0] + v * n[1] + w * n[2];
    f[0] = rho * un;
    f[1] = rho * u * un + p * n[0];
    f[2] = rho * v * un + p * n[1];
    f[3] = rho * w * un + p * n[2];
    f[4] = (q[4] + p) * un;
    return 0;
}

/* Roe flux with Harten's entropy fix (delta = 0.1 a~) on every wave,
 * F = (F_L + F_R - R |Lambda| L (Q_R - Q_L)) / 2 with L, R of
 * buildCharacteristicMatrix at the Roe-averaged state.  Follows
 * RoeFlux::flux, src/methods/numerics/Flux/SF_roe.cpp:20-111, including
 * two of its choices: the eigenvectors are built at the state
 * (rho~, rho~ u~, rho~ h~), whose energy slot holds rho~ h~, so their sound
 * speed is sqrt(gamma) a~ while the wave speeds use a~; and the speeds come
 * in the order (u_n, u_n, u_n, u_n + a~, u_n - a~) against the columns of
 * R, whose fourth column is the u_n - c wave (SF_riemann.cpp:117-122). */
int sfo_roe(const double qL[5], const double qR[5], const double n[3],
            double gamma, double out[5])
{
    const double gm1 = gamma - 1.0;
    double pL, pR;
    if (legacy_pressure(qL, gamma, &pL) || legacy_pressure(qR, gamma, &pR)) return -1;
    const double invRL = 1.0 / qL[0];
    const double uL = qL[1] * invRL, vL = qL[2] * invRL, wL = qL[3] * invRL;
    const double hL = (qL[4] + pL) * invRL;
    const double invRR = 1.0 / qR[0];
    const double uR = qR[1] * invRR, vR = qR[2] * invRR, wR = qR[3] * invRR;
    const double hR = (qR[4] + pR) * invRR;
    const double sqrtRL = sqrt(qL[0]), sqrtRR = sqrt(qR[0]);
    const double sumSqrt = sqrtRL + sqrtRR;
    const double uRoe = (sqrtRL * uL + sqrtRR * uR) / sumSqrt;
    const double vRoe = (sqrtRL * vL + sqrtRR * vR) / sumSqrt;
    const double wRoe = (sqrtRL * wL + sqrtRR * wR) / sumSqrt;
    const double hRoe = (sqrtRL * hL + sqrtRR * hR) / sumSqrt;
    const double q2Roe = uRoe * uRoe + vRoe * vRoe + wRoe * wRoe;
    const double a2Roe = gm1 * (hRoe - 0.5 * q2Roe);
    if (!isfinite(a2Roe) || a2Roe <= 0.0) return -1;
    const double aRoe = sqrt(a2Roe);
    const double unRoe = uRoe * n[0] + vRoe * n[1] + wRoe * n[2];
    const double lambda[5] = {unRoe, unRoe, unRoe, unRoe + aRoe, unRoe - aRoe};
    double absLambda[5];
    const double entropyEps = 0.1 * aRoe;
    for (int m = 0; m < 5; ++m) {
        const double value = fabs(lambda[m]);
        absLambda[m] = value < 2.0 * entropyEps
            ? 0.5 * (value * value / (2.0 * entropyEps) + entropyEps) : value;
    }
    double dQ[5];
    for (int v = 0; v < 5; ++v) dQ[v] = qR[v] - qL[v];
    const double rhoRoe = sqrtRL * sqrtRR;
    double qRoe[5];
    qRoe[0] = rhoRoe;
    qRoe[1] = rhoRoe * uRoe;
    qRoe[2] = rhoRoe * vRoe;
    qRoe[3] = rhoRoe * wRoe;
    qRoe[4] = rhoRoe * (0.5 * q2Roe + a2Roe / gm1);
    double L[25], R[25];
    if (sfo_char_matrix(qRoe, n, gamma, L, R)) return -1;
    double alpha[5];
    sfo_gemv5(L, dQ, alpha);
    double dis[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int wave = 0; wave < 5; ++wave) {
        const double strength = absLambda[wave] * alpha[wave];
        for (int v = 0; v < 5; ++v) dis[v] += R[v * 5 + wave] * strength;
    }
    double fL[5], fR[5];
    if (euler_flux(qL, n, gamma, fL) || euler_flux(qR, n, gamma, fR)) return -1;
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fL[v] + fR[v] - dis[v]);
    return 0;
}

/* LaxWendroffFlux::flux, src/methods/numerics/Flux/SF_lw.cpp:17-40: the
 * physical flux of the Richtmyer half state
 * q* = (qL + qR)/2 - (dt/spacing)/2 (F(qR) - F(qL)); dt is the step's
 * (AssemblyContext::timeStep = state.dt, SF_conservativeRHS.cpp:167-173). */
int sfo_lax_wendroff(const double qL[5], const double qR[5], const double n[3],
                     double gamma, double dt, double spacing, double out[5])
{
    if (!(dt > 0.0) || !(spacing > 0.0)) return -1;
    double fL[5], fR[5];
    if (euler_flux(qL, n, gamma, fL) || euler_flux(qR, n, gamma, fR)) return -1;
    const double lambda = dt / spacing;
    double qh[5];
    for (int v = 0; v < 5; ++v)
        qh[v] = 0.5 * (qL[v] + qR[v]) - 0.5 * lambda * (fR[v] - fL[v]);
    return euler_flux(qh, n, gamma, out);
}

int sfo_face_flux(int flux, const double qL[5], const double qR[5], const double n[3],
                  double gamma, double out[5])
{
    switch (flux) {
    case SFO_RUSANOV: return sfo_rusanov(qL, qR, n, gamma, out);
    case SFO_HLLC: return sfo_hllc(qL, qR, n, gamma, out);
    case SFO_STEGER: return sfo_steger_warming(qL, qR, n, gamma, out);
    case SFO_LAXF: return sfo_lax_friedrichs(qL, qR, n, gamma, out);
    case SFO_ROE: return sfo_roe(qL, qR, n, gamma, out);
    }
    return -1;
}

/* Flux::characteristicWENOFaceFlux<NS> with the scalar-stencil mapping of
 * fillScalarWENOStencils<NS> (left: nodes 0..w-1, right: nodes NS-1..NS-w,
 * w = 4, 5, 7 for NS = 4, 6, 8),
 * src/solver/discretization/convection/SF_WENO.h:64-98,185-245. */
int sfo_char_face_flux_n(const double *st, const double cof[3], int scheme,
                         int flux, double gamma, double out[5])
{
    return sfo_char_face_flux_dt(st, cof, scheme, flux, gamma, 0.0, 0.0, out);
}

int sfo_char_face_flux_dt(const double *st, const double cof[3], int scheme,
                          int flux, double gamma, double dt, double spacing,
                          double out[5])
{
    const int ns = sfo_stencil_size(scheme);
    const int width = ns == 4 ? 4 : ns == 6 ? 5 : 7;
    const double smag = sqrt(cof[0] * cof[0] + cof[1] * cof[1] + cof[2] * cof[2]);
    if (!isfinite(smag) || smag <= 1.0e-300) return SFO_ERR_ARG;
    const double n[3] = {cof[0] / smag, cof[1] / smag, cof[2] / smag};
    const double area = smag;
    const int cl = ns / 2 - 1;
    double qa[5];
    for (int v = 0; v < 5; ++v) qa[v] = 0.5 * (st[cl * 5 + v] + st[(cl + 1) * 5 + v]);
    double L[25], R[25];
    if (sfo_char_matrix(qa, n, gamma, L, R)) return SFO_ERR_CHAR;
    double ch[8][5];
    for (int s = 0; s < ns; ++s) sfo_gemv5(L, st + 5 * s, ch[s]);
    double wL[5], wR[5];
    unsigned err = 0;
    for (int v = 0; v < 5; ++v) {
        double left[8], right[8];
        for (int s = 0; s < width; ++s) {
            left[s] = ch[s][v];
            right[s] = ch[ns - 1 - s][v];
        }
        switch (scheme) {
        case SFO_TENO5:
            wL[v] = sfo_teno5_core(left, &err);
            wR[v] = sfo_teno5_core(right, &err);
            break;
        case SFO_WENO3:
            wL[v] = sfo_weno3_core(left, &err);
            wR[v] = sfo_weno3_core(right, &err);
            break;
        case SFO_WENO7:
            wL[v] = sfo_weno7_core(left, &err);
            wR[v] = sfo_weno7_core(right, &err);
            break;
        default:
            wL[v] = sfo_weno5_core(left, &err);
            wR[v] = sfo_weno5_core(right, &err);
        }
        if (err) return (int)err;
        if (!isfinite(wL[v]) || !isfinite(wR[v])) return SFO_ERR_WENO;
    }
    double qL[5], qR[5], f[5];
    sfo_gemv5(R, wL, qL);
    sfo_gemv5(R, wR, qR);
    const int frc = flux == SFO_LAXW ? sfo_lax_wendroff(qL, qR, n, gamma, dt, spacing, f)
                                     : sfo_face_flux(flux, qL, qR, n, gamma, f);
    if (frc) return SFO_ERR_FLUX;
    for (int v = 0; v < 5; ++v) out[v] = f[v] * area;
    return 0;
}

int sfo_char_face_flux(const double st[30], const double cof[3], int scheme,
                       int flux, double gamma, double out[5])
{
    if (sfo_stencil_size(scheme) != 6) return SFO_ERR_ARG;
    return sfo_char_face_flux_n(st, cof, scheme, flux, gamma, out);
}

/* ---------------------------------------------------------------------- */
/* Grid                                                                    */
/* ---------------------------------------------------------------------- */

struct sfo_grid {
    sfo_desc desc;
    int g[3], m[3];
    size_t cells;
    double *q, *q0, *k[4], *fx[3];
    double jac, metric[3], cof[3];
    double dt;           /* the step's dt (Lax-Wendroff reads it) */
    unsigned err;
};

static size_t at(const sfo_grid *G, int i, int j, int k)
{
    return ((size_t)k * G->m[1] + j) * G->m[0] + i;
}

/* A node is "solved" (Math::forFluidInterior,
 * src/methods/numerics/structured/SF_iteration.h:37-58) unless it lies on
 * the boundary plane of a transmissive axis, which the boundary
 * applicator marks as a solver-boundary point (SF_applicator.h:162-187). */
/* zeroGradient on the lower (side 0) or upper (side 1) end of axis d */
static int zg_end(const sfo_grid *G, int d, int side)
{
    const int bc = G->desc.bc[d];
    if (!G->desc.active[d]) return 0;
    if (bc == SFO_TRANSMISSIVE) return 1;
    if (bc == SFO_NONE_LO) return side == 1;
    if (bc == SFO_NONE_HI) return side == 0;
    return 0;
}

static int solved(const sfo_grid *G, const int p[3])
{
    for (int d = 0; d < 3; ++d) {
        if (zg_end(G, d, 0) && p[d] == G->g[d]) return 0;
        if (zg_end(G, d, 1) && p[d] == G->g[d] + G->desc.n[d] - 1) return 0;
    }
    return 1;
}

sfo_grid *sfo_grid_create(const sfo_desc *desc)
{
    for (int d = 0; d < 3; ++d) {
        if (desc->n[d] < 1) return NULL;
        if (desc->active[d] && desc->n[d] < 4) return NULL;
        if (!desc->active[d] && desc->n[d] != 1) return NULL;
        if (!(desc->d[d] > 0.0)) return NULL;
    }
    if (desc->scheme < SFO_WENO5 || desc->scheme > SFO_WENO7) return NULL;
    for (int d = 0; d < 3; ++d)
        if (desc->bc[d] < SFO_TRANSMISSIVE || desc->bc[d] > SFO_NONE) return NULL;
    if (desc->flux < SFO_RUSANOV || desc->flux > SFO_LAXW) return NULL;
    sfo_grid *G = (sfo_grid *)calloc(1, sizeof *G);
    if (!G) return NULL;
    G->desc = *desc;
    G->cells = 1;
    for (int d = 0; d < 3; ++d) {
        G->g[d] = desc->active[d] ? sfo_halo(desc->scheme) : 0;
        G->m[d] = desc->n[d] + 2 * G->g[d];
        G->cells *= (size_t)G->m[d];
    }
    size_t bytes = G->cells * 5 * sizeof(double);
    G->q = (double *)calloc(1, bytes);
    G->q0 = (double *)calloc(1, bytes);
    for (int s = 0; s < 4; ++s) G->k[s] = (double *)calloc(1, bytes);
    for (int d = 0; d < 3; ++d) G->fx[d] = (double *)calloc(1, bytes);
    /* uniform Cartesian metrics: xi_x = 1/dx, ..., Jac = 1/J = 1/(dx dy dz)
     * (Mesh::computeMetrics, src/infrastructure/mesh/SF_mesh.cpp:320-338),
     * face cofactor = average of metric/Jac at both nodes
     * (CanonicalFace::geometry, SF_canonicalFace.h:87-97). */
    G->jac = 1.0 / (desc->d[0] * desc->d[1] * desc->d[2]);
    for (int d = 0; d < 3; ++d) {
        G->metric[d] = 1.0 / desc->d[d];
        double c = G->metric[d] / G->jac;
        G->cof[d] = 0.5 * (c + c);
    }
    return G;
}

void sfo_grid_destroy(sfo_grid *G)
{
    if (!G) return;
    free(G->q); free(G->q0);
    for (int s = 0; s < 4; ++s) free(G->k[s]);
    for (int d = 0; d < 3; ++d) free(G->fx[d]);
    free(G);
}

unsigned sfo_error(const sfo_grid *G) { return G->err; }

#define FOR_NODES(G, p)                                                    \
    for (p[2] = (G)->g[2]; p[2] < (G)->g[2] + (G)->desc.n[2]; ++p[2])      \
        for (p[1] = (G)->g[1]; p[1] < (G)->g[1] + (G)->desc.n[1]; ++p[1])  \
            for (p[0] = (G)->g[0]; p[0] < (G)->g[0] + (G)->desc.n[0]; ++p[0])

int sfo_set_state(sfo_grid *G, const double *soa)
{
    const size_t N = (size_t)G->desc.n[0] * G->desc.n[1] * G->desc.n[2];
    int p[3];
    FOR_NODES(G, p) {
        size_t u = ((size_t)(p[2] - G->g[2]) * G->desc.n[1] + (p[1] - G->g[1]))
                 * G->desc.n[0] + (p[0] - G->g[0]);
        double *q = G->q + 5 * at(G, p[0], p[1], p[2]);
        for (int v = 0; v < 5; ++v) q[v] = soa[v * N + u];
    }
    /* an end without a condition: its ghost layers hold the end node's
     * state from here on (nothing refills them) */
    for (int d = 0; d < 3; ++d) {
        if (!G->desc.active[d] || G->desc.bc[d] == SFO_TRANSMISSIVE
            || G->desc.bc[d] == SFO_PERIODIC)
            continue;
        for (int side = 0; side < 2; ++side) {
            if (zg_end(G, d, side)) continue;
            FOR_NODES(G, p) {
                const int end = side == 0 ? G->g[d] : G->g[d] + G->desc.n[d] - 1;
                if (p[d] != end) continue;
                for (int l = 1; l <= G->g[d]; ++l) {
                    int gp[3] = {p[0], p[1], p[2]};
                    gp[d] += side == 0 ? -l : l;
                    memcpy(G->q + 5 * at(G, gp[0], gp[1], gp[2]),
                           G->q + 5 * at(G, p[0], p[1], p[2]), 5 * sizeof(double));
                }
            }
        }
    }
    return 0;
}

int sfo_get_state(const sfo_grid *G, double *soa)
{
    const size_t N = (size_t)G->desc.n[0] * G->desc.n[1] * G->desc.n[2];
    int p[3];
    FOR_NODES(G, p) {
        size_t u = ((size_t)(p[2] - G->g[2]) * G->desc.n[1] + (p[1] - G->g[1]))
                 * G->desc.n[0] + (p[0] - G->g[0]);
        const double *q = G->q + 5 * at(G, p[0], p[1], p[2]);
        for (int v = 0; v < 5; ++v) soa[v * N + u] = q[v];
    }
    return 0;
}

/* Iterate the two boundary planes of axis d over the physical range of the
 * other axes; body gets the boundary node, its inner neighbour and the
 * direction (+1 upward into ghosts on the high side, -1 on the low side). */
typedef void (*plane_fn)(sfo_grid *, const int b[3], const int src[3], int d,
                         int dir);

static void for_planes(sfo_grid *G, int d, plane_fn fn)
{
    int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = G->g[a]; hi[a] = G->g[a] + G->desc.n[a]; }
    for (int side = 0; side < 2; ++side) {
        if (!zg_end(G, d, side)) continue;
        int p[3];
        for (p[e2] = lo[e2]; p[e2] < hi[e2]; ++p[e2])
            for (p[e1] = lo[e1]; p[e1] < hi[e1]; ++p[e1]) {
                int b[3] = {p[0], p[1], p[2]}, s[3];
                b[d] = side == 0 ? G->g[d] : G->g[d] + G->desc.n[d] - 1;
                memcpy(s, b, sizeof s);
                s[d] += side == 0 ? 1 : -1;
                fn(G, b, s, d, side == 0 ? -1 : 1);
            }
    }
}

/* zeroGradient for density and momentum: the boundary node takes its
 * inner neighbour's value and the ghosts take the boundary node's value
 * (Algebraic::ZeroGradient::applyScalar/applyVector3,
 * src/solver/boundary/reconstruction/linear/SF_zeroGradient.cpp:88-130). */
static void zg_copy4(sfo_grid *G, const int b[3], const int s[3], int d, int dir)
{
    double *qb = G->q + 5 * at(G, b[0], b[1], b[2]);
    const double *qs = G->q + 5 * at(G, s[0], s[1], s[2]);
    for (int v = 0; v < 4; ++v) qb[v] = qs[v];
    for (int l = 1; l <= G->g[d]; ++l) {
        int gp[3] = {b[0], b[1], b[2]};
        gp[d] += dir * l;
        double *qg = G->q + 5 * at(G, gp[0], gp[1], gp[2]);
        for (int v = 0; v < 4; ++v) qg[v] = qb[v];
    }
}

static double pressure_at(const sfo_grid *G, const double *q)
{
    closed_t s;
    if (close_state(q, G->desc.gamma, &s)) return NAN;
    return s.p;
}

/* zeroGradient for pressure, converted to total energy:
 * updateEnergyFromPressure ZERO_GRADIENT branch,
 * src/solver/boundary/SF_thermodynamicClosure.cpp:219-281. */
static void zg_energy(sfo_grid *G, const int b[3], const int s[3], int d, int dir)
{
    const double gamma = G->desc.gamma;
    double *qb = G->q + 5 * at(G, b[0], b[1], b[2]);
    const double *qs = G->q + 5 * at(G, s[0], s[1], s[2]);
    double ps = pressure_at(G, qs);
    if (!isfinite(ps) || ps <= 0.0) { G->err |= SFO_ERR_STATE; return; }
    qb[4] = energy_from_pressure(qb, gamma, ps);
    double pb = pressure_at(G, qb);
    if (!isfinite(pb) || pb <= 0.0) { G->err |= SFO_ERR_STATE; return; }
    for (int l = 1; l <= G->g[d]; ++l) {
        int gp[3] = {b[0], b[1], b[2]};
        gp[d] += dir * l;
        double *qg = G->q + 5 * at(G, gp[0], gp[1], gp[2]);
        qg[4] = energy_from_pressure(qg, gamma, pb);
    }
}

/* Periodic fill (no reference counterpart): ghost layer l on either side
 * of axis d copies the unique node at (index mod n). */
static void periodic_fill(sfo_grid *G, int d)
{
    int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
    const int n = G->desc.n[d], g = G->g[d];
    int p[3];
    for (p[e2] = G->g[e2]; p[e2] < G->g[e2] + G->desc.n[e2]; ++p[e2])
        for (p[e1] = G->g[e1]; p[e1] < G->g[e1] + G->desc.n[e1]; ++p[e1])
            for (int l = 0; l < g; ++l) {
                int lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
                int slo[3] = {p[0], p[1], p[2]}, shi[3] = {p[0], p[1], p[2]};
                lo[d] = l;                 slo[d] = g + ((l - g) % n + n) % n;
                hi[d] = g + n + l;         shi[d] = g + l % n;
                memcpy(G->q + 5 * at(G, lo[0], lo[1], lo[2]),
                       G->q + 5 * at(G, slo[0], slo[1], slo[2]), 5 * sizeof(double));
                memcpy(G->q + 5 * at(G, hi[0], hi[1], hi[2]),
                       G->q + 5 * at(G, shi[0], shi[1], shi[2]), 5 * sizeof(double));
            }
}

/* Boundary::Applicator::apply order (SF_applicator.h:51-83): density and
 * momentum for every set, then energy from pressure for every set; sets
 * are visited axis by axis (x, y, z).  Periodic axes are filled last. */
int sfo_apply_bc(sfo_grid *G)
{
    for (int d = 0; d < 3; ++d) for_planes(G, d, zg_copy4);
    for (int d = 0; d < 3; ++d) for_planes(G, d, zg_energy);
    for (int d = 0; d < 3; ++d)
        if (G->desc.active[d] && G->desc.bc[d] == SFO_PERIODIC)
            periodic_fill(G, d);
    return G->err ? -1 : 0;
}

/* Flux::computeDirectionalFlux (SF_WENO.h:257-364) over Math::forFaces
 * (SF_iteration.h:61-82): faces i+1/2 for i = g-1 .. g+n-1 on the axis,
 * physical range on the others; the lower node owns the face
 * (Math::storeFlux, SF_faceAssembly.h:26-40). */
static int compute_fluxes(sfo_grid *G, int d)
{
    int p[3], lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = G->g[a]; hi[a] = G->g[a] + G->desc.n[a]; }
    lo[d] = G->g[d] - 1;
    double cof[3] = {0.0, 0.0, 0.0};
    cof[d] = G->cof[d];
    const int ns = sfo_stencil_size(G->desc.scheme);
    for (p[2] = lo[2]; p[2] < hi[2]; ++p[2])
        for (p[1] = lo[1]; p[1] < hi[1]; ++p[1])
            for (p[0] = lo[0]; p[0] < hi[0]; ++p[0]) {
                double st[40];
                for (int s = 0; s < ns; ++s) {
                    int r[3] = {p[0], p[1], p[2]};
                    r[d] += s + 1 - ns / 2;
                    memcpy(st + 5 * s, G->q + 5 * at(G, r[0], r[1], r[2]),
                           5 * sizeof(double));
                }
                double f[5];
                int rc = sfo_char_face_flux_dt(st, cof, G->desc.scheme, G->desc.flux,
                                               G->desc.gamma, G->dt, G->desc.d[d], f);
                if (rc) { G->err |= (unsigned)rc; return -1; }
                for (int v = 0; v < 5; ++v) {
                    if (!isfinite(f[v])) { G->err |= SFO_ERR_FLUX; return -1; }
                    G->fx[d][5 * at(G, p[0], p[1], p[2]) + v] = f[v];
                }
            }
    return 0;
}

/* Math::localSpatialResidual (SF_residualAssembly.h:12-32) on solved
 * nodes, stored like Time::Explicit::snapshotRHS (SF_explicit.cpp:181-191):
 * R = Jac * sum_active (F_d(i) - F_d(i-1)) - source (source = 0 here). */
static int assemble_rhs(sfo_grid *G, double *out)
{
    for (int d = 0; d < 3; ++d)
        if (G->desc.active[d] && compute_fluxes(G, d)) return -1;
    memset(out, 0, G->cells * 5 * sizeof(double));
    int p[3];
    FOR_NODES(G, p) {
        if (!solved(G, p)) continue;
        size_t c = at(G, p[0], p[1], p[2]);
        for (int v = 0; v < 5; ++v) {
            double div = 0.0;
            for (int d = 0; d < 3; ++d) {
                if (!G->desc.active[d]) continue;
                int r[3] = {p[0], p[1], p[2]};
                r[d] -= 1;
                div += G->fx[d][5 * c + v] - G->fx[d][5 * at(G, r[0], r[1], r[2]) + v];
            }
            out[5 * c + v] = G->jac * div - 0.0;
        }
    }
    return 0;
}

int sfo_residual(sfo_grid *G, double *soa)
{
    if (sfo_apply_bc(G)) return -1;
    if (assemble_rhs(G, G->k[0])) return -1;
    const size_t N = (size_t)G->desc.n[0] * G->desc.n[1] * G->desc.n[2];
    int p[3];
    FOR_NODES(G, p) {
        size_t u = ((size_t)(p[2] - G->g[2]) * G->desc.n[1] + (p[1] - G->g[1]))
                 * G->desc.n[0] + (p[0] - G->g[0]);
        const double *r = G->k[0] + 5 * at(G, p[0], p[1], p[2]);
        for (int v = 0; v < 5; ++v) soa[v * N + u] = r[v];
    }
    return 0;
}

/* RusanovEOS::deltaT (SF_rusanovEOS.cpp:110-147) over solved nodes. */
int sfo_compute_dt(sfo_grid *G, double cfl, double *dt)
{
    double result = DBL_MAX;
    int p[3];
    FOR_NODES(G, p) {
        if (!solved(G, p)) continue;
        closed_t s;
        if (close_state(G->q + 5 * at(G, p[0], p[1], p[2]), G->desc.gamma, &s)) {
            G->err |= SFO_ERR_DT;
            return -1;
        }
        double spectral = 0.0;
        for (int d = 0; d < 3; ++d) {
            if (!G->desc.active[d]) continue;
            double m[3] = {0.0, 0.0, 0.0};
            m[d] = G->metric[d];
            double un = fabs(s.u[0] * m[0] + s.u[1] * m[1] + s.u[2] * m[2]);
            spectral += un + s.c * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        }
        if (!isfinite(spectral) || spectral <= 0.0) { G->err |= SFO_ERR_DT; return -1; }
        double local = cfl / spectral;
        if (local < result) result = local;
    }
    if (!isfinite(result) || result == DBL_MAX) { G->err |= SFO_ERR_DT; return -1; }
    *dt = result;
    return 0;
}

/* validateEquationState (SF_equationState.h:25-57) over solved nodes. */
static int validate(sfo_grid *G)
{
    int p[3];
    FOR_NODES(G, p) {
        if (!solved(G, p)) continue;
        closed_t s;
        if (close_state(G->q + 5 * at(G, p[0], p[1], p[2]), G->desc.gamma, &s)) {
            G->err |= SFO_ERR_STATE;
            return -1;
        }
    }
    return 0;
}

/* Stage combination of one value (Time::Explicit::applySSPStage /
 * applyStage / applyFinalRK4 / forwardEuler, SF_explicit.cpp:193-238,
 * with the tableaux of SF_builtinProviders.cpp:85-127).  q is the current
 * value, q0 the step-start value, k[0..2] the stored stage residuals of
 * RK4 and r the residual of this stage (dq/dt = -r). */
static const double ssp_base[3] = {0.0, 0.75, 1.0 / 3.0};
static const double ssp_incr[3] = {1.0, 0.25, 2.0 / 3.0};
static const double rk4_incr[3] = {0.5, 0.5, 1.0};
static const double rk4_w[4] = {1.0, 2.0, 2.0, 1.0};

double sfo_stage_combine(int time, int stage, double q, double q0,
                         const double k[3], double r, double dt)
{
    if (time == SFO_EULER) return q - dt * r;
    if (time == SFO_SSPRK3) {
        const double euler = q - dt * r;
        return ssp_base[stage] * q0 + ssp_incr[stage] * euler;
    }
    if (stage < 3) return q0 - rk4_incr[stage] * dt * r;
    return q0 - dt / 6.0 * (rk4_w[0] * k[0] + rk4_w[1] * k[1] + rk4_w[2] * k[2]
                            + rk4_w[3] * r);
}

/* Stage abscissae c_s (state.time + c_s * dt is the stage time). */
double sfo_stage_abscissa(int time, int stage)
{
    static const double ssp[3] = {0.0, 1.0, 0.5};
    static const double rk4[4] = {0.0, 0.5, 0.5, 1.0};
    if (time == SFO_EULER) return 0.0;
    return time == SFO_SSPRK3 ? ssp[stage] : rk4[stage];
}

/* One physical step: Time::Explicit::begin/executeStage
 * (SF_explicit.cpp:193-238,287-530) with the tableaux of
 * SF_builtinProviders.cpp:85-127; the boundary state is prepared before
 * every right-hand side and once more at commit
 * (SF_conservativeRHS.cpp:137-138, SF_flowOperations.cpp:648-652). */
int sfo_step(sfo_grid *G, double dt)
{
    G->dt = dt;
    int p[3];
    const int scheme = G->desc.time;
    if (scheme == SFO_EULER) {
        if (sfo_apply_bc(G) || assemble_rhs(G, G->k[0])) return -1;
        FOR_NODES(G, p) {
            if (!solved(G, p)) continue;
            size_t c = 5 * at(G, p[0], p[1], p[2]);
            for (int v = 0; v < 5; ++v) G->q[c + v] -= dt * G->k[0][c + v];
            double pr;
            if (legacy_pressure(G->q + c, G->desc.gamma, &pr)) {
                G->err |= SFO_ERR_STATE;
                return -1;
            }
        }
        if (validate(G)) return -1;
        return sfo_apply_bc(G);
    }
    /* begin: snapshot q0 on solved nodes */
    memset(G->q0, 0, G->cells * 5 * sizeof(double));
    FOR_NODES(G, p) {
        if (!solved(G, p)) continue;
        size_t c = 5 * at(G, p[0], p[1], p[2]);
        memcpy(G->q0 + c, G->q + c, 5 * sizeof(double));
    }
    if (scheme == SFO_SSPRK3 || scheme == SFO_RK4) {
        const int stages = scheme == SFO_SSPRK3 ? 3 : 4;
        for (int st = 0; st < stages; ++st) {
            double *r = scheme == SFO_RK4 ? G->k[st] : G->k[0];
            if (sfo_apply_bc(G) || assemble_rhs(G, r)) return -1;
            if (scheme == SFO_RK4 && validate(G)) return -1;
            FOR_NODES(G, p) {
                if (!solved(G, p)) continue;
                size_t c = 5 * at(G, p[0], p[1], p[2]);
                for (int v = 0; v < 5; ++v) {
                    const double k[3] = {G->k[0][c + v], G->k[1][c + v], G->k[2][c + v]};
                    G->q[c + v] = sfo_stage_combine(scheme, st, G->q[c + v], G->q0[c + v],
                                                    k, r[c + v], dt);
                }
            }
            if (validate(G)) return -1;
        }
        return sfo_apply_bc(G);
    }
    G->err |= SFO_ERR_ARG;
    return -1;
}

/* flow.step.prepare -> flow.dt.compute (min with maxDeltaT) -> step
 * (SF_flowOperations.cpp:471-497). */
int sfo_run(sfo_grid *G, int nsteps, double cfl, double max_dt, double *dt_seq)
{
    for (int s = 0; s < nsteps; ++s) {
        double dt;
        if (sfo_apply_bc(G) || sfo_compute_dt(G, cfl, &dt)) return -1;
        if (max_dt < dt) dt = max_dt;
        if (dt_seq) dt_seq[s] = dt;
        if (sfo_step(G, dt)) return -1;
    }
    return 0;
}

/* Time::Driver::run (src/solver/algorithm/time/SF_time.cpp:42-93) over
 * flow.step.prepare / flow.dt.compute (dt = min(CFL dt, maxDeltaT,
 * limit), SF_flowOperations.cpp:487-497) / the step.  Steps until end_time
 * (within 64 eps) or end_step; the limit is the time left to the end and,
 * for time-based output, to the next output instant.  Returns the number
 * of steps, or -1 on error. */
int sfo_run_driver(sfo_grid *G, double start, double end, int end_step,
                   int write_by_step, int write_steps, double write_time,
                   double cfl, double max_dt, double *dt_seq, int cap)
{
    double time = start;
    int step = 0;
    double next_out = start + write_time;
    int next_out_step = write_steps;
    const double eps = DBL_EPSILON;
    for (;;) {
        const double remaining = end - time;
        double scale = fabs(time) > fabs(end) ? fabs(time) : fabs(end);
        if (scale < 1.0) scale = 1.0;
        if (remaining <= 64.0 * eps * scale || (end_step > 0 && step >= end_step)) break;
        ++step;
        double limit = remaining;
        if (!write_by_step && next_out - time < limit) limit = next_out - time;
        if (!(limit > 0.0)) return -1;
        double dt;
        if (sfo_apply_bc(G) || sfo_compute_dt(G, cfl, &dt)) return -1;
        if (max_dt < dt) dt = max_dt;
        if (limit < dt) dt = limit;
        if (dt_seq && step - 1 < cap) dt_seq[step - 1] = dt;
        if (sfo_step(G, dt)) return -1;
        const double new_time = time + dt;
        double tscale = fabs(time) > fabs(new_time) ? fabs(time) : fabs(new_time);
        if (tscale < 1.0) tscale = 1.0;
        time = new_time;
        if (write_by_step && step == next_out_step) next_out_step += write_steps;
        else if (!write_by_step && time >= next_out - 1.0e-11 * tscale) next_out += write_time;
    }
    return step;
}

