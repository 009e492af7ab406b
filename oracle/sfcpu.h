/*
 * sfcpu.h -- CPU FP64 restatement of sonicSolver's density-based
 * single-fluid explicit hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * path is compared against; nothing in the shipped package may link,
 * load or call it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * Every function restates one reference function; the citation is given
 * next to the definition in sfcpu.c (paths relative to the sonicSolver
 * source tree).  The storage keeps the reference's traits: conserved
 * variables interleaved per node ([k][j][i][v], SF_field.h:58-67), a face
 * flux store per direction, then residual, stage update and validation as
 * separate passes.
 *
 * Two pieces have no counterpart in the reference and are pinned by
 * mathematics instead of by reference output: the HLLC flux (Toro,
 * "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
 * ch. 10) and periodic ghost fill.
 */
#ifndef SONIC_ORACLE_SFCPU_H
#define SONIC_ORACLE_SFCPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SFO_WENO5 = 0, SFO_TENO5 = 1, SFO_WENO3 = 2, SFO_WENO7 = 3 };
enum { SFO_RUSANOV = 0, SFO_HLLC = 1, SFO_STEGER = 2, SFO_LAXF = 3, SFO_ROE = 4,
       SFO_LAXW = 5 };
enum { SFO_EULER = 0, SFO_SSPRK3 = 1, SFO_RK4 = 2 };
/* per axis: zeroGradient on both ends, periodic, or an end that carries
 * no condition (the case declares none for that set): its node is advanced
 * like an interior node and its ghost layers keep the values they were
 * given with the state (copies of the end node), never refilled. */
enum { SFO_TRANSMISSIVE = 0, SFO_PERIODIC = 1, SFO_NONE_LO = 2, SFO_NONE_HI = 3,
       SFO_NONE = 4 };

/* error bits (the reference terminates the process on each of these) */
enum {
    SFO_ERR_STATE = 1u,   /* non-physical state after a stage            */
    SFO_ERR_CHAR = 2u,    /* non-physical face-average state             */
    SFO_ERR_TENO = 4u,    /* TENO5 rejected every substencil             */
    SFO_ERR_FLUX = 8u,    /* non-finite face flux / wave speed           */
    SFO_ERR_DT = 16u,     /* invalid spectral radius in the CFL scan     */
    SFO_ERR_WENO = 32u,   /* non-finite stencil or weight sum            */
    SFO_ERR_ARG = 64u     /* invalid argument                            */
};

/* ---- point kernels ---------------------------------------------------- */
double sfo_weno5_core(const double v[5], unsigned *err);
double sfo_teno5_core(const double v[5], unsigned *err);
double sfo_weno3_core(const double v[4], unsigned *err);
double sfo_weno7_core(const double v[7], unsigned *err);
/* the WENO7 smoothness indicator of one four-node candidate over [a, b] */
double sfo_weno7_smoothness(const double v[4], double a, double b);
/* stencil nodes of a scheme (4, 6 or 8) and its ghost depth (2, 3, 4) */
int sfo_stencil_size(int scheme);
int sfo_halo(int scheme);
int sfo_char_matrix(const double q[5], const double n[3], double gamma,
                    double L[25], double R[25]);
void sfo_gemv5(const double M[25], const double v[5], double out[5]);
int sfo_rusanov(const double qL[5], const double qR[5], const double n[3],
                double gamma, double out[5]);
int sfo_hllc(const double qL[5], const double qR[5], const double n[3],
             double gamma, double out[5]);
int sfo_steger_warming(const double qL[5], const double qR[5], const double n[3],
                       double gamma, double out[5]);
int sfo_lax_friedrichs(const double qL[5], const double qR[5], const double n[3],
                       double gamma, double out[5]);
int sfo_roe(const double qL[5], const double qR[5], const double n[3],
            double gamma, double out[5]);
/* Lax-Wendroff: needs the step's dt and the face's node spacing */
int sfo_lax_wendroff(const double qL[5], const double qR[5], const double n[3],
                     double gamma, double dt, double spacing, double out[5]);
/* any of the five face fluxes by id */
int sfo_face_flux(int flux, const double qL[5], const double qR[5], const double n[3],
                  double gamma, double out[5]);
int sfo_physical_flux(const double q[5], const double n[3], double gamma,
                      double out[5]);
/* stencil is 6 nodes x 5 variables (offsets -2..+3 around face i+1/2);
 * cofactor is the face cofactor vector; out is the area-weighted flux. */
int sfo_char_face_flux(const double stencil[30], const double cofactor[3],
                       int scheme, int flux, double gamma, double out[5]);
/* the same for any scheme: sfo_stencil_size(scheme) nodes x 5 variables at
 * offsets 1-ns/2 .. ns/2 around face i+1/2 */
int sfo_char_face_flux_n(const double *stencil, const double cofactor[3],
                         int scheme, int flux, double gamma, double out[5]);
/* the same with the step's dt and the face's node spacing (Lax-Wendroff) */
int sfo_char_face_flux_dt(const double *stencil, const double cofactor[3],
                          int scheme, int flux, double gamma, double dt,
                          double spacing, double out[5]);

/* stage combination of one value, dq/dt = -r (see sfcpu.c) */
double sfo_stage_combine(int time, int stage, double q, double q0,
                         const double k[3], double r, double dt);
double sfo_stage_abscissa(int time, int stage);

/* ---- grid ------------------------------------------------------------- */
typedef struct sfo_grid sfo_grid;

typedef struct {
    int n[3];        /* unique nodes per axis (1 on an inactive axis)       */
    int active[3];   /* 1 = axis takes part in the discretisation           */
    int bc[3];       /* SFO_TRANSMISSIVE, SFO_PERIODIC or SFO_NONE_* per axis */
    double d[3];     /* node spacing per axis                               */
    double gamma;
    int scheme, flux, time;
} sfo_desc;

sfo_grid *sfo_grid_create(const sfo_desc *desc);
void sfo_grid_destroy(sfo_grid *g);
/* state in/out as variable-major arrays over the unique nodes:
 * soa[v*N + (k*ny + j)*nx + i], N = nx*ny*nz */
int sfo_set_state(sfo_grid *g, const double *soa);
int sfo_get_state(const sfo_grid *g, double *soa);
int sfo_apply_bc(sfo_grid *g);
/* unscaled-by-dt residual R = Jac * sum_d (F_d(i) - F_d(i-1)) on the
 * unique nodes, variable-major; boundary nodes of transmissive axes get 0 */
int sfo_residual(sfo_grid *g, double *soa_out);
int sfo_compute_dt(sfo_grid *g, double cfl, double *dt);
int sfo_step(sfo_grid *g, double dt);
/* run nsteps with dt = min(cfl dt, max_dt); dt_seq may be NULL */
int sfo_run(sfo_grid *g, int nsteps, double cfl, double max_dt,
            double *dt_seq);
unsigned sfo_error(const sfo_grid *g);
/* Time::Driver::run semantics (end time, end step, output instants) */
int sfo_run_driver(sfo_grid *g, double start, double end, int end_step,
                   int write_by_step, int write_steps, double write_time,
                   double cfl, double max_dt, double *dt_seq, int cap);

#ifdef __cplusplus
}
#endif
#endif
