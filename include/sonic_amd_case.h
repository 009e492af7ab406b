/*
 * sonic_amd_case.h -- a SonicFile case directory (the reference's case.yaml
 * tree) read and lowered onto the explicit density-based path: mesh from
 * mesh/generation.yaml, initial fields and boundary sets from fields/,
 * recipes and run control from solvers/, the equation of state from
 * models/.  Stands in for SF::CaseIO::read
 * (infrastructure/io/case/SF_case.cpp:251-290) followed by the decode of
 * CaseAdapter (app/application/model/compatibility/SF_nativeDecode.cpp:
 * 906-1072, 1323-1556) for the subset of cases this path serves: one
 * structured block with uniform spacing, perfect gas, fields rho / U / p
 * with fixed-value initial sets, zeroGradient or empty boundary sets, an
 * explicit time recipe and a characteristic WENO5 / TENO5 convection recipe.
 * Anything else is rejected with a message that names the file and entry.
 */
#ifndef SONIC_AMD_CASE_H
#define SONIC_AMD_CASE_H

#include "sonic_amd.h"
#include "sonic_amd_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_case sf_case;

/* Read and validate a case directory. */
int sf_case_open(const char *case_dir, sf_case **out);
int sf_case_close(sf_case *c);
/* The lowered grid: active axes, transmissive ends, spacing, gamma and the
 * scheme / flux / time of the recipes. */
int sf_case_grid(const sf_case *c, sf_grid_desc *desc);
/* Initial conserved state on the unique nodes of the lowered grid
 * (soa[v*N + (k*ny + j)*nx + i]); N from sf_case_grid. */
int sf_case_state(const sf_case *c, double *soa);
/* Recipe names as the case spells them (after resolution), CFL and
 * maxDeltaT. */
int sf_case_numerics(const sf_case *c, char *convection, size_t convection_len, char *time,
                     size_t time_len, double *cfl, double *max_dt);
/* startTime / endTime / endStep / writeControl / writeInterval. */
int sf_case_run_control(const sf_case *c, sf_run_control *control);
/* A text summary of everything the reader decided. */
int sf_case_describe(const sf_case *c, char *buf, size_t len);
/* Run the case on the device with Time::Driver
 * (solver/algorithm/time/SF_time.cpp:42-93); when output_dir is not NULL,
 * write a .vts file at every output instant (and at the start when
 * writeInitial is set) plus <jobName>.pvd there, named and laid out as the
 * reference's ResultWriter does (<jobName>_t<time> for runTime output,
 * <jobName>_<step> for timeStep output).  final_soa (N x 5, may be NULL)
 * receives the final state; steps / end_time may be NULL. */
int sf_case_run(sf_case *c, const char *output_dir, double *final_soa, int *steps,
                double *end_time);
/* Write `count` states (states[w] = N x 5 SoA as sf_case_state lays it out,
 * stored one after another) as the .vts files of steps[w] / times[w] plus
 * the <jobName>.pvd index, exactly as sf_case_run writes its output
 * instants; host memory only, no device work.  Stands in for
 * ResultWriter::save (infrastructure/io/SF_vtkOutput.cpp:59-67,113-125). */
int sf_case_write_vts(const sf_case *c, const char *output_dir, int count, const int *steps,
                      const double *times, const double *states);
/* The block: node layers per axis (inactive axes included), the scaled
 * hex vertices (vertex[3 v + a]) and the gas constant. */
int sf_case_mesh(const sf_case *c, int nodes[3], double vertex[24], double *gas_constant);
/* The boundary sets the settings of field "rho", "U" or "p" name,
 * comma-separated in the order the output numbers them (BC_ID). */
int sf_case_boundary_sets(const sf_case *c, const char *field, char *buf, size_t len);
/* Membership of every block node (i fastest) in the named node set
 * (MeshGen's set definitions, SF_meshGen.cpp:514-541). */
int sf_case_set_mask(const sf_case *c, const char *set, uint8_t *mask);

#ifdef __cplusplus
}
#endif
#endif
