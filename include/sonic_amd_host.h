/*
 * sonic_amd_host.h -- provider-level C ABI over sonic_amd.h: resolves the
 * reference's recipe names (WHICH), lowers the explicit step onto the same
 * operation ids and plan shape (HOW), and runs the reference's time-driver
 * loop (end time / end step / output instants) on a device grid.
 *
 * Citations are relative to the sonicSolver source tree.
 */
#ifndef SONIC_AMD_HOST_H
#define SONIC_AMD_HOST_H

#include "sonic_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A convection term recipe by name -> scheme, flux and halo width.
 * Names are compared after the reference's token normalisation (quotes,
 * '_', '-' and spaces removed, lower case): weno5Rusanov, teno5Rusanov
 * (resolveConvectionTermRecipe, app/application/model/parsing/
 * SF_numericsParser.h:140-162) and weno5HLLC, teno5HLLC (new).  Recipes the
 * reference registers but this provider does not serve (WENO3/WENO7,
 * Steger-Warming, Lax-Friedrichs, Roe, Lax-Wendroff, primitive upwind) fail
 * with SF_E_ARG and say so. */
int sf_recipe_resolve(const char *convection_recipe, int *scheme, int *flux,
                      int *halo_width);
/* The legacy pair form `convection.default` + `convection.flux`
 * (parseConvectionScheme / parseFluxSplitter, SF_numericsParser.h:11-94,
 * then builtInConvectionRecipe, core/config/types/SF_termRecipe.h:204-221). */
int sf_recipe_from_pair(const char *scheme_token, const char *flux_token,
                        int *scheme, int *flux, int *halo_width);
/* A time recipe token (forwardEuler, SSPRK3, classicalRK4; empty selects
 * classicalRK4) -> SF_TIME_* and its stage count (resolveTimeRecipe,
 * SF_numericsParser.h:115-125). */
int sf_time_recipe_resolve(const char *token, int *time, int *stages);

/* Text of the compiled explicit plan for a recipe pair: the op sequence of
 * explicitFragment (solver/system/SF_builtinProviders.cpp:29-82) with the
 * ConservativeResidual commit leaves (:162-175), one leaf per line. */
int sf_plan_describe(const char *convection_recipe, const char *time_recipe,
                     char *buf, size_t len);

/* Time::Driver run control (core/config/SF_runControl.h:9-16). */
typedef struct {
    double start_time;
    double end_time;
    int end_step;               /* 0: no step limit                         */
    int write_by_step;          /* 1: output every write_interval_steps     */
    int write_interval_steps;
    double write_interval_time; /* runTime output interval                  */
} sf_run_control;

/* Called at each output instant (after the step that reaches it); a
 * nonzero return stops the run. */
typedef int (*sf_output_fn)(void *user, int step, double time, double dt);

typedef struct sf_solver sf_solver;

/* A device grid plus the plan compiled for the given recipes, with the
 * operations bound as FlowOperations::bindExplicitOps binds them for the
 * "flow.conservative" provider (solver/algorithm/SF_flowOperations.cpp:
 * 466-522,648).  The scheme/flux/time fields of `geometry` are ignored. */
int sf_solver_create(const sf_grid_desc *geometry, const char *convection_recipe,
                     const char *time_recipe, double cfl, double max_dt,
                     sf_solver **out);
int sf_solver_destroy(sf_solver *solver);
/* The solver's grid (for state upload / download); owned by the solver. */
int sf_solver_grid(sf_solver *solver, sf_grid **grid);
/* Time::Driver::run (solver/algorithm/time/SF_time.cpp:42-93): step until
 * end_time (within 64 eps) or end_step, limiting dt to the time left to the
 * end and to the next output instant. */
int sf_solver_run(sf_solver *solver, const sf_run_control *control,
                  sf_output_fn on_output, void *user, int *steps_done,
                  double *final_time);
/* The dt of every step the solver has taken, oldest first. */
int sf_solver_dt_history(sf_solver *solver, double *out, int capacity, int *count);

#ifdef __cplusplus
}
#endif
#endif
