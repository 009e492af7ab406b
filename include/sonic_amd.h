/*
 * sonic_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * sonicSolver's density-based single-fluid explicit hot path:
 * characteristic WENO5-JS / TENO5 face reconstruction, Rusanov / HLLC face
 * fluxes, the ConservativeResidual flux divergence, the Euler / SSPRK3 /
 * RK4 stage updates, on-device ghost fill and the CFL time-step reduction.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative SF_E_* code on
 *    failure; sf_last_error() returns the message of the last failure on
 *    the calling thread.
 *  - State crosses the boundary as plain FP64 arrays, variable-major over
 *    the unique nodes: soa[v*N + (k*ny + j)*nx + i], N = nx*ny*nz,
 *    v = {rho, rhoU, rhoV, rhoW, rhoE}.  Host arrays are owned by the
 *    caller; device memory is owned by the library.
 *  - A handle is used from one host thread at a time; all its work is
 *    queued on one HIP stream owned by the handle.
 *
 * Citations name the reference interface each entry point stands in for
 * (paths relative to the sonicSolver source tree).
 */
#ifndef SONIC_AMD_H
#define SONIC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enumerations ------------------------------------------------------ */
/* FDM::ConvectionScheme, core/config/types/SF_termRecipe.h:14 */
enum { SF_SCHEME_WENO5 = 0, SF_SCHEME_TENO5 = 1, SF_SCHEME_WENO3 = 2, SF_SCHEME_WENO7 = 3 };
/* FDM::FluxSplitter, SF_termRecipe.h:34-41; HLLC has no reference member */
enum { SF_FLUX_RUSANOV = 0, SF_FLUX_HLLC = 1, SF_FLUX_STEGER_WARMING = 2,
       SF_FLUX_LAX_FRIEDRICHS = 3, SF_FLUX_ROE = 4, SF_FLUX_LAX_WENDROFF = 5 };
/* FDM::TimeRecipeId, core/config/types/SF_timeRecipe.h:10 */
enum { SF_TIME_EULER = 0, SF_TIME_SSPRK3 = 1, SF_TIME_RK4 = 2 };
/* Boundary kind of an axis: zeroGradient (transmissive) sets on both
 * ends; periodic (no reference counterpart: BCType has no periodic kind);
 * or an end whose set carries no density / velocity / pressure condition
 * (lower, upper or both; the other end zeroGradient).  Such an end is not
 * a solver-boundary set (SF_applicator.h:162-187): its node is advanced
 * and enters the CFL scan like an interior node, and its ghost layers keep
 * the values sf_state_upload gives them (copies of the end node), since no
 * boundary law refills them. */
enum { SF_BC_TRANSMISSIVE = 0, SF_BC_PERIODIC = 1, SF_BC_NONE_LO = 2, SF_BC_NONE_HI = 3,
       SF_BC_NONE = 4 };

/* error codes */
enum {
    SF_OK = 0,
    SF_E_ARG = -1,        /* invalid argument                               */
    SF_E_HIP = -2,        /* HIP runtime failure                            */
    SF_E_NUMERIC = -3,    /* device raised a numerical fault (see flags)    */
    SF_E_MEMORY = -4,     /* device allocation failed                       */
    SF_E_STATE = -5       /* call out of order                              */
};

/* device error flags (the reference terminates the process on each) */
enum {
    SF_FLAG_STATE = 1u,   /* non-physical state after a stage:
                             validateEquationState, SF_equationState.h:25-57 */
    SF_FLAG_CHAR = 2u,    /* non-physical face-average state:
                             buildCharacteristicMatrix, SF_riemann.cpp:72-74 */
    SF_FLAG_TENO = 4u,    /* TENO5 rejected all substencils:
                             SF_TENO5.cpp:79-89                              */
    SF_FLAG_FLUX = 8u,    /* non-finite face flux: storeFlux,
                             SF_faceAssembly.h:26-40                         */
    SF_FLAG_DT = 16u      /* invalid spectral radius: RusanovEOS::deltaT,
                             SF_rusanovEOS.cpp:141-146                       */
};

/* Grid description.  Replaces the Field set-up that the reference does
 * through Field::setup + Mesh::computeMetrics + the boundary applicator
 * (core/field/SF_field.h:89-96, infrastructure/mesh/SF_mesh.cpp:219-346,
 * solver/boundary/SF_applicator.h:51-83) for a uniform Cartesian grid. */
typedef struct {
    int n[3];        /* unique nodes per axis; 1 on an inactive axis          */
    int active[3];   /* 1 = axis is discretised (Math::isDirectionActive)     */
    int bc[3];       /* SF_BC_* for both ends of each active axis             */
    double d[3];     /* node spacing per axis (1.0 on an inactive axis)       */
    double gamma;    /* perfect-gas ratio of specific heats                   */
    int scheme;      /* SF_SCHEME_*                                           */
    int flux;        /* SF_FLUX_*                                             */
    int time;        /* SF_TIME_*                                             */
    int device;      /* HIP device ordinal                                    */
} sf_grid_desc;

typedef struct sf_grid sf_grid;

/* ---- library ----------------------------------------------------------- */
int sf_last_error(char *buf, size_t len);
int sf_version(int *major, int *minor);

/* ---- grid lifetime ----------------------------------------------------- */
/* Allocates device-resident state (three state buffers, a residual
 * accumulator and, for RK4, three stage buffers).  Replaces Field::setup
 * (SF_field.h:94) and PatchWorkspace::ensureFor
 * (solver/algorithm/SF_patchWorkspace.h:32-44). */
int sf_grid_create(const sf_grid_desc *desc, sf_grid **out);
int sf_grid_destroy(sf_grid *grid);
/* Device bytes held by the grid. */
int sf_grid_device_bytes(const sf_grid *grid, size_t *bytes);

/* ---- state transfer ------------------------------------------------------ */
/* Host <-> device copies of the conserved state on the unique nodes
 * (the reference reads and writes Field::operator(), SF_field.h:271-272). */
int sf_state_upload(sf_grid *grid, const double *soa);
int sf_state_download(sf_grid *grid, double *soa);
/* Same, from / to device memory (a pointer valid on the grid's device). */
int sf_state_upload_device(sf_grid *grid, const double *soa_device);
int sf_state_download_device(sf_grid *grid, double *soa_device);

/* ---- the explicit step, operation by operation ---------------------------
 * These stand in for the op ids that FlowOperations::bindExplicitOps binds
 * for the "flow.conservative" runtime provider
 * (solver/algorithm/SF_flowOperations.cpp:466-522,648-652;
 *  core/system/SF_operationIds.h:29-40). */

/* flow.step.prepare: boundary state at the start of a step
 * (prepareBoundaryState -> Boundary::Applicator::apply). */
int sf_step_prepare(sf_grid *grid);
/* flow.dt.compute: dt = min(CFL scan, max_dt, limit)
 * (RusanovEOS::deltaT, SF_rusanovEOS.cpp:110-147, then
 *  SF_flowOperations.cpp:487-497).  The value stays on the device for the
 * following stages; if dt_out is not NULL it is also copied to the host. */
int sf_compute_dt(sf_grid *grid, double cfl, double max_dt, double limit,
                  double *dt_out);
/* Use a caller-chosen dt for the next step instead of the CFL scan. */
int sf_set_dt(sf_grid *grid, double dt);
/* flow.step.begin: Time::Explicit::begin (SF_explicit.cpp:287-321). */
int sf_step_begin(sf_grid *grid);
/* explicit.stage.execute: Time::Explicit::executeStage
 * (SF_explicit.cpp:331-530) with the RHS assembly of
 * ConservativeRHS::assembleAllPatches (SF_conservativeRHS.cpp:106-247). */
int sf_stage(sf_grid *grid, int stage);
/* flow.step.commit + time.commit: boundary state at t+dt, clock advance
 * (SF_flowOperations.cpp:648-652, SF_singleFluidStepper.cpp:40-46). */
int sf_step_commit(sf_grid *grid);
/* Number of stages of the grid's time recipe (1, 3 or 4). */
int sf_stage_count(const sf_grid *grid, int *count);

/* ---- whole steps --------------------------------------------------------- */
/* nsteps physical steps with dt = min(CFL dt, max_dt) each, all queued on
 * the device without host synchronisation between steps.  If dt_seq is not
 * NULL it receives the nsteps dt values.  Checks the device error flags at
 * the end and returns SF_E_NUMERIC if any was raised, as sf_fault_check
 * does: the kernels queued after the fault did nothing, the grid holds the
 * state at the start of the faulting step and its clock counts the steps
 * before it (dt_seq entries from the faulting step on are not meaningful).
 * The reference stops at the first non-physical stage
 * (validateEquationState, SF_equationState.h:25-57; TENO5,
 * SF_TENO5.cpp:79-89). */
int sf_run(sf_grid *grid, int nsteps, double cfl, double max_dt,
           double *dt_seq);
/* Physical time and step count of the grid's clock. */
int sf_clock(sf_grid *grid, double *time, long long *steps);

/* ---- a whole run that allocates nothing ------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  sf_grid_create + sf_state_upload_device + sf_run +
 * sf_state_download_device + sf_grid_destroy in one call, on the same
 * kernels, with no hipMalloc, hipFree or other allocation: the grid's state
 * buffers, flux accumulator, RK4 stage residuals, clock and fault record
 * live in one file-scope device array of 64 MiB per GPU (held from the
 * library's first use of that GPU until the process ends).  A grid needing
 * more (its *scratch_bytes, if not NULL, is set either way) is rejected
 * with SF_E_ARG before any work.  soa_device (unique-node SoA, device
 * memory, 5 N doubles) holds the initial state and receives the result;
 * dt_seq_device (device memory, nsteps doubles, may be NULL) receives the
 * dt values; the work is queued on `stream` (a hipStream_t, NULL for the
 * null stream) and waited for before returning.  Results equal sf_run's bit
 * for bit, including a fault: SF_E_NUMERIC, with the state at the start of
 * the faulting step in soa_device.  Every call on one GPU uses that GPU's
 * one array (each call clears what it needs), so calls that overlap on one
 * GPU race: call from one host thread per GPU, or serialise the calls. */
int sf_run_static(const sf_grid_desc *desc, double *soa_device, int nsteps, double cfl,
                  double max_dt, double *dt_seq_device, void *stream, size_t *scratch_bytes);

/* ---- diagnostics ---------------------------------------------------------- */
/* Residual R = Jac * sum_d (F_d(i) - F_d(i-1)) of the current state after
 * boundary preparation, on the unique nodes (0 on solver-boundary nodes);
 * Math::localSpatialResidual, SF_residualAssembly.h:12-32. */
int sf_residual(sf_grid *grid, double *soa_out);
/* Device error flags (SF_FLAG_*); clear != 0 resets them (and the record
 * of where they were raised).  sf_state_upload* also resets them. */
int sf_error_flags(sf_grid *grid, uint32_t *flags, int clear);
/* Waits for the grid's queued work and checks the device error flags, as
 * the reference's fail-fast checks do after every stage
 * (SF_equationState.h:25-57, SF_TENO5.cpp:79-89, SF_rusanovEOS.cpp:141-146).
 * Once a flag is raised every later stage, fill and clock kernel of the
 * grid returns at entry, so the buffer holding the state at the start of
 * the faulting step is untouched; this call makes it the current state
 * again.  Returns 0 with *flags = 0 when no flag is raised, else
 * SF_E_NUMERIC with the flags, the 1-based number of the faulting step on
 * the grid's clock and its phase (0..3 = stage 1..4, 6 = ghost fill at
 * the start or commit of the step, 7 = the CFL scan before it).  Any
 * output may be NULL. */
int sf_fault_check(sf_grid *grid, uint32_t *flags, long long *step, int *phase);
/* Nodes per segment of the y / z marching sweeps (1..4096), or 0 for the
 * automatic choice (64 where that leaves at least 8192 one-wave blocks,
 * else 32 halved down to 8 until there are 4096).  No reference
 * counterpart; results do not depend on it. */
int sf_grid_set_march_segment(sf_grid *grid, int nodes);
/* How the x sweep walks the grid: 1 by whole x lines in 64-node tiles,
 * 0 by tiles of 63 faces over all lines numbered as one sequence, -1 (the
 * default) by lines where nx >= 64 and the grid has at least 256 x lines
 * per compute unit (65536 lines on a 256-CU device: 256^3 and up), else by
 * flattened tiles.  Grids whose
 * variables exceed 2^32 - 1024 bytes always take the flattened tiles.  No
 * reference counterpart; results do not depend on it. */
int sf_grid_set_x_walk(sf_grid *grid, int mode);
/* Wait for the grid's queued work. */
int sf_synchronize(sf_grid *grid);
/* The HIP stream (hipStream_t) the grid queues its work on. */
int sf_grid_stream(sf_grid *grid, void **stream);

/* ---- point evaluations of the face arithmetic ------------------------------
 * The device functions the sweeps use, applied to caller-given data, for
 * checking them against a CPU restatement.  Arrays are host memory,
 * array-of-structures: state i at [5 i .. 5 i + 4]; flags (may be NULL)
 * receives each point's SF_FLAG_* bits.  Blocking. */
/* The face flux of `flux` (SF_FLUX_*) between qL[i] and qR[i] on a face
 * whose unit normal is e_axis: the FaceFlux interface of the reference's
 * recipes (RusanovEOS::faceFlux, SF_rusanovEOS.cpp:47-77; StegerWarming,
 * LaxFriedrichs, Roe and LaxWendroff fluxes in methods/numerics/Flux;
 * HLLC has no reference counterpart).  lw = dt / spacing (Lax-Wendroff
 * only). */
int sf_riemann_flux(int device, int flux, int axis, double gamma, double lw, int count,
                    const double *qL, const double *qR, double *out, uint32_t *flags);
/* Flux::characteristicWENOFaceFlux<6> (SF_WENO.h:185-245) of the six-node
 * stencils[i] (nodes at offsets -2..3 around the face, [(6 i + s) 5 + v])
 * on a face of unit normal e_axis and the given area; WENO5 / TENO5
 * (SF_SCHEME_*) with the Rusanov or HLLC flux. */
int sf_face_flux(int device, int scheme, int flux, int axis, double gamma, double area,
                 int count, const double *stencils, double *out, uint32_t *flags);

/* ---- kernel timing -----------------------------------------------------------
 * No reference counterpart (the reference has no timers).  When enabled,
 * every launch of the kernels below is bracketed by HIP events recorded on
 * the grid's stream; sf_profile_read waits for them and returns the summed
 * duration and the launch count.  Enabling resets the counters. */
enum {
    SF_KERNEL_SWEEP_X = 0,   /* x-face flux sweep                               */
    SF_KERNEL_SWEEP_Y = 1,   /* y-face flux sweep                               */
    SF_KERNEL_SWEEP_Z = 2,   /* z-face flux sweep                               */
    SF_KERNEL_GHOST = 3,     /* one ghost fill (all its boundary kernels)       */
    SF_KERNEL_CFL = 4,       /* standalone CFL scan                             */
    SF_KERNEL_COUNT = 5
};
int sf_profile_enable(sf_grid *grid, int enable);
int sf_profile_read(sf_grid *grid, int kernel, double *total_ms, long long *launches);

/* ---- batched solves ----------------------------------------------------------
 * No reference counterpart: the reference advances one case per process
 * (app/application/SF_application.cpp:111-118 runs one case).  A batch
 * holds many independent 1D or 2D problems of uneven size on one device:
 * each has its own scheme, flux, CFL number, dt and clock, transmissive
 * boundaries on every side (the reference's zeroGradient sets,
 * SF_zeroGradient.cpp:88-130) and SSPRK3 time stepping
 * (SF_builtinProviders.cpp:98-121).  One step of the batch advances every
 * live problem by one SSPRK3 step with its own CFL dt; problems never
 * interact, so each result is the same whichever problems share the batch.
 * State crosses the boundary per problem as soa[v*N + j*nx + i],
 * N = nx*ny. */
typedef struct {
    int n[2];        /* unique nodes along x and y; n[1] = 1 for a 1D problem */
    double d[2];     /* node spacing (d[1] ignored for a 1D problem)           */
    int scheme;      /* SF_SCHEME_*                                            */
    int flux;        /* SF_FLUX_*                                              */
    double cfl;      /* CFL number of this problem's dt                        */
} sf_batch_item;

typedef struct sf_batch sf_batch;

/* Padded device nodes a problem occupies in the arena (per variable); the
 * arena holds 4 values of 5 variables per node (three state buffers and the
 * flux accumulator), 160 bytes. */
int sf_batch_item_nodes(const sf_batch_item *item, size_t *nodes);
/* A batch with an arena of node_capacity padded nodes and up to
 * item_capacity problems. */
int sf_batch_create(int device, double gamma, size_t node_capacity, int item_capacity,
                    sf_batch **out);
int sf_batch_destroy(sf_batch *batch);
/* Admit a problem: first-fit placement of its footprint in the arena
 * (SF_E_MEMORY when it does not fit).  max_steps > 0 freezes the problem
 * after that many steps; 0 leaves it running until removed.  The slot
 * numbers the problem until sf_batch_remove. */
int sf_batch_add(sf_batch *batch, const sf_batch_item *item, const double *soa,
                 long long max_steps, int *slot);
int sf_batch_remove(sf_batch *batch, int slot);
int sf_batch_download(sf_batch *batch, int slot, double *soa);
/* Advance every live problem by nsteps steps.  use_graph 0 launches each
 * step's kernels directly; 1 replays one captured hipGraph per step
 * (captured again after admissions/removals); 2 replays the graph when one
 * is current or nsteps is long enough to pay for capturing it, and
 * launches directly otherwise.  Results do not depend on it. */
int sf_batch_step(sf_batch *batch, int nsteps, int use_graph);
/* Physical time and step count of one problem. */
int sf_batch_clock(sf_batch *batch, int slot, double *time, long long *steps);
/* Step counts of all slots (item_capacity entries; -1 for a free slot). */
int sf_batch_steps(sf_batch *batch, long long *steps);
/* Arena occupancy in padded nodes, and the number of admitted problems. */
int sf_batch_memory(const sf_batch *batch, size_t *used_nodes, size_t *free_nodes,
                    int *items);
/* Tile shape of the persistent sweeps: x tiles of xlen nodes of one line
 * (63, 127 or 255), y tiles of 64 columns by ylen nodes (4..4096).  Takes
 * effect at the next step; results do not depend on it. */
int sf_batch_set_tiles(sf_batch *batch, int xlen, int ylen);
/* Persistent grid size of the sweeps in workgroups per compute unit
 * (1..64), or 0 for as many as each sweep kernel keeps resident (the
 * runtime's occupancy figure).  Results do not depend on it. */
int sf_batch_set_occupancy(sf_batch *batch, int blocks_per_cu);
/* Sweep schedule: fused != 0 runs one launch per sweep for every
 * (scheme, flux) variant (workgroup ranges assigned per variant); 0 runs one
 * launch per variant present (fewer registers per kernel).  Results do not
 * depend on it. */
int sf_batch_set_schedule(sf_batch *batch, int fused);
/* Per-problem device fault flags (SF_FLAG_*), item_capacity entries (0 for
 * free slots).  A problem that raises a flag during a step finishes that
 * step's stages on its own nodes, is not counted as having taken the step
 * and is frozen from the next step on; the others keep advancing, and
 * sf_batch_step returns SF_E_NUMERIC. */
int sf_batch_faults(sf_batch *batch, uint32_t *flags);

/* Launch shape chosen by sf_batch_autotune. */
typedef struct {
    int xlen, ylen, blocks_per_cu, fused;
    double ms_per_step;      /* mean step time of the chosen shape            */
    int candidates;          /* shapes timed                                  */
} sf_batch_tuning;
/* Time every launch shape (schedule, x tile, y tile, persistent grid size)
 * on the admitted problems for trial_steps steps each, from a saved state
 * that is restored after every trial, and keep the fastest.  The problems'
 * states and clocks are unchanged on return. */
int sf_batch_autotune(sf_batch *batch, int trial_steps, sf_batch_tuning *out);

#ifdef __cplusplus
}
#endif
#endif
