/*
 * ramses2_hip_ext.h -- extensions of the gfx950 hydro library beyond the
 * drop-in ABI of ramses2_hip.h.  None of these has a counterpart in the
 * reference's Metal or CUDA bridges; they serve callers that drive whole level
 * steps on the device (benchmarks, the batched AMR path, ensemble serving).
 */
#ifndef RAMSES2_HIP_EXT_H
#define RAMSES2_HIP_EXT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision of the Godunov kernel used by rh_godunov and rh_level_step:
 * 0 = FP64 throughout; 1 = FP32 reconstruction and Riemann fluxes, FP64
 * primitive recovery p=(gamma-1)(E-rho v^2/2), FP64 accumulation of the
 * conserved variables, FP64 dt reduction and conservation sums. */
int rh_set_precision(int precision);

/* Courant pass of the next step inside the Godunov kernel (rh_level_run, 3D):
 * -1 automatic (on: it measures faster in both precisions), 0 a separate
 * Courant kernel before each step, 1 the Godunov kernel reduces the Courant
 * condition of the state it writes (the FP64 kernel from each cell's new state
 * as it stores it, the FP32 kernel in an epilogue that re-reads its cells). */
int rh_set_fused_courant(int mode);

/* Artificial diffusion coefficient (hydro_params_t%difmag, amr/amr_commons.f90:
 * 690-702; the Metal ABI has no such argument): difmag > 0 adds the diffusive
 * fluxes of cmpdivu/cmpdiff (hydro/umuscl.f90:2392-2540) to every 3D level
 * step of rh_godunov and rh_level_run (with a separate Courant pass); the
 * batched AMR and serving steps refuse it.  Default 0. */
int rh_set_difmag(double difmag);

/* Deterministic coarse-fine refluxes: 1 = each face between a level cell and
 * an interpolated ghost cell stores its flux in a slot of its own (24 per
 * ghost oct), and a fold adds the slots to the parent cell in a fixed order
 * after the level's launch, so that repeated runs give the same bits (the
 * reference adds them serially, hydro/godunov_fine.f90:894-1052); 0 (the
 * default) = FP64 atomics, whose order varies from run to run.  Applies to
 * rh_godunov and to the next rh_amr_prepare. */
int rh_set_deterministic(int on);

/* 1 (the default): levels that are periodic in every dimension, with no
 * refined cell and no coarse-fine faces, run a build of the Godunov brick body
 * without the boundary-ghost gather, refined-face zeroing and reflux code;
 * 0: every level runs the general body (the results are the same bits). */
int rh_set_specialized(int on);

/* Oct-brick shape (octs per brick in x, y, z) used by the next rh_build_nbor
 * for 3D levels; the autotuner (ramses2_amd/autotune.py) picks it. */
int rh_set_brick_shape(int bx, int by, int bz);
int rh_get_brick_shape(int32_t *bx, int32_t *by, int32_t *bz);

/* Copy one of a level's device index tables to the host (for checking the
 * indexing against a brute-force key lookup).  which: 0 neighbour table
 * [noct][3^ndim] (oct index, -(reference oct) for a boundary ghost, 0 none);
 * 1 father [noct][2] (parent oct, cell 1..2^ndim); 2 brick origins
 * [nbricks][3]; 3 brick windows [nbricks][(bx+2)(by+2)(bz+2)]; 4 ghost octs
 * of the level [nghost][4] (oct index, key x, y, z); 5 their parent cells
 * [nghost][2].  *len receives the number of int32 values of the table; the
 * copy happens when capacity >= *len (else RH_ERR_ARG). */
int rh_get_level_table(int ilevel, int which, int32_t *out, int64_t capacity, int64_t *len);

/* The brick lattice rh_build_nbor gives a 3D level whose octs have these keys
 * ([noct][3], each >= box_min): bricks of bx*by*bz octs with origins at
 * box_min + off + multiples of the shape, off chosen per level to minimise
 * the modelled march work, returned in *work (per brick: 0.35(2bx+4)(2by+4)
 * P cells and (2bx+2)(2by+2) columns per marched plane).  Host only: no
 * device call; the autotuner ranks brick shapes with it. */
int rh_brick_lattice(const int32_t *ckey, int noct, const int32_t box_min[3], const int32_t box_max[3], int bx,
                     int by, int bz, int32_t off[3], double *work);

/* HIP-event timer on the library stream: milliseconds between start and stop. */
int rh_timer_start(void);
int rh_timer_stop(float *ms);

/* nsteps full time steps of one level that is the whole mesh (uniform grid):
 * courant_fine (dt reduced on the device) -> godunov_fine with set_unew and
 * set_uold folded in (unew = uold + dF, then the two arrays swap roles).
 * dt_fixed > 0 imposes dt instead of the Courant value.  dt_last, t_total and
 * the conservation sums (mass, total energy, internal energy, 0) of the last
 * Courant pass are returned when non-NULL.  With the Courant pass in the
 * Godunov kernel (rh_set_fused_courant, the default) that pass is over the
 * state after the last step, and dt_last is the dt the next step would take;
 * with a separate Courant kernel it is over the state before the last step,
 * and dt_last is the dt that step took.
 * use_graph != 0 replays the step pair as a captured hipGraph; the graph is
 * kept and replayed again by the next call with the same arguments on the
 * same arrays and tables (rh_build_nbor, rh_alloc_amr and rh_finalize drop
 * it).  kernel_ms, when non-NULL (and use_graph == 0), receives the mean
 * duration of the Godunov launches, timed with HIP events on the library
 * stream inside the steps. */
int rh_level_run(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr,
                 double smallc2, double courant_factor, const double g[3], int slope, int riemann,
                 double dt_fixed, int nsteps, int use_graph, double *dt_last, double *t_total,
                 double sums[4], float *kernel_ms);

/* A plain device copy of nbytes (multiple of 32) with a vector-load kernel,
 * timed with HIP events: the calibration kernel for HBM byte counters. */
int rh_copy_probe(int64_t nbytes, float *ms);

/* Kernel-only timing helper: nlaunch Godunov launches of the level (fused mode,
 * dt fixed), timed with HIP events on the library stream; ms_per_launch is the
 * mean launch duration.  The state is advanced nlaunch times. */
int rh_time_godunov(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr,
                    double smallc2, double courant_factor, const double g[3], int slope, int riemann,
                    double dt, int nlaunch, float *ms_per_launch);

/* ---- Batched AMR stepping -----------------------------------------------------
 * Prepare the levels levelmin..levelmax of the current context (octs of level l
 * at head[l-1] .. head[l-1]+noct[l-1]-1, levels contiguous coarse to fine, each
 * built with rh_build_nbor) for batched steps with one shared dt: stores one
 * Godunov descriptor per level and the work list of all their bricks, and
 * computes the first dt (courant_fine over the leaf cells of all levels,
 * min with dx_finest/smallc).  dx of level l is boxlen/2^l. */
int rh_amr_prepare(int levelmin, int levelmax, const int32_t *head, const int32_t *noct, double boxlen,
                   double gamma, double smallr, double smallc2, double courant_factor, const double g[3],
                   int slope, int riemann);

/* nsteps steps of every prepared level: ghost octs refreshed from the parent
 * level, one work-list Godunov launch over the bricks of all levels (refluxes
 * gathered per ghost oct, then added to the parent cells), upload_fine finest
 * first, courant_fine of the new state over all levels (dt of the next step).
 * use_graph != 0 replays pairs of steps as a captured hipGraph.  Returns the dt
 * of the next step, the time advanced since rh_amr_prepare, the conservation
 * sums over all leaf cells {mass, ekin, eint, emag}, and the mean step time
 * (HIP events on the library stream). */
int rh_amr_run(int nsteps, int precision, int use_graph, double *dt_last, double *t_total, double sums[4],
               float *ms_per_step);

/* ---- Contexts and ensemble serving ----------------------------------------
 * A context holds one mesh (arrays, hash, neighbour and brick tables, dt and
 * time on the device).  Context 0 always exists; every entry point of
 * ramses2_hip.h acts on the current context.  All contexts share the library
 * stream, so work of different jobs can share launches. */
int rh_ctx_create(int32_t *id);          /* new, empty context (not selected)      */
int rh_ctx_select(int id);               /* make it current                         */
int rh_ctx_current(int32_t *id);
int rh_ctx_destroy(int id);              /* free its device memory; id 0 is reset   */
int rh_ctx_footprint(int id, int64_t *bytes);   /* device bytes the context holds  */

/* Give the current (empty) context a copy of the mesh of context src_id: level
 * parameters, oct records, device hash, neighbour, ghost and brick tables, by
 * device-to-device copies on the library stream; uold/unew are zeroed.  Serving
 * builds one template context per grid size and clones it per job. */
int rh_ctx_clone_mesh(int src_id);

/* Device blocks released by destroyed contexts are kept for reuse by requests
 * of the same size (no device synchronisation on release).  This hands them
 * back to the device and reports how many bytes were cached. */
int rh_cache_release(int64_t *bytes_released);

/* Device bytes a context needs for a uniform 3D level of n^3 octs (n octs per
 * dimension) with the current brick shape: the admission estimate of the
 * serving loop (rh_ctx_footprint reports the bytes actually held). */
int rh_estimate_uniform_bytes(int n_octs_per_dim, int ncachemax, int hash_size, int64_t *bytes);

/* Initial state of octs head_idx.. of the current context, written on the
 * device into uold and unew (ncell cells per dimension, gamma):
 *   kind 1, blast:   params = {rho0, p0, p_blast, lo, -}: (rho0, 0, p0), and p_blast in
 *                    the 2^3 cells with every integer coordinate in [lo, lo+2);
 *   kind 2, Riemann: params = {rho_l, p_l, rho_r, p_r, axis}: left state where
 *                    (i_axis + 0.5)/ncell < 0.5, velocity 0. */
int rh_init_ic(int head_idx, int num_octs, int ncell, int kind, const double params[5], double gamma);

/* Make the current context's level a job for rh_batch_step: stores its Godunov
 * parameters and computes the dt of its first step with a Courant pass; the
 * job time starts at 0.  The level must be built (rh_build_nbor). */
int rh_job_prepare(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr,
                   double smallc2, double courant_factor, const double g[3], int slope, int riemann);

/* nsteps time steps of each of the njobs prepared contexts ids[] with shared
 * launches: per step one work-list Godunov kernel over the bricks of
 * all the jobs (each job's Courant condition of the new state reduced in the
 * kernel's epilogue) and one per-job final reduction (dt of the job's next
 * step, its time).  Jobs advance independently: a job's result does not depend
 * on the other jobs of the batch.  Asynchronous. */
int rh_batch_step(const int32_t *ids, int njobs, int precision, int nsteps);

/* Synchronous read of a job's {t, dt of the next step, mass, ekin, eint, emag,
 * dt} (the last four from the last Courant reduction). */
int rh_job_status(int id, double out[7]);

/* A fence on the library stream, for a request loop that keeps one batched
 * step in flight while it retires the jobs of the one before: records an
 * event after every piece of work enqueued so far and, in stream order, copies
 * the rh_job_status values of the n contexts ids[] (n may be 0) into pinned
 * host memory the library holds.  Asynchronous; *fence names it until
 * rh_fence_wait.  At most 64 fences are outstanding (RH_ERR_STATE beyond). */
int rh_fence_record(const int32_t *ids, int n, int32_t *fence);

/* Block until the fence has passed, write its n statuses (7 values each, in
 * the order of ids[]) to out (may be null), and release the fence. */
int rh_fence_wait(int32_t fence, double *out);

#ifdef __cplusplus
}
#endif
#endif
