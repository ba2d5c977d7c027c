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

/* Oct-brick shape (octs per brick in x, y, z) used by the next rh_build_nbor
 * for 3D levels; the autotuner (ramses2_amd/autotune.py) picks it. */
int rh_set_brick_shape(int bx, int by, int bz);
int rh_get_brick_shape(int32_t *bx, int32_t *by, int32_t *bz);

/* HIP-event timer on the library stream: milliseconds between start and stop. */
int rh_timer_start(void);
int rh_timer_stop(float *ms);

/* nsteps full time steps of one level that is the whole mesh (uniform grid):
 * courant_fine (dt reduced on the device) -> godunov_fine with set_unew and
 * set_uold folded in (unew = uold + dF, then the two arrays swap roles).
 * dt_fixed > 0 imposes dt instead of the Courant value.  dt_last, t_total and
 * the conservation sums of the last Courant pass are returned when non-NULL.
 * use_graph != 0 replays the step pair as a captured hipGraph.  kernel_ms, when
 * non-NULL (and use_graph == 0), receives the mean duration of the Godunov
 * launches, timed with HIP events on the library stream inside the steps. */
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

#ifdef __cplusplus
}
#endif
#endif
