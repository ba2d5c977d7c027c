/*
 * ramses2_hip.h -- C-ABI of the MI355X (gfx950) fine-level Godunov hydro step.
 *
 * Drop-in boundary for the RAMSES II level kernels.  The entry points mirror
 * the reference's Metal bridge (metal/metal_interface.f90, bind(C) "mtl_*")
 * with FP64 arguments in place of the Metal backend's FP32 ones, so that a
 * Fortran "#elif defined(_HIP)" branch at the existing call sites
 * (hydro/godunov_fine.f90:29-35,100-106,168-175, hydro/courant_fine.f90:41-49,
 * amr/adaptive_loop.f90:110-116) can call them through iso_c_binding exactly as
 * metal/metal_runner.f90 calls mtl_*.
 *
 * Conventions (same as the reference ABI):
 *   - oct indices are 1-based Fortran indices (head_idx, ngridmax);
 *   - hydro arrays use the Fortran layout u(1:twotondim, 1:nvar, 1:ngrid);
 *   - scalars by value, small arrays by pointer, outputs by pointer;
 *   - after rh_set_grid_device the library owns device-resident uold/unew/grid,
 *     hash and neighbour tables; host copies are stale until
 *     rh_transfer_grid_host.
 * Every entry point returns 0 on success or a nonzero code; rh_last_error()
 * returns the message of the last failure.  Calls are not re-entrant: one host
 * thread per process drives the library, as one MPI rank drives the reference.
 */
#ifndef RAMSES2_HIP_H
#define RAMSES2_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define RH_OK 0
#define RH_ERR_HIP 1        /* a HIP runtime call failed                     */
#define RH_ERR_ARG 2        /* invalid argument (sizes, indices, ids)        */
#define RH_ERR_STATE 3      /* call order (e.g. godunov before build_nbor)   */
#define RH_ERR_MESH 4       /* mesh inconsistency (missing neighbour, hash full) */
#define RH_ERR_NOMEM 5      /* device allocation failed                      */

/* Riemann solver ids, hydro/hydro_parameters.f90:30-32 */
#define RH_RIEMANN_LLF 1
#define RH_RIEMANN_HLL 2
#define RH_RIEMANN_HLLC 3

/* Oct record, the layout of type(oct) (amr/oct_commons.f90:5-10) for NDIM=3,
 * nhilbert=1: 32 bytes; same as oct_t of metal/metal_types.h:1-17.
 * For NDIM=1 the reference record is 24 bytes: {int64 hkey; int32 ckey;
 * uint8 refined[2]; (2 pad); int32 lev; (4 pad)}; rh_set_grid_device takes
 * the record size from ndim. */
typedef struct {
    int64_t hkey;
    int32_t ckey[3];
    uint8_t refined[8];
    int32_t lev;
} rh_oct3_t;

/* Replaces mtl_init (metal/metal_interface.f90:8; metal/metal_bridge.mm).
 * Selects the current HIP device and creates the library stream. */
int rh_init(void);

/* Replaces mtl_alloc_amr (metal/metal_interface.f90:14-18).  Allocates
 * uold/unew/grid/neighbour storage for ngridmax+ncachemax octs and the device
 * hash table.  ndim is taken from twotondim (2 -> 1D, 8 -> 3D). */
int rh_alloc_amr(int ngridmax, int ncachemax, int nvar, int twotondim, int hash_size);

/* Replaces mtl_upload_level_params (metal/metal_interface.f90:31-41).
 * ckey_max(1:nlevelmax+1), key_off(1:nlevelmax+1), box_ckey_min/max(1:3,1:nlevelmax+1). */
int rh_upload_level_params(const int32_t *ckey_max, const int64_t *key_off,
                           const int32_t *box_ckey_min, const int32_t *box_ckey_max,
                           const int32_t periodic[3], int nlevelmax);

/* Boundary types for non-periodic directions, &BOUNDARY_PARAMS bound_type
 * (amr/boundaries.f90:39-165): bound_type[2*d+0] low side, [2*d+1] high side,
 * 1 = reflexive, 2 = zero gradient.  No reference counterpart in the Metal
 * ABI (the Metal backend supports periodic boxes only); the CUDA backend
 * builds the same ghost octs in gpu/gpu_refine.cuf:2231 make_cache_octs. */
int rh_set_boundaries(const int32_t bound_type[6]);

/* Replaces mtl_set_grid_device (metal/metal_interface.f90:46-55): H->D copy of
 * uold, unew and the oct records of the first ngridmax octs.  bold must be NULL
 * (no MHD).  grid points to ngridmax oct records of the reference layout. */
int rh_set_grid_device(const double *uold, const double *unew, const void *bold,
                       const void *grid, int ngridmax, int nvar, int twotondim);

/* Replaces mtl_transfer_grid_host (metal/metal_interface.f90:67-74): D->H copy
 * of uold for the first ngridmax octs (synchronous). */
int rh_transfer_grid_host(double *uold, void *bold, int ngridmax, int nvar, int twotondim);

/* D->H copy of unew (same layout); no Metal counterpart, used by tests and by
 * callers that inspect unew between godunov and set_uold. */
int rh_transfer_unew_host(double *unew, int ngridmax, int nvar, int twotondim);

/* Replaces mtl_insert_hash_all (metal/metal_runner.f90:551-554, refine.metal:796):
 * inserts octs head_idx..head_idx+num_octs-1 into the device hash, keyed by
 * (lev, ckey) as key_off(lev) + ix + iy*nx + iz*nx^2 (gpu/gpu_hash.cuf:240-245). */
int rh_insert_hash_all(int head_idx, int num_octs, int hash_size);

/* Replaces mtl_build_nbor (metal/metal_interface.f90:121-133): resolves, through
 * the device hash, the 3^ndim neighbour octs of every oct of the level
 * (slot 1+i+3j+9k, periodic wrap as hydro/godunov_fine.f90:463-468) and the
 * oct-brick windows the Godunov kernel gathers its stencil from.  Keys outside
 * a non-periodic domain resolve to boundary ghosts built from the adjacent
 * interior oct (amr/nbors_utils.f90:356-420). */
int rh_build_nbor(int head_idx, int num_octs, int hash_size, int ckey_max_l, int64_t key_off_l,
                  const int32_t box_min[3], const int32_t box_max[3], const int32_t periodic[3]);

/* Replace mtl_set_unew / mtl_set_uold (metal/metal_interface.f90:86-97),
 * i.e. set_unew / set_uold of hydro/godunov_fine.f90:114-148,183-214. */
int rh_set_unew(int head_idx, int num_octs);
int rh_set_uold(int head_idx, int num_octs);

/* Replaces mtl_cmpdt (metal/metal_interface.f90:101-116), i.e. courant_fine
 * (hydro/courant_fine.f90:57-168): Courant time step and conservation sums over
 * the leaf cells of the level.  Synchronises (the results are host scalars). */
int rh_cmpdt(int head_idx, int num_octs, double dx, double gamma, double smallr,
             double smallc2, double courant_factor, const double g[3],
             double *mass, double *ekin, double *eint, double *emag, double *dt);

/* No counterpart in the reference: rh_cmpdt for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc).  Same arguments and results as rh_cmpdt; its per-block
 * partial sums and result live in a static device array of the library
 * (4096 partials of 4 doubles + 8 doubles, 128 KB per GPU, held until the
 * process ends) instead of library-allocated memory, and it calls no
 * hipMalloc, hipFree or allocating routine.  The reduction covers any oct
 * range (at most 4096 blocks, each striding over the cells).  Calls that
 * overlap on one GPU share that array and race; nothing in it needs clearing
 * (each call overwrites what it reads). */
int rh_cmpdt_noalloc(int head_idx, int num_octs, double dx, double gamma, double smallr, double smallc2,
                     double courant_factor, const double g[3], double *mass, double *ekin, double *eint,
                     double *emag, double *dt);

/* Replaces mtl_godunov (metal/metal_interface.f90:537-558), i.e. godunov_fine /
 * godfine1 / unsplit (hydro/godunov_fine.f90:43-79,244-1342, hydro/umuscl.f90:20-292):
 * unew += conservative update of every oct of the level, plus the refluxing of
 * coarse-fine faces into the parent cells when ilevel > levelmin.
 * num_subgrids is the oct count (NSUBGRID=1 numbering).  smax_cap is the LLF/HLL
 * wave-speed cap courant_factor*dx/dt that unsplit sets (umuscl.f90:99) and the
 * Metal runner passes precomputed (metal/metal_runner.f90:284). */
int rh_godunov(int head_idx, int num_subgrids, int ngridmax, int ilevel, int levelmin, int levelmax,
               double gamma, double smallr, double smallc2, double dt, double dx, double smax_cap,
               int slope, int riemann, const double g[3]);

/* Replaces mtl_upload (metal/metal_interface.f90:520-527), i.e. upload_fine
 * (hydro/upload.f90:74-283): average each oct of the fine level
 * head_idx..head_idx+num_octs-1 into its parent cell of uold;
 * internal_energy != 0 averages the internal rather than the total energy
 * (interpol_var == 1).  rh_build_nbor of the fine level must have run. */
int rh_upload(int head_idx, int num_octs, int internal_energy, double gamma, double smallr, double smallc2);

/* &REFINE_PARAMS interpol_var (0: rho, rho u, E; 1: rho, rho u, rho e) and
 * interpol_type (0 injection, 1 MinMod, 2 MonCen, 3 central) used to fill the
 * ghost octs of missing coarse-fine neighbours (interpol_hydro,
 * hydro/interpol_hydro.f90:5-113, as godfine1 calls it at
 * hydro/godunov_fine.f90:555).  No Metal counterpart: the CUDA backend reads
 * them from its module variables (gpu/gpu_refine.cuf:2231 make_cache_octs). */
int rh_set_interpolation(int interpol_var, int interpol_type);

/* ---- Refinement flags (the flagging step of the refine path) -------------
 * flag1(1:8, 1:ngridmax) int32 on the device, as m%flag1 (amr/amr_commons.f90).
 * Replaces mtl_upload_flag1 (metal/metal_interface.f90:59-63): host -> device. */
int rh_upload_flag1(const int32_t *flag1, int ngridmax);
/* Device -> host copy of flag1 (the reverse of rh_upload_flag1; the Metal
 * bridge's flag1 is host-shared memory, metal/metal_bridge.mm:1016). */
int rh_transfer_flag1_host(int32_t *flag1, int ngridmax);
/* Replaces mtl_reset_flag1 (metal/metal_interface.f90:200-203): flag1 = 0 for
 * the cells of octs head_idx..head_idx+num_octs-1. */
int rh_reset_flag1(int head_idx, int num_octs);
/* Replaces mtl_count_flag1 (metal/metal_interface.f90:218-222): the number of
 * flagged cells of the octs (an int function there, an output here). */
int rh_count_flag1(int head_idx, int num_octs, int32_t *count);
/* Replaces mtl_hydro_flag (metal/metal_interface.f90:225-241), FP64: the
 * gradient criteria of hydro_flag / hydro_refine (hydro/hydro_flag.f90:7-159,
 * hydro/godunov_utils.f90:6-227) on the level's current uold: flag1 = 1 where
 * 2 max(|d_R - d|/(d_R + d + floor_d), |d - d_L|/(d + d_L + floor_d)) >
 * err_grad_d, or the same of the pressure > err_grad_p (-1 disables a
 * criterion), d_L/d_R the face neighbours in each direction (from the coarser
 * level where the level has no oct there).  Set flags stay set.  3D only. */
int rh_hydro_flag(int head_idx, int num_octs, double gamma, double smallr, double smallc2, double err_grad_d,
                  double err_grad_p, double floor_d, double floor_p);

/* Replaces mtl_device_sync (metal/metal_interface.f90:529-531). */
int rh_device_sync(void);

/* Message of the last failing call ("" if none). */
const char *rh_last_error(void);

/* Free all device state (no reference counterpart; process exit also frees). */
int rh_finalize(void);

#ifdef __cplusplus
}
#endif
#endif
