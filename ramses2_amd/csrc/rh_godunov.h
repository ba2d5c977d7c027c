// Kernels of the fine-level Godunov step: oct-brick z-marching kernel (3D),
// cell kernel (1D).  Shared by the drop-in entry points and by the batched /
// serving extensions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rh_physics.h"

namespace rh {

// Arguments of one Godunov launch over the bricks of one level (or, for the
// batched path, over a work list of bricks of several levels/jobs).
struct GodArgs {
    const double *uold;       // stencil source
    double *unew;             // destination
    const double *ubase;      // value the update is added to: unew (accumulate) or uold (fused)
    const uint8_t *refined;   // [oct][8]
    const int32_t *win;       // brick windows
    const int32_t *org;       // brick origins (oct keys)
    const int32_t *zr;        // per brick: octs present in z layers zlo..zhi, zlo | zhi << 16 (null: all)
    int nbricks;
    int head, noct;           // 1-based active oct range of the level
    int boxmin[3], boxmax[3];
    int btype[3][2];
    const double *dt_dev;     // if non-null, dt is read from device memory
    double dt;
    double dx, inv_dx;
    double gamma, smallr, smallc2, courant_factor;
    double smax_cap;          // used when use_smax != 0, else courant_factor*dx/dt
    int use_smax;
    double g[3];
    int slope, riemann;
    int xcd_remap;            // remap block ids so each XCD gets contiguous bricks
    // set by launch_godunov3d: bricks at the end of each XCD's range that run
    // as two half-brick work items (0 for levels with refluxes)
    int nsplit = 0;
    // coarse-fine refluxing (levels above levelmin): octs > ngridmax are
    // interpolated ghosts whose parent cell is ghost_parent[2*(v-ngridmax-1)+{0,1}]
    int reflux;
    // 0 when no cell of the level is refined: the kernel then skips the
    // refined flags (their loads, the zeroing of refined faces)
    int any_ref;
    // 0 when the level is periodic in every dimension: its brick windows hold
    // no boundary ghosts, and the gather skips their synthesis
    int bnd = 1;
    int ngridmax;
    const int32_t *ghost_parent;
    // fused courant_fine of the updated state: per-brick partials
    // (mass, ekin, eint, dt) at cpart[4*brick]
    int courant;
    double *cpart;
    // when non-null, refluxes accumulate into racc[5*(ghost index)] (one slot
    // per ghost oct, i.e. per parent cell) instead of into the parent cell of
    // unew; a later fold adds them (several levels in one launch)
    double *racc;
    // 1: racc holds kRefluxSlotsPerGhost slots per ghost, each written by one
    // face (deterministic reflux); 0: one accumulator per ghost (atomics)
    int rdet = 0;
};

// Work list of bricks over several descriptors (levels / jobs): item i is
// (descriptor index, brick index); range[0..8] splits the items into one
// contiguous range per XCD.  Descriptors must stay unchanged while a launch
// runs (the kernel reads them through the constant address space).
struct WorkList {
    const GodArgs *descs;
    const int2 *items;
    int range[9];
};

// 1D arguments: one thread per cell of the level, 3-slot neighbour table.
struct God1Args {
    const double *uold;
    double *unew;
    const double *ubase;
    const uint8_t *refined;
    const int32_t *nbor;      // [noct][3]
    int head, noct;
    int boxmin, boxmax;
    int btype[2];
    const double *dt_dev;
    double dt, dx, inv_dx;
    double gamma, smallr, smallc2, courant_factor, smax_cap;
    int use_smax;
    double g;
    int slope, riemann;
    int reflux;
    int ngridmax;
    const int32_t *ghost_parent;
    double *racc = nullptr;   // as GodArgs::racc / rdet
    int rdet = 0;
};

// Launch the 3D brick kernel for brick shape (bx,by,bz) in precision prec
// (0 = FP64, 1 = FP32 reconstruction/fluxes with FP64 accumulation).
// *nparts (if non-null) receives the number of Courant partials the launch
// writes (bricks plus split halves), at most 2 nbricks + 8.
hipError_t launch_godunov3d(const GodArgs &a, int bx, int by, int bz, int prec, hipStream_t s,
                            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, int *nparts = nullptr);
hipError_t launch_godunov1d(const God1Args &a, int prec, hipStream_t s);
// The specialised brick bodies (GEN = false) are used where a level allows
// them unless this is 0 (rh_set_specialized; the general body then runs
// everywhere, for checking that both give the same bits).
extern int g_specialized;
// Work-list launch: items[i] = (descriptor index, brick index), one workgroup per item.
// gen = false: every descriptor is a periodic level without refined cells or
// refluxes (bnd, any_ref and reflux all 0) whose update adds to uold (ubase ==
// uold), which the specialised body needs.
hipError_t launch_godunov3d_wl(const WorkList &w, int bx, int by, int bz, int slope, int riemann, int prec,
                               hipStream_t s, bool gen = true);

// Courant reduction over cells of octs [head, head+noct): per-block partials,
// then a fixed-order final pass -> out[0..4] = mass, ekin, eint, emag, dt.
// Fixed-order reduction of fused-Courant partials: t_acc += the dt the step
// used (*dt_inout on entry), then *dt_inout = min(dtmax, partial dts).
// Per-job final Courant reduction of a batch (one block per job).
struct CourantJob {
    const double *part;   // per-brick partials of the job
    int nparts;
    double dtmax;         // dx / smallc
    double *out;          // {mass, ekin, eint, emag, dt}
    double *dt_inout;     // dt of the step just taken in, dt of the next step out
    double *t_acc;        // time of the job, advanced by the dt taken
};
// courant_fine over the leaf cells of the contiguous oct range [head,
// head+noct) spanning several levels: dx of each oct is boxlen*2^-lev.
// dx_finest sets the initial dt bound dx/smallc.  Partials + final reduction
// as launch_courant.  mixed: each cell's signal speed and dt in FP32 (the
// mixed-precision steps); the sums and the min-reduction stay FP64.
hipError_t launch_courant_ml(const double *uold, const uint8_t *refined, const int32_t *lev, int head, int noct,
                             double boxlen, double dx_finest, double gamma, double smallr, double smallc,
                             double courant_factor, const double g[3], double *partials, int max_blocks, double *out,
                             double *dt_out, hipStream_t s, double *t_acc, bool mixed = false);
// racc[0 .. 5*n) = 0 (a kernel: safe inside captured graphs)
hipError_t launch_zero(double *p, long n, hipStream_t s);
// unew(parent cell of ghost g) += the nslot accumulators of g (racc[5*nslot*g
// ...], added in slot order) for the ghosts order[0 .. nghost), which are
// cleared; nslot = 1 (atomic refluxes) or kRefluxSlots (deterministic mode)
hipError_t launch_reflux_fold(double *unew, double *racc, const int32_t *ghost_parent, const int32_t *order,
                              int nghost, int ttd, hipStream_t s, int nslot = 1);
// Slots per ghost oct of the deterministic reflux: 6 faces x 4 cells
constexpr int kRefluxSlotsPerGhost = 24;
hipError_t launch_courant_reduce_batch(const CourantJob *jobs, int njobs, hipStream_t s);
hipError_t launch_courant_reduce(const double *partials, int nparts, double dtmax, double *out, double *dt_inout,
                                 double *t_acc, hipStream_t s);

hipError_t launch_courant(const double *uold, const uint8_t *refined, int head, int noct, int ndim,
                          double dx, double gamma, double smallr, double smallc, double courant_factor,
                          const double g[3], double *partials, int max_blocks, double *out,
                          double *dt_out, hipStream_t s, double *t_acc = nullptr);

// Ghost octs of a level: fill each from its parent cell and the 2*ndim
// neighbouring parent cells by interpol_hydro (hydro/interpol_hydro.f90:5-113).
// src[g*7 + j] = (oct code, cell index 1-based) of u1(j); keys for boundary codes.
// Ghost refresh of one or more levels in one launch: segment k covers
// threads [seg_t[k], seg_t[k+1]) and the ghosts order[seg_o[k] ...] of one
// level, whose parent level's box is seg_box[k] (min xyz, max xyz).
constexpr int kMaxGhostSegs = 16;
struct GhostArgs {
    double *uold;
    const uint8_t *refined;
    const int32_t *src;      // [14][src_ld]: word e (slot e/2, oct code or cell) of each ghost, by order position
    int src_ld;
    const int32_t *srckey;   // [nghost][7][3]
    const int32_t *order;    // ghost indices (0-based, all levels) in parent-cell order
    int gi0;                 // ghost index of src[0] (order values are offset by it)
    int first;               // oct index (1-based) of ghost gi0
    int nghost;              // threads: ghosts of all segments
    int ndim;
    int interpol_var, interpol_type;
    double smallr;
    int btype[3][2];
    int nseg;
    int seg_t[kMaxGhostSegs + 1];
    int seg_o[kMaxGhostSegs];
    int seg_box[kMaxGhostSegs][6];
};
hipError_t launch_refresh_ghosts(const GhostArgs &a, hipStream_t s);

// Restriction (upload_fine, hydro/upload.f90:74-283): average each fine oct of
// [head, head+noct) into its parent cell father[2*i], father[2*i+1].
hipError_t launch_upload(double *uold, const int32_t *father, int head, int noct, int ndim, int internal_energy,
                         double smallr, hipStream_t s);

// Artificial diffusion of a 3D level (difmag > 0), after its Godunov launch:
// unew += the cmpdiff fluxes of uold (hydro/umuscl.f90:2392-2540).
struct DiffArgs {
    const double *uold;
    double *unew;
    const uint8_t *refined;
    const int32_t *nbor;   // [noct][27] of the level
    const int32_t *ckey;   // [ngrid][3]
    int head, noct, ndim;
    int boxmin[3], boxmax[3];
    int refl;              // reflexive faces, bit 2*dim+side
    const double *dt_dev;  // if non-null, dt is read from device memory
    double dt, dx, difmag, smallr, g[3];
    int reflux, ngridmax;
    const int32_t *ghost_parent;
    double *racc;          // as GodArgs::racc / rdet
    int rdet;
};
hipError_t launch_diffusion3d(const DiffArgs &a, hipStream_t s);

// Ordered brick-shape support list: returns true when (bx,by,bz) has a compiled kernel.
bool brick_shape_supported(int bx, int by, int bz);

}  // namespace rh
