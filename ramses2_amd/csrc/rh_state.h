// Library state and shared declarations of the gfx950 hydro library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/ramses2_hip.h"
#include "rh_godunov.h"

namespace rh {

// Status of the last call, reported through rh_last_error().
void set_error(const std::string &msg);

#define RH_HIP_CHECK(expr)                                                                  \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::rh::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));             \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

// Oct-brick decomposition of one level: bricks of BX*BY*BZ octs, each with a
// window of (BX+2)*(BY+2)*(BZ+2) oct slots (the brick plus a one-oct halo).
// Window values: >0 oct index (1-based), <0 boundary ghost built from oct
// -value, 0 no oct.
struct LevelBricks {
    int ilevel = 0;
    int head = 0, noct = 0;
    int bx = 0, by = 0, bz = 0;
    int nbricks = 0;
    int32_t *win = nullptr;   // [nbricks][bz+2][by+2][bx+2]
    int32_t *org = nullptr;   // [nbricks][3] oct key of the brick's first oct
    int32_t *zr = nullptr;    // [nbricks] z layers holding octs of the level: zlo | zhi << 16
    std::vector<int32_t> h_zr;  // host copy of zr
    int any_ref = 1;          // some cell of the level's octs is refined
    int bnd = 1;              // some dimension is not periodic (boundary ghosts in the windows)
    int off[3] = {0, 0, 0};   // brick lattice offset: origins at boxmin + off + multiples of the shape
    int boxmin[3] = {0, 0, 0};
    int boxmax[3] = {0, 0, 0};
};

// Neighbour table of one level: 3^ndim slots per oct of the level, the
// parent cell of each oct (father, for restriction), and the level's ghost
// octs (missing neighbours interpolated from the parent level).
struct LevelNbor {
    int ilevel = 0;
    int head = 0, noct = 0;
    int32_t *nbor = nullptr;    // [noct][3^ndim]
    int32_t *father = nullptr;  // [noct][2] (igrid, icell) at ilevel-1, or 0
    int ghost_first = 0;        // first ghost oct index (1-based) of the level
    int nghost = 0;
    int boxmin[3] = {0, 0, 0};
    int boxmax[3] = {0, 0, 0};
};

// Batched multi-level stepping of an AMR mesh (rh_amr_prepare / rh_amr_run):
// one work-list Godunov launch over the bricks of every level per step.
struct AmrPlan {
    bool ready = false;
    int levelmin = 0, levelmax = 0;
    int head0 = 0, noct_all = 0;   // the levels' octs form one contiguous range
    double boxlen = 1.0, dx_finest = 0.0;
    double gamma = 1.4, smallr = 0.0, smallc2 = 0.0, courant = 0.0, g[3] = {0, 0, 0};
    int slope = 0, riemann = 0;
    int ndesc = 0, nitems = 0;
    GodArgs *d_descs = nullptr;    // [2][ndesc]: arrays in both orders
    std::vector<GodArgs> h_descs;  // host copy of d_descs
    int2 *d_items = nullptr;       // [nitems] (level descriptor, brick), grouped by brick shape
    // one work-list launch per brick shape among the levels
    struct Group {
        int bx, by, bz;
        int first, count;          // items [first, first+count) of d_items
        int range[9];              // the items of XCD x: [range[x], range[x+1]), equal modelled work
    };
    std::vector<Group> groups;
    double *racc = nullptr;        // [max(nghost_total,1)][nslot][5] reflux sums per ghost oct
    int nslot = 1;                 // 1, or kRefluxSlotsPerGhost in deterministic mode
    std::vector<int> levels;       // levels with octs, coarse to fine
};

struct State {
    bool inited = false;
    int device = 0;
    hipStream_t stream = nullptr;
    int ndim = 3, nvar = 5, ttd = 8;
    int ngridmax = 0, ncachemax = 0, ngridtot = 0;
    int64_t hash_cap = 0;   // power of two
    double *uold = nullptr, *unew = nullptr;
    int32_t *ckey = nullptr;       // [ngridtot][3]
    int32_t *lev = nullptr;        // [ngridtot]
    uint8_t *refined = nullptr;    // [ngridtot][8]
    uint64_t *hkeys = nullptr;     // hash keys (empty = ~0)
    int32_t *hvals = nullptr;
    // per-level parameters (index l-1)
    int nlevelmax = 0;
    std::vector<int32_t> ckey_max;
    std::vector<int64_t> key_off;
    std::vector<int32_t> box_min, box_max;  // [l][3]
    int periodic[3] = {1, 1, 1};
    int btype[3][2] = {{2, 2}, {2, 2}, {2, 2}};
    // host copy of the oct keys (brick enumeration)
    std::vector<int32_t> h_ckey, h_lev;
    std::vector<uint8_t> h_refany;  // per oct: some cell refined (from the last grid upload)
    std::map<int, LevelBricks> bricks;  // by level
    std::map<int, LevelNbor> nbors;     // by level
    // brick shape used for 3D levels
    int brick_bx = 8, brick_by = 6, brick_bz = 8;
    // Godunov precision: 0 FP64, 1 FP32 reconstruction/fluxes + FP64 accumulation
    int precision = 0;
    // Courant pass of the next step fused into the Godunov kernel (rh_level_run):
    // -1 automatic (fused, in both precisions), 0 off, 1 on
    int fused_courant = -1;
    // rh_level_run's captured pairs of steps, kept for later calls with the
    // same arguments, arrays and tables (key: the bits of all of them); two
    // entries, one per order of the two state arrays (an odd step count
    // leaves them swapped)
    std::vector<uint64_t> graph_key[2];
    hipGraph_t level_graph[2] = {nullptr, nullptr};
    hipGraphExec_t level_exec[2] = {nullptr, nullptr};
    // timer events
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // single-level device time stepping: dt and accumulated time on the device
    double *dt_dev = nullptr;
    double *t_dev = nullptr;
    // ghost octs (interpolated missing neighbours), indices ngridmax+1..ngridmax+nghost_total
    int nghost_total = 0;
    int32_t *ghost_parent = nullptr;  // [ncachemax][2]
    int32_t *ghost_src = nullptr;     // [7][2][ncachemax]: parent stencil of each ghost, by position in ghost_order
    int32_t *ghost_srckey = nullptr;  // [ncachemax][7][3]
    int32_t *ghost_order = nullptr;   // [ncachemax] ghost indices, each level's range in parent-cell order
    int32_t *ghost_key = nullptr;     // [ncachemax][3]
    int32_t *d_counter = nullptr;     // device scratch counters [4]
    int32_t *flag1 = nullptr;         // [ngridmax][8] refinement flags (allocated at first use)
    // deterministic refluxes (rh_set_deterministic): each coarse-fine face
    // writes its own slot, folded in a fixed order; racc_det holds the slots
    // of the per-level drop-in calls
    int deterministic = 0;
    // artificial diffusion coefficient (hydro_params_t%difmag); > 0 adds the
    // cmpdiff pass after each 3D Godunov launch
    double difmag = 0.0;
    double *racc_det = nullptr;
    long racc_det_cap = 0;            // ghosts it has room for
    int interpol_var = 0, interpol_type = 1;
    // reduction scratch
    double *red_part = nullptr;
    int red_cap = 0;
    double *red_out = nullptr;     // device, 8 doubles
    double *h_red = nullptr;       // pinned host, 8 doubles
    std::string last_error;
    // device allocations of this context (bytes by pointer) and their total
    std::map<void *, size_t> allocs;
    int64_t dev_bytes = 0;
    // a job prepared for batched stepping (one 3D level, dt on the device)
    bool job_ready = false;
    GodArgs job;
    int job_bx = 0, job_by = 0, job_bz = 0;
    double job_dtmax = 0.0;
    AmrPlan amr;
};

// Modelled work of a brick of bx x by octs that marches `planes` cell planes
// (its occupied z layers): per plane, plus two halo planes each side, the P
// phase converts (2bx+4)(2by+4) cells and the T/F/U phases run (2bx+2)(2by+2)
// columns, a P cell costing about a third of a column-plane.
inline double brick_work(int bx, int by, double planes)
{
    const double pcell = 0.35 * (2 * bx + 4) * (2 * by + 4), col = (double)(2 * bx + 2) * (2 * by + 2);
    return pcell * (planes + 4) + col * (planes + 2);
}

// Leading dimension of the ghost tables laid out by ghost slot.
inline int ghost_ld(const State &s) { return s.ncachemax > 0 ? s.ncachemax : 1; }

// Godunov arguments of a 3D level of the current context (fused mode,
// unew = uold + dF); lb receives the level's bricks.  Returns an RH_ERR code.
int fill_args3(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
               double courant_factor, const double g[3], int slope, int riemann, const double *dt_dev, double dt,
               GodArgs &a, const LevelBricks *&lb);
int fill_args1(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
               double courant_factor, const double g[3], int slope, int riemann, const double *dt_dev, double dt,
               God1Args &a);
// Grow the current context's per-brick partial-sum buffer to n entries.
int ensure_partials(int n);
// The artificial-diffusion pass of a 3D level after its Godunov launch
// (unew += cmpdiff fluxes of uold), when difmag > 0.
hipError_t enqueue_diffusion(State &s, int ilevel, int head_idx, int num_octs, double dx, double smallr,
                             const double g[3], const double *dt_dev, double dt, int reflux, double *racc,
                             int rdet);
// Destroy the context's cached rh_level_run graph (its arrays or tables change).
void drop_level_graph(State &s);

// The current context (the drop-in entry points act on it).
State &st();

// Device allocation / release accounted to the current context.
hipError_t dmalloc(void **p, size_t n);
void dfree(void *p);

// Bytes held in the allocation cache; release them to the device.
int64_t cached_bytes();
void release_cached();

// Context table: id 0 is the default context, always present.
int ctx_count();
State *ctx_get(int id);
int ctx_add(State *s);
void ctx_set_slot(int id, State *s);
void ctx_select(State *s);

}  // namespace rh
