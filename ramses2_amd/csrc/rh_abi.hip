// C-ABI of the gfx950 hydro library (include/ramses2_hip.h): device state,
// device hash, neighbour / brick tables and the drop-in rh_* entry points.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <map>
#include <string>
#include <vector>

#include "../../include/ramses2_hip.h"
#include "rh_godunov.h"
#include "rh_mesh.h"
#include "rh_state.h"

namespace rh {

// Contexts: each holds one mesh with its device arrays and tables.  Context 0
// always exists; the drop-in entry points act on the current context.
static State g_default;
static State *g_cur = &g_default;
static std::vector<State *> g_ctx{&g_default};
static hipStream_t g_stream = nullptr;  // library stream, shared by all contexts
State &st() { return *g_cur; }
void set_error(const std::string &msg) { g_cur->last_error = msg; }

// Device blocks released by a context are kept for reuse by a later request
// of exactly the same size (contexts of one mesh size come and go in ensemble
// serving), up to kCacheCap bytes.  hipFree synchronises the device; a cached
// block needs no synchronisation because every use of it is ordered on the
// one library stream.
static std::multimap<size_t, void *> g_cache;
static size_t g_cache_bytes = 0;
static constexpr size_t kCacheCap = size_t(32) << 30;

static void cache_release_all()
{
    for (auto &kv : g_cache) (void)hipFree(kv.second);
    g_cache.clear();
    g_cache_bytes = 0;
}

hipError_t dmalloc(void **p, size_t n)
{
    auto it = g_cache.find(n);
    hipError_t e = hipSuccess;
    if (it != g_cache.end()) {
        *p = it->second;
        g_cache.erase(it);
        g_cache_bytes -= n;
    } else {
        e = hipMalloc(p, n);
        if (e == hipErrorOutOfMemory && !g_cache.empty()) {
            (void)hipGetLastError();
            cache_release_all();
            e = hipMalloc(p, n);
        }
    }
    if (e == hipSuccess) {
        g_cur->allocs[*p] = n;
        g_cur->dev_bytes += (int64_t)n;
    } else {
        *p = nullptr;
    }
    return e;
}

void dfree(void *p)
{
    if (!p) return;
    auto it = g_cur->allocs.find(p);
    if (it == g_cur->allocs.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t n = it->second;
    g_cur->dev_bytes -= (int64_t)n;
    g_cur->allocs.erase(it);
    if (g_cache_bytes + n <= kCacheCap) {
        g_cache.emplace(n, p);
        g_cache_bytes += n;
    } else {
        (void)hipFree(p);
    }
}

int64_t cached_bytes() { return (int64_t)g_cache_bytes; }
void release_cached() { cache_release_all(); }

// Pinned host buffer of small synchronous reads (shared: one host thread
// drives the library).
static double *g_hred = nullptr;

int ctx_count() { return (int)g_ctx.size(); }
State *ctx_get(int id) { return id >= 0 && id < (int)g_ctx.size() ? g_ctx[id] : nullptr; }
int ctx_add(State *s)
{
    g_ctx.push_back(s);
    return (int)g_ctx.size() - 1;
}
void ctx_set_slot(int id, State *s) { g_ctx[id] = s; }
void ctx_select(State *s) { g_cur = s; }

static int fail(int code, const std::string &msg)
{
    st().last_error = msg;
    return code;
}

struct LevelTab {
    int64_t key_off[64];
    int32_t nx[64];
};

__global__ void k_hash_insert(const int32_t *__restrict__ ckey, const int32_t *__restrict__ lev, int head, int n,
                              int ndim, LevelTab lt, int nlev, uint64_t *keys, int32_t *vals, uint64_t cap,
                              int *overflow)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = head - 1 + i;
    const int l = lev[o];
    if (l < 1 || l > nlev) {
        atomicOr(overflow, 2);
        return;
    }
    const uint64_t key = oct_key(lt.key_off[l - 1], lt.nx[l - 1], ndim, ckey[3 * o], ckey[3 * o + 1], ckey[3 * o + 2]);
    uint64_t h = hmix(key) & (cap - 1);
    for (uint64_t p = 0; p < cap; p++) {
        unsigned long long prev = atomicCAS((unsigned long long *)&keys[h], (unsigned long long)kEmpty,
                                            (unsigned long long)key);
        if (prev == kEmpty || prev == key) {
            vals[h] = o + 1;
            return;
        }
        h = (h + 1) & (cap - 1);
    }
    atomicOr(overflow, 1);
}

// 3^ndim neighbour slots of each oct of the level (slot 1+i+3j+9k, 0-based here).
__global__ void k_build_nbor(const int32_t *__restrict__ ckey, int head, int noct, ResolveCtx c, int nslot,
                             int32_t *__restrict__ nbor)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)noct * nslot) return;
    const int o = head - 1 + (int)(i / nslot);
    const int s = (int)(i % nslot);
    const int di = s % 3 - 1, dj = (s / 3) % 3 - 1, dk = s / 9 - 1;
    nbor[i] = resolve_oct(c, ckey[3 * o] + di, ckey[3 * o + 1] + dj, ckey[3 * o + 2] + dk);
}

// Brick windows: (bx+2)(by+2)(bz+2) slots per brick, slot (ox,oy,oz) = oct at
// org + (ox-1, oy-1, oz-1).
__global__ void k_build_bricks(const int32_t *__restrict__ org, int nbricks, int bx, int by, int bz, ResolveCtx c,
                               int32_t *__restrict__ win)
{
    const int wx = bx + 2, wy = by + 2, wz = bz + 2, nw = wx * wy * wz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nbricks * nw) return;
    const int b = (int)(i / nw), s = (int)(i % nw);
    const int ox = s % wx, oy = (s / wx) % wy, oz = s / (wx * wy);
    win[i] = resolve_oct(c, org[3 * b] + ox - 1, org[3 * b + 1] + oy - 1, org[3 * b + 2] + oz - 1);
}

// Count the in-domain neighbour slots of the level that no oct occupies.
__global__ void k_count_missing(const int32_t *__restrict__ nbor, long n, int *counter)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && nbor[i] == 0) atomicAdd(counter, 1);
}

// For every missing neighbour slot, insert its (wrapped) key into the device
// hash once and give it a ghost oct index in the cache region
// ngridmax+1..ngridmax+ncachemax (the role of make_cache_octs /
// insert_hash_cache, gpu/gpu_refine.cuf:937,2231).  counter[0] counts ghosts,
// counter[1] flags overflow.
__global__ void k_make_ghosts(const int32_t *__restrict__ ckey, const int32_t *__restrict__ nbor, int head, int noct,
                              ResolveCtx c, int nslot, int ilevel, int ngridmax, int ghost_base, int ghost_cap,
                              uint64_t *keys, int32_t *vals, int *counter, int32_t *gkey, int32_t *g_ckey,
                              int32_t *g_lev, uint8_t *g_ref)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)noct * nslot || nbor[i] != 0) return;
    const int o = head - 1 + (int)(i / nslot);
    const int s = (int)(i % nslot);
    int k[3] = {ckey[3 * o] + s % 3 - 1, ckey[3 * o + 1] + (s / 3) % 3 - 1, ckey[3 * o + 2] + s / 9 - 1};
    int nout = 0;
    for (int d = 0; d < c.ndim; d++) {
        if (c.periodic[d]) {
            const int n = c.boxmax[d] - c.boxmin[d];
            int r = (k[d] - c.boxmin[d]) % n;
            if (r < 0) r += n;
            k[d] = c.boxmin[d] + r;
        }
        nout += k[d] < c.boxmin[d] || k[d] >= c.boxmax[d];
    }
    // outside the box in one dimension (the boundary reference oct is absent,
    // or the slot would not be empty): interpolated from the parent level's
    // boundary octs, as the reference's get_grid leaves it to get_parent_cell
    // (amr/nbors_utils.f90:580-640, hydro/godunov_fine.f90:480-560)
    if (nout > 1 || (nout == 1 && c.ndim < 3)) return;
    const uint64_t key = nout ? ext_key(c.nx, k[0], k[1], k[2]) : oct_key(c.koff, c.nx, c.ndim, k[0], k[1], k[2]);
    uint64_t h = hmix(key) & (c.cap - 1);
    for (uint64_t p = 0; p < c.cap; p++) {
        unsigned long long prev = atomicCAS((unsigned long long *)&keys[h], (unsigned long long)kEmpty,
                                            (unsigned long long)key);
        if (prev == key) return;  // another slot created it
        if (prev == kEmpty) {
            const int gi = atomicAdd(&counter[0], 1);
            if (ghost_base + gi >= ghost_cap) {
                atomicOr(&counter[1], 1);
                vals[h] = 0;
                return;
            }
            const int idx = ngridmax + 1 + ghost_base + gi;  // 1-based oct index
            vals[h] = idx;
            gkey[3 * (ghost_base + gi) + 0] = k[0];
            gkey[3 * (ghost_base + gi) + 1] = k[1];
            gkey[3 * (ghost_base + gi) + 2] = k[2];
            g_ckey[3 * (idx - 1) + 0] = k[0];
            g_ckey[3 * (idx - 1) + 1] = k[1];
            g_ckey[3 * (idx - 1) + 2] = k[2];
            g_lev[idx - 1] = ilevel;
            for (int q = 0; q < 8; q++) g_ref[8 * (size_t)(idx - 1) + q] = 0;
            return;
        }
        h = (h + 1) & (c.cap - 1);
    }
    atomicOr(&counter[1], 2);
}

// Parent cell (get_parent_cell, amr/nbors_utils.f90:222-249) and the 2*ndim
// neighbouring parent cells (get_twondim_nbor_parent_cell, :149-217) of each
// new ghost: the stencil interpol_hydro reads.  counter[2] flags a missing
// parent ("GODUNOV: parent_cell should exist", hydro/godunov_fine.f90:491).
// gsrc receives the level's new ghosts only ([n][7][2], ghost first_gi first);
// k_order_sources rearranges it for the interpolation.
__global__ void k_ghost_sources(const int32_t *__restrict__ gkey, int first_gi, int n, ResolveCtx fine,
                                ResolveCtx coarse, int32_t *gparent, int32_t *gsrc, int32_t *gsrckey, int *counter)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int gi = first_gi + j;
    const int nd = fine.ndim;
    int k[3] = {gkey[3 * gi], gkey[3 * gi + 1], gkey[3 * gi + 2]};
    for (int slot = 0; slot <= 2 * nd; slot++) {
        int kn[3] = {k[0], k[1], k[2]};
        if (slot > 0) {
            const int d = (slot - 1) / 2;
            kn[d] += (slot % 2 == 1) ? -1 : 1;
            if (fine.periodic[d]) {
                const int nn = fine.boxmax[d] - fine.boxmin[d];
                int r = (kn[d] - fine.boxmin[d]) % nn;
                if (r < 0) r += nn;
                kn[d] = fine.boxmin[d] + r;
            }
        }
        // parent key and cell bit by floor division: the reference's integer
        // division for its non-negative keys, and the boundary oct -1 (cell
        // bit 1) for the key -1 next to a non-periodic box that starts at 0
        int kf[3] = {0, 0, 0}, ind = 1;
        for (int d = 0; d < nd; d++) {
            kf[d] = kn[d] >> 1;
            ind += (1 << d) * (kn[d] & 1);
        }
        const int code = resolve_oct(coarse, kf[0], kf[1], kf[2]);
        gsrc[(j * 7 + slot) * 2] = code;
        gsrc[(j * 7 + slot) * 2 + 1] = ind;
        gsrckey[(gi * 7 + slot) * 3] = kf[0];
        gsrckey[(gi * 7 + slot) * 3 + 1] = kf[1];
        gsrckey[(gi * 7 + slot) * 3 + 2] = kf[2];
        if (slot == 0) {
            // a negative code is a boundary oct of the coarse level (a ghost
            // beyond a non-periodic face): its refluxes are dropped
            gparent[2 * gi] = code;
            gparent[2 * gi + 1] = ind;
            if (code == 0) atomicOr(&counter[2], 1);
        } else if (code == 0) {
            atomicOr(&counter[2], 2);
        }
    }
}

// The ghost sources of one level rearranged for the interpolation, which walks
// the ghosts in parent-cell order: word e (slot e/2, its oct code or its cell)
// of the ghost at position base+i of that order goes to src[e*ld + base+i], so
// that the lanes of a wave, which take consecutive positions, read
// consecutive words.
__global__ void k_order_sources(const int32_t *__restrict__ tmp, const int32_t *__restrict__ order, int base, int n,
                                int ld, int32_t *__restrict__ src)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = order[base + i] - base;
    for (int e = 0; e < 14; e++) src[(size_t)e * ld + base + i] = tmp[(size_t)g * 14 + e];
}

// father(oct) = parent cell of each oct of the level (update_father_array,
// gpu/gpu_refine.cuf:1964).
__global__ void k_build_father(const int32_t *__restrict__ ckey, int head, int noct, ResolveCtx coarse, int ndim,
                               int32_t *father)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noct) return;
    const int o = head - 1 + i;
    int kf[3] = {0, 0, 0}, ind = 1;
    for (int d = 0; d < ndim; d++) {
        kf[d] = ckey[3 * o + d] >> 1;
        ind += (1 << d) * (ckey[3 * o + d] & 1);
    }
    const int code = resolve_oct(coarse, kf[0], kf[1], kf[2]);
    father[2 * i] = code > 0 ? code : 0;
    father[2 * i + 1] = ind;
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static int ensure_init()
{
    if (!st().inited) return fail(RH_ERR_STATE, "rh_init has not been called");
    return 0;
}

static int ensure_mesh()
{
    if (int r = ensure_init()) return r;
    if (!st().uold) return fail(RH_ERR_STATE, "rh_alloc_amr has not been called");
    return 0;
}

static void free_level_tables()
{
    for (auto &kv : st().bricks) {
        dfree(kv.second.win);
        dfree(kv.second.org);
        dfree(kv.second.zr);
    }
    st().bricks.clear();
    for (auto &kv : st().nbors) {
        dfree(kv.second.nbor);
        dfree(kv.second.father);
    }
    st().nbors.clear();
}

static int level_of_range(int head_idx, int num_octs, int &ilevel)
{
    State &s = st();
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridtot)
        return fail(RH_ERR_ARG, "oct range out of bounds");
    if (num_octs == 0) {
        ilevel = 0;
        return 0;
    }
    ilevel = s.h_lev[head_idx - 1];
    if (ilevel < 1 || ilevel > s.nlevelmax) return fail(RH_ERR_MESH, "oct level outside the level parameters");
    return 0;
}

static ResolveCtx make_resolve(int ilevel, const int32_t box_min[3], const int32_t box_max[3],
                               const int32_t periodic[3])
{
    State &s = st();
    ResolveCtx c;
    c.keys = s.hkeys;
    c.vals = s.hvals;
    c.cap = (uint64_t)s.hash_cap;
    c.koff = s.key_off[ilevel - 1];
    c.nx = s.ckey_max[ilevel - 1];
    c.ndim = s.ndim;
    c.ngridmax = s.ngridmax;
    for (int d = 0; d < 3; d++) {
        c.boxmin[d] = d < s.ndim ? box_min[d] : 0;
        c.boxmax[d] = d < s.ndim ? box_max[d] : 1;
        c.periodic[d] = d < s.ndim ? periodic[d] : 1;
    }
    return c;
}

// Lattice offset of a level's bricks.  The octs of a refined region come in
// the 2x2x2 blocks of the parent octs they refine, so where a level is sparse
// the number of bricks that cover it, and the planes they march, depend on
// where the brick lattice starts.  Every offset in [0,B) per dimension that
// keeps all origins inside the box (no oct below boxmin + offset) is scored
// by the modelled work of its bricks: per brick, the P phase converts
// (2BX+4)(2BY+4) cells and the T/F/U phases run (2BX+2)(2BY+2) columns on
// each marched plane (the occupied z layers plus the halo planes), a P cell
// costing about a third of a column-plane.  Levels filling at least half
// their box keep offset 0; large searches try only the aligned and
// half-shifted lattices.  Returns the modelled work of the chosen lattice.
static double choose_brick_offset(const int32_t *key, int n, const int32_t boxmin[3], const int32_t boxmax[3],
                                  const int B[3], int off[3])
{
    off[0] = off[1] = off[2] = 0;
    if (n <= 0) return 0.0;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], key[3 * i + d]);
            hi[d] = std::max(hi[d], key[3 * i + d]);
        }
    double vol = 1;
    for (int d = 0; d < 3; d++) vol *= (double)boxmax[d] - (double)boxmin[d];
    const bool dense = 2.0 * n > vol;
    std::vector<int> cand[3];
    double ncand = 1;
    for (int d = 0; d < 3; d++) {
        for (int o = 0; o < (dense ? 1 : B[d]); o++)
            if (o == 0 || o <= lo[d] - boxmin[d]) cand[d].push_back(o);
        ncand *= (double)cand[d].size();
    }
    if (ncand * n > 6e7)
        for (int d = 0; d < 3; d++) {
            std::vector<int> c2{0};
            if (B[d] > 1 && B[d] / 2 <= lo[d] - boxmin[d]) c2.push_back(B[d] / 2);
            cand[d] = c2;
        }
    std::vector<int32_t> occ;
    double best = -1;
    for (int oz : cand[2])
        for (int oy : cand[1])
            for (int ox : cand[0]) {
                const int o[3] = {ox, oy, oz};
                int g0[3], ng[3];
                for (int d = 0; d < 3; d++) {
                    g0[d] = (lo[d] - boxmin[d] - o[d]) / B[d];
                    ng[d] = (hi[d] - boxmin[d] - o[d]) / B[d] - g0[d] + 1;
                }
                occ.assign((size_t)ng[0] * ng[1] * ng[2], -1);
                for (int i = 0; i < n; i++) {
                    int b[3];
                    for (int d = 0; d < 3; d++) b[d] = (key[3 * i + d] - boxmin[d] - o[d]) / B[d] - g0[d];
                    const int z = (key[3 * i + 2] - boxmin[2] - o[2]) % B[2];
                    int32_t &v = occ[((size_t)b[2] * ng[1] + b[1]) * ng[0] + b[0]];
                    v = v < 0 ? (z | z << 16) : (std::min(v & 0xffff, z) | std::max(v >> 16, z) << 16);
                }
                double cost = 0;
                for (int32_t v : occ)
                    if (v >= 0) {
                        cost += brick_work(B[0], B[1], 2.0 * ((v >> 16) - (v & 0xffff) + 1));
                    }
                if (best < 0 || cost < best) {
                    best = cost;
                    off[0] = ox, off[1] = oy, off[2] = oz;
                }
            }
    return best;
}

}  // namespace rh

using namespace rh;

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
extern "C" {

const char *rh_last_error(void) { return st().last_error.c_str(); }

// Scratch of rh_cmpdt_noalloc: Courant partials (kNoAllocParts x 4 doubles)
// and the 5 results, one copy per GPU.  Host code uses the per-device
// address from hipGetSymbolAddress (g_scratch_dev), never the symbol itself.
static constexpr int kNoAllocParts = 4096;
static __device__ __align__(256) double scratch[kNoAllocParts * 4 + 8];
static double *g_scratch_dev[64];

// The device address of `scratch` on the current device, fetched once per
// device (with that device current); null and an error message on failure.
static double *scratch_ptr()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        set_error("rh_cmpdt_noalloc: no current HIP device (or device id >= 64)");
        return nullptr;
    }
    if (!g_scratch_dev[dev]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("rh_cmpdt_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

int rh_init(void)
{
    State &s = st();
    if (s.inited) return 0;
    int n = 0;
    RH_HIP_CHECK(hipGetDeviceCount(&n));
    if (n < 1) return fail(RH_ERR_HIP, "no HIP device visible");
    RH_HIP_CHECK(hipGetDevice(&s.device));
    if (!g_stream) RH_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    s.stream = g_stream;
    RH_HIP_CHECK(dmalloc((void **)&s.red_out, 8 * sizeof(double)));
    if (!g_hred) RH_HIP_CHECK(hipHostMalloc(&g_hred, 8 * sizeof(double), hipHostMallocDefault));
    s.h_red = g_hred;
    s.red_cap = 4096;
    RH_HIP_CHECK(dmalloc((void **)&s.red_part, (size_t)s.red_cap * 4 * sizeof(double)));
    RH_HIP_CHECK(dmalloc((void **)&s.dt_dev, 2 * sizeof(double)));
    RH_HIP_CHECK(dmalloc((void **)&s.d_counter, 4 * sizeof(int)));
    RH_HIP_CHECK(dmalloc((void **)&s.t_dev, 2 * sizeof(double)));
    RH_HIP_CHECK(hipEventCreate(&s.ev0));
    RH_HIP_CHECK(hipEventCreate(&s.ev1));
    // the no-allocation Courant entry point's scratch address on this device
    if (!scratch_ptr()) return RH_ERR_HIP;
    s.inited = true;
    s.last_error.clear();
    return 0;
}

int rh_finalize(void)
{
    State &s = st();
    if (!s.inited) return 0;
    drop_level_graph(s);
    // no device synchronisation: released blocks go to the allocation cache
    // and are reused only by work ordered after this context's on the stream
    free_level_tables();
    dfree(s.uold);
    dfree(s.unew);
    dfree(s.ckey);
    dfree(s.lev);
    dfree(s.refined);
    dfree(s.flag1);
    s.flag1 = nullptr;
    dfree(s.racc_det);
    s.racc_det = nullptr;
    s.racc_det_cap = 0;
    dfree(s.hkeys);
    dfree(s.hvals);
    dfree(s.ghost_parent);
    dfree(s.ghost_src);
    dfree(s.ghost_srckey);
    dfree(s.ghost_order);
    dfree(s.ghost_key);
    dfree(s.d_counter);
    dfree(s.red_part);
    dfree(s.red_out);
    dfree(s.dt_dev);
    dfree(s.t_dev);
    (void)hipEventDestroy(s.ev0);
    (void)hipEventDestroy(s.ev1);
    // anything else the context still holds (the shared stream stays)
    while (!s.allocs.empty()) dfree(s.allocs.begin()->first);
    s = State();
    return 0;
}

int rh_alloc_amr(int ngridmax, int ncachemax, int nvar, int twotondim, int hash_size)
{
    if (int r = ensure_init()) return r;
    State &s = st();
    if (nvar != 5) return fail(RH_ERR_ARG, "nvar must be 5 (hydro, no passive scalars)");
    if (twotondim != 2 && twotondim != 8) return fail(RH_ERR_ARG, "twotondim must be 2 (1D) or 8 (3D)");
    if (ngridmax < 1 || ncachemax < 0 || (int64_t)ngridmax + ncachemax > (1 << 29))
        return fail(RH_ERR_ARG, "ngridmax+ncachemax must be in [1, 2^29]");
    if (hash_size < 1) return fail(RH_ERR_ARG, "hash_size must be positive");
    (void)hipStreamSynchronize(s.stream);
    drop_level_graph(s);
    free_level_tables();
    dfree(s.uold);
    dfree(s.unew);
    dfree(s.ckey);
    dfree(s.lev);
    dfree(s.refined);
    dfree(s.flag1);
    s.flag1 = nullptr;
    dfree(s.racc_det);
    s.racc_det = nullptr;
    s.racc_det_cap = 0;
    dfree(s.hkeys);
    dfree(s.hvals);
    s.ndim = twotondim == 8 ? 3 : 1;
    s.ttd = twotondim;
    s.nvar = nvar;
    s.ngridmax = ngridmax;
    s.ncachemax = ncachemax;
    s.ngridtot = ngridmax + ncachemax;
    int64_t cap = 1;
    while (cap < hash_size) cap <<= 1;
    s.hash_cap = cap;
    const size_t nd = (size_t)s.ngridtot * nvar * twotondim;
    if (dmalloc((void **)&s.uold, nd * sizeof(double)) != hipSuccess || dmalloc((void **)&s.unew, nd * sizeof(double)) != hipSuccess ||
        dmalloc((void **)&s.ckey, (size_t)s.ngridtot * 3 * sizeof(int32_t)) != hipSuccess ||
        dmalloc((void **)&s.lev, (size_t)s.ngridtot * sizeof(int32_t)) != hipSuccess ||
        dmalloc((void **)&s.refined, (size_t)s.ngridtot * 8) != hipSuccess ||
        dmalloc((void **)&s.hkeys, (size_t)cap * sizeof(uint64_t)) != hipSuccess ||
        dmalloc((void **)&s.hvals, (size_t)cap * sizeof(int32_t)) != hipSuccess)
        return fail(RH_ERR_NOMEM, "device allocation failed in rh_alloc_amr");
    RH_HIP_CHECK(hipMemsetAsync(s.uold, 0, nd * sizeof(double), s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.unew, 0, nd * sizeof(double), s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.ckey, 0, (size_t)s.ngridtot * 3 * sizeof(int32_t), s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.lev, 0, (size_t)s.ngridtot * sizeof(int32_t), s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.refined, 0, (size_t)s.ngridtot * 8, s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.hkeys, 0xFF, (size_t)cap * sizeof(uint64_t), s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.hvals, 0, (size_t)cap * sizeof(int32_t), s.stream));
    {
        const size_t ng = (size_t)(ncachemax > 0 ? ncachemax : 1);
        dfree(s.ghost_parent);
        dfree(s.ghost_src);
        dfree(s.ghost_srckey);
        dfree(s.ghost_order);
        dfree(s.ghost_key);
        if (dmalloc((void **)&s.ghost_parent, ng * 2 * sizeof(int32_t)) != hipSuccess ||
            dmalloc((void **)&s.ghost_src, ng * 14 * sizeof(int32_t)) != hipSuccess ||
            dmalloc((void **)&s.ghost_srckey, ng * 21 * sizeof(int32_t)) != hipSuccess ||
            dmalloc((void **)&s.ghost_order, ng * sizeof(int32_t)) != hipSuccess ||
            dmalloc((void **)&s.ghost_key, ng * 3 * sizeof(int32_t)) != hipSuccess)
            return fail(RH_ERR_NOMEM, "device allocation failed for the ghost-oct tables");
        s.nghost_total = 0;
    }
    s.h_ckey.assign((size_t)s.ngridtot * 3, 0);
    s.h_lev.assign((size_t)s.ngridtot, 0);
    s.h_refany.clear();
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

int rh_upload_level_params(const int32_t *ckey_max, const int64_t *key_off, const int32_t *box_ckey_min,
                           const int32_t *box_ckey_max, const int32_t periodic[3], int nlevelmax)
{
    if (int r = ensure_init()) return r;
    State &s = st();
    if (nlevelmax < 1 || nlevelmax > 63) return fail(RH_ERR_ARG, "nlevelmax must be in [1, 63]");
    if (!ckey_max || !key_off || !box_ckey_min || !box_ckey_max || !periodic)
        return fail(RH_ERR_ARG, "null level parameter array");
    s.nlevelmax = nlevelmax;
    s.ckey_max.assign(ckey_max, ckey_max + nlevelmax + 1);
    s.key_off.assign(key_off, key_off + nlevelmax + 1);
    s.box_min.assign(box_ckey_min, box_ckey_min + 3 * (nlevelmax + 1));
    s.box_max.assign(box_ckey_max, box_ckey_max + 3 * (nlevelmax + 1));
    for (int d = 0; d < 3; d++) s.periodic[d] = periodic[d] ? 1 : 0;
    return 0;
}

int rh_set_boundaries(const int32_t bound_type[6])
{
    if (!bound_type) return fail(RH_ERR_ARG, "null bound_type");
    for (int d = 0; d < 3; d++)
        for (int side = 0; side < 2; side++) {
            const int t = bound_type[2 * d + side];
            if (t != 1 && t != 2) return fail(RH_ERR_ARG, "bound_type must be 1 (reflexive) or 2 (zero gradient)");
            st().btype[d][side] = t;
        }
    return 0;
}

int rh_set_grid_device(const double *uold, const double *unew, const void *bold, const void *grid, int ngridmax,
                       int nvar, int twotondim)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (bold) return fail(RH_ERR_ARG, "bold must be NULL (no MHD)");
    if (ngridmax < 0 || ngridmax > s.ngridtot || nvar != s.nvar || twotondim != s.ttd)
        return fail(RH_ERR_ARG, "rh_set_grid_device: sizes differ from rh_alloc_amr");
    const size_t per = (size_t)nvar * twotondim;
    if (uold) RH_HIP_CHECK(hipMemcpyAsync(s.uold, uold, per * ngridmax * sizeof(double), hipMemcpyHostToDevice, s.stream));
    if (unew) RH_HIP_CHECK(hipMemcpyAsync(s.unew, unew, per * ngridmax * sizeof(double), hipMemcpyHostToDevice, s.stream));
    if (grid) {
        // unpack the reference oct records into SoA arrays
        const size_t rec = s.ndim == 3 ? 32 : 24;
        const size_t off_ckey = 8, off_ref = s.ndim == 3 ? 20 : 12, off_lev = s.ndim == 3 ? 28 : 16;
        std::vector<uint8_t> refb((size_t)ngridmax * 8, 0);
        s.h_refany.assign((size_t)ngridmax, 0);
        const uint8_t *g = (const uint8_t *)grid;
        for (int i = 0; i < ngridmax; i++) {
            const uint8_t *r = g + rec * i;
            int32_t ck[3] = {0, 0, 0};
            memcpy(ck, r + off_ckey, sizeof(int32_t) * s.ndim);
            memcpy(&s.h_ckey[3 * (size_t)i], ck, sizeof(ck));
            memcpy(&s.h_lev[i], r + off_lev, sizeof(int32_t));
            memcpy(&refb[8 * (size_t)i], r + off_ref, s.ttd);
            for (int c = 0; c < s.ttd; c++) s.h_refany[i] |= refb[8 * (size_t)i + c] != 0;
        }
        RH_HIP_CHECK(hipMemcpyAsync(s.ckey, s.h_ckey.data(), (size_t)ngridmax * 3 * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
        RH_HIP_CHECK(hipMemcpyAsync(s.lev, s.h_lev.data(), (size_t)ngridmax * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
        RH_HIP_CHECK(hipMemcpyAsync(s.refined, refb.data(), (size_t)ngridmax * 8, hipMemcpyHostToDevice, s.stream));
        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
        // mesh changed: tables and ghost octs must be rebuilt
        free_level_tables();
        s.nghost_total = 0;
        RH_HIP_CHECK(hipMemsetAsync(s.hkeys, 0xFF, (size_t)s.hash_cap * sizeof(uint64_t), s.stream));
    }
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

int rh_transfer_grid_host(double *uold, void *bold, int ngridmax, int nvar, int twotondim)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    (void)bold;
    if (!uold || ngridmax < 0 || ngridmax > s.ngridtot || nvar != s.nvar || twotondim != s.ttd)
        return fail(RH_ERR_ARG, "rh_transfer_grid_host: bad arguments");
    RH_HIP_CHECK(hipMemcpyAsync(uold, s.uold, (size_t)nvar * twotondim * ngridmax * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

int rh_transfer_unew_host(double *unew, int ngridmax, int nvar, int twotondim)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (!unew || ngridmax < 0 || ngridmax > s.ngridtot || nvar != s.nvar || twotondim != s.ttd)
        return fail(RH_ERR_ARG, "rh_transfer_unew_host: bad arguments");
    RH_HIP_CHECK(hipMemcpyAsync(unew, s.unew, (size_t)nvar * twotondim * ngridmax * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

int rh_insert_hash_all(int head_idx, int num_octs, int hash_size)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    (void)hash_size;
    if (s.nlevelmax < 1) return fail(RH_ERR_STATE, "rh_upload_level_params has not been called");
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridtot)
        return fail(RH_ERR_ARG, "rh_insert_hash_all: oct range out of bounds");
    if (num_octs == 0) return 0;
    LevelTab lt;
    for (int l = 0; l < 64; l++) {
        lt.key_off[l] = l < s.nlevelmax ? s.key_off[l] : 0;
        lt.nx[l] = l < s.nlevelmax ? s.ckey_max[l] : 0;
    }
    int *d_over = nullptr;
    RH_HIP_CHECK(hipMalloc(&d_over, sizeof(int)));
    RH_HIP_CHECK(hipMemsetAsync(d_over, 0, sizeof(int), s.stream));
    hipLaunchKernelGGL(k_hash_insert, dim3((num_octs + 255) / 256), dim3(256), 0, s.stream, s.ckey, s.lev, head_idx,
                       num_octs, s.ndim, lt, s.nlevelmax, s.hkeys, s.hvals, (uint64_t)s.hash_cap, d_over);
    RH_HIP_CHECK(hipGetLastError());
    int over = 0;
    RH_HIP_CHECK(hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    (void)hipFree(d_over);
    if (over & 1) return fail(RH_ERR_MESH, "device hash table full");
    if (over & 2) return fail(RH_ERR_MESH, "oct with a level outside [1, nlevelmax]");
    return 0;
}

int rh_build_nbor(int head_idx, int num_octs, int hash_size, int ckey_max_l, int64_t key_off_l,
                  const int32_t box_min[3], const int32_t box_max[3], const int32_t periodic[3])
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    (void)hash_size;
    (void)ckey_max_l;
    (void)key_off_l;
    if (!box_min || !box_max || !periodic) return fail(RH_ERR_ARG, "null box arrays");
    drop_level_graph(s);  // its tables are rebuilt
    if (s.ndim == 3 && !periodic[0] + !periodic[1] + !periodic[2] > 1)
        // stencil octs beyond an edge or a corner of the box: the reference
        // builds them from whichever neighbouring boundary oct its cache holds
        // at that moment (amr/nbors_utils.f90:580-640), so no result is defined
        return fail(RH_ERR_ARG, "3D boxes with more than one non-periodic dimension are not supported");
    int ilevel = 0;
    if (int r = level_of_range(head_idx, num_octs, ilevel)) return r;
    if (num_octs == 0) return 0;
    for (int i = head_idx - 1; i < head_idx - 1 + num_octs; i++)
        if (s.h_lev[i] != ilevel) return fail(RH_ERR_MESH, "octs of the range are not all on one level");
    ResolveCtx c = make_resolve(ilevel, box_min, box_max, periodic);

    // per-oct neighbour table (3^ndim slots), father table and ghost octs
    {
        int prev_first = 0, prev_n = 0;
        auto it = s.nbors.find(ilevel);
        if (it != s.nbors.end()) {
            prev_first = it->second.ghost_first;
            prev_n = it->second.nghost;
            dfree(it->second.nbor);
            dfree(it->second.father);
            s.nbors.erase(it);
        }
        LevelNbor ln;
        ln.ilevel = ilevel;
        ln.head = head_idx;
        ln.noct = num_octs;
        ln.ghost_first = prev_first;
        ln.nghost = prev_n;
        const int nslot = s.ndim == 3 ? 27 : 3;
        RH_HIP_CHECK(dmalloc((void **)&ln.nbor, (size_t)num_octs * nslot * sizeof(int32_t)));
        const long n = (long)num_octs * nslot;
        const unsigned nblk = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_build_nbor, dim3(nblk), dim3(256), 0, s.stream, s.ckey, head_idx, num_octs, c, nslot,
                           ln.nbor);
        RH_HIP_CHECK(hipGetLastError());
        for (int d = 0; d < 3; d++) {
            ln.boxmin[d] = c.boxmin[d];
            ln.boxmax[d] = c.boxmax[d];
        }
        if (ilevel > 1) {
            ResolveCtx cc = make_resolve(ilevel - 1, &s.box_min[3 * (ilevel - 2)], &s.box_max[3 * (ilevel - 2)],
                                         periodic);
            RH_HIP_CHECK(dmalloc((void **)&ln.father, (size_t)num_octs * 2 * sizeof(int32_t)));
            hipLaunchKernelGGL(k_build_father, dim3((num_octs + 255) / 256), dim3(256), 0, s.stream, s.ckey,
                               head_idx, num_octs, cc, s.ndim, ln.father);
            RH_HIP_CHECK(hipGetLastError());
            int hc[4] = {0, 0, 0, 0};
            RH_HIP_CHECK(hipMemsetAsync(s.d_counter, 0, 4 * sizeof(int), s.stream));
            hipLaunchKernelGGL(k_count_missing, dim3(nblk), dim3(256), 0, s.stream, ln.nbor, n, s.d_counter);
            RH_HIP_CHECK(hipGetLastError());
            RH_HIP_CHECK(hipMemcpyAsync(hc, s.d_counter, 4 * sizeof(int), hipMemcpyDeviceToHost, s.stream));
            RH_HIP_CHECK(hipStreamSynchronize(s.stream));
            if (hc[0] > 0) {  // (the edges of a 1D box count here and make no ghost)
                const int base = s.nghost_total;
                RH_HIP_CHECK(hipMemsetAsync(s.d_counter, 0, 4 * sizeof(int), s.stream));
                hipLaunchKernelGGL(k_make_ghosts, dim3(nblk), dim3(256), 0, s.stream, s.ckey, ln.nbor, head_idx,
                                   num_octs, c, nslot, ilevel, s.ngridmax, base, s.ncachemax, s.hkeys, s.hvals,
                                   s.d_counter, s.ghost_key, s.ckey, s.lev, s.refined);
                RH_HIP_CHECK(hipGetLastError());
                RH_HIP_CHECK(hipMemcpyAsync(hc, s.d_counter, 4 * sizeof(int), hipMemcpyDeviceToHost, s.stream));
                RH_HIP_CHECK(hipStreamSynchronize(s.stream));
                if (hc[1]) {
                    s.nbors[ilevel] = ln;
                    return fail(RH_ERR_MESH, "ncachemax too small for the ghost octs of the level (or hash full)");
                }
                const int nnew = hc[0];
                if (nnew > 0) {
                    int32_t *tmp = nullptr;  // the new ghosts' sources in ghost order, until rearranged
                    RH_HIP_CHECK(dmalloc((void **)&tmp, (size_t)nnew * 14 * sizeof(int32_t)));
                    RH_HIP_CHECK(hipMemsetAsync(s.d_counter, 0, 4 * sizeof(int), s.stream));
                    hipLaunchKernelGGL(k_ghost_sources, dim3((nnew + 127) / 128), dim3(128), 0, s.stream, s.ghost_key,
                                       base, nnew, c, cc, s.ghost_parent, tmp, s.ghost_srckey, s.d_counter);
                    RH_HIP_CHECK(hipGetLastError());
                    RH_HIP_CHECK(hipMemcpyAsync(hc, s.d_counter, 4 * sizeof(int), hipMemcpyDeviceToHost, s.stream));
                    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
                    ln.ghost_first = s.ngridmax + 1 + base;
                    ln.nghost = nnew;
                    s.nghost_total += nnew;
                    {
                        // the level's ghosts in parent-cell order: the interpolation
                        // and the reflux fold walk them in this order, so that the
                        // ghosts of one wave (and of one XCD) share parent octs
                        std::vector<int32_t> par((size_t)nnew * 2), ord((size_t)nnew);
                        RH_HIP_CHECK(hipMemcpyAsync(par.data(), s.ghost_parent + 2 * (size_t)base,
                                                    par.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream));
                        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
                        for (int g = 0; g < nnew; g++) ord[g] = g;
                        std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
                            const int64_t kx = (int64_t)par[2 * x] * 8 + par[2 * x + 1];
                            const int64_t ky = (int64_t)par[2 * y] * 8 + par[2 * y + 1];
                            return kx < ky;
                        });
                        for (int g = 0; g < nnew; g++) ord[g] += base;
                        RH_HIP_CHECK(hipMemcpyAsync(s.ghost_order + base, ord.data(), ord.size() * sizeof(int32_t),
                                                    hipMemcpyHostToDevice, s.stream));
                        hipLaunchKernelGGL(k_order_sources, dim3((nnew + 127) / 128), dim3(128), 0, s.stream, tmp,
                                           s.ghost_order, base, nnew, ghost_ld(s), s.ghost_src);
                        RH_HIP_CHECK(hipGetLastError());
                        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
                        dfree(tmp);
                    }
                    for (int g = 0; g < nnew; g++) s.h_lev[s.ngridmax + base + g] = ilevel;
                    if (hc[2] & 1) {
                        s.nbors[ilevel] = ln;
                        return fail(RH_ERR_MESH, "GODUNOV: parent_cell should exist");
                    }
                    if (hc[2] & 2) {
                        s.nbors[ilevel] = ln;
                        return fail(RH_ERR_MESH, "GODUNOV: parent neighbors should exist");
                    }
                    // the ghosts now resolve through the hash
                    hipLaunchKernelGGL(k_build_nbor, dim3(nblk), dim3(256), 0, s.stream, s.ckey, head_idx, num_octs, c,
                                       nslot, ln.nbor);
                    RH_HIP_CHECK(hipGetLastError());
                }
            }
        }
        s.nbors[ilevel] = ln;
    }
    if (s.ndim == 3) {
        auto it = s.bricks.find(ilevel);
        if (it != s.bricks.end()) {
            dfree(it->second.win);
            dfree(it->second.org);
            dfree(it->second.zr);
            s.bricks.erase(it);
        }
        LevelBricks lb;
        lb.ilevel = ilevel;
        lb.head = head_idx;
        lb.noct = num_octs;
        lb.bx = s.brick_bx;
        lb.by = s.brick_by;
        lb.bz = s.brick_bz;
        lb.bnd = !(c.periodic[0] && c.periodic[1] && c.periodic[2]);
        // enumerate the bricks that hold at least one oct of the level, on
        // the lattice boxmin + off + multiples of the shape
        const int B[3] = {lb.bx, lb.by, lb.bz};
        choose_brick_offset(s.h_ckey.data() + 3 * (size_t)(head_idx - 1), num_octs, c.boxmin, c.boxmax, B, lb.off);
        const int32_t lmin[3] = {c.boxmin[0] + lb.off[0], c.boxmin[1] + lb.off[1], c.boxmin[2] + lb.off[2]};
        std::vector<int64_t> keys;
        const int64_t nbx = (c.boxmax[0] - lmin[0] + lb.bx - 1) / lb.bx;
        const int64_t nby = (c.boxmax[1] - lmin[1] + lb.by - 1) / lb.by;
        const int64_t nbz = (c.boxmax[2] - lmin[2] + lb.bz - 1) / lb.bz;
        auto brick_of = [&](size_t i) {
            const int64_t bx = (s.h_ckey[3 * i] - lmin[0]) / lb.bx;
            const int64_t by = (s.h_ckey[3 * i + 1] - lmin[1]) / lb.by;
            const int64_t bz = (s.h_ckey[3 * i + 2] - lmin[2]) / lb.bz;
            return (bz * nby + by) * nbx + bx;
        };
        std::vector<int32_t> zr;  // per brick: z layers holding octs, zlo | zhi << 16
        auto zlayer = [&](size_t i) { return (s.h_ckey[3 * i + 2] - lmin[2]) % lb.bz; };
        if (nbx * nby * nbz <= (int64_t(1) << 26)) {
            // occupancy map of the level's brick grid: linear in octs + bricks
            std::vector<int32_t> occ((size_t)(nbx * nby * nbz), -1);
            for (int i = head_idx - 1; i < head_idx - 1 + num_octs; i++) {
                int32_t &o = occ[(size_t)brick_of((size_t)i)];
                const int z = zlayer((size_t)i);
                o = o < 0 ? (z | z << 16) : (std::min(o & 0xffff, z) | std::max(o >> 16, z) << 16);
            }
            for (size_t k = 0; k < occ.size(); k++)
                if (occ[k] >= 0) {
                    keys.push_back((int64_t)k);
                    zr.push_back(occ[k]);
                }
        } else {
            keys.reserve(num_octs);
            for (int i = head_idx - 1; i < head_idx - 1 + num_octs; i++) keys.push_back(brick_of((size_t)i));
            std::sort(keys.begin(), keys.end());
            keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
            zr.assign(keys.size(), -1);
            for (int i = head_idx - 1; i < head_idx - 1 + num_octs; i++) {
                const size_t b = std::lower_bound(keys.begin(), keys.end(), brick_of((size_t)i)) - keys.begin();
                const int z = zlayer((size_t)i);
                int32_t &o = zr[b];
                o = o < 0 ? (z | z << 16) : (std::min(o & 0xffff, z) | std::max(o >> 16, z) << 16);
            }
        }
        lb.nbricks = (int)keys.size();
        // refined cells in the level (unknown before a grid upload: assume some)
        lb.any_ref = 1;
        if (s.h_refany.size() >= (size_t)(head_idx - 1 + num_octs)) {
            lb.any_ref = 0;
            for (int i = head_idx - 1; i < head_idx - 1 + num_octs && !lb.any_ref; i++) lb.any_ref = s.h_refany[i];
        }
        std::vector<int32_t> org(3 * keys.size());
        for (size_t b = 0; b < keys.size(); b++) {
            const int64_t k = keys[b];
            org[3 * b + 0] = lmin[0] + (int32_t)(k % nbx) * lb.bx;
            org[3 * b + 1] = lmin[1] + (int32_t)((k / nbx) % nby) * lb.by;
            org[3 * b + 2] = lmin[2] + (int32_t)(k / (nbx * nby)) * lb.bz;
        }
        const int nw = (lb.bx + 2) * (lb.by + 2) * (lb.bz + 2);
        RH_HIP_CHECK(dmalloc((void **)&lb.org, org.size() * sizeof(int32_t)));
        RH_HIP_CHECK(dmalloc((void **)&lb.zr, zr.size() * sizeof(int32_t)));
        RH_HIP_CHECK(dmalloc((void **)&lb.win, (size_t)lb.nbricks * nw * sizeof(int32_t)));
        RH_HIP_CHECK(hipMemcpyAsync(lb.org, org.data(), org.size() * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
        RH_HIP_CHECK(hipMemcpyAsync(lb.zr, zr.data(), zr.size() * sizeof(int32_t), hipMemcpyHostToDevice, s.stream));
        lb.h_zr = zr;
        const long n = (long)lb.nbricks * nw;
        hipLaunchKernelGGL(k_build_bricks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, lb.org,
                           lb.nbricks, lb.bx, lb.by, lb.bz, c, lb.win);
        RH_HIP_CHECK(hipGetLastError());
        for (int d = 0; d < 3; d++) {
            lb.boxmin[d] = c.boxmin[d];
            lb.boxmax[d] = c.boxmax[d];
        }
        s.bricks[ilevel] = lb;
        // every oct of the level must be found in its own brick window
        std::vector<int32_t> hwin((size_t)n);
        RH_HIP_CHECK(hipMemcpyAsync(hwin.data(), lb.win, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream));
        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
        long found = 0;
        const int wx = lb.bx + 2, wy = lb.by + 2;
        for (int b = 0; b < lb.nbricks; b++)
            for (int oz = 1; oz <= lb.bz; oz++)
                for (int oy = 1; oy <= lb.by; oy++)
                    for (int ox = 1; ox <= lb.bx; ox++) {
                        if (org[3 * b] + ox - 1 >= c.boxmax[0] || org[3 * b + 1] + oy - 1 >= c.boxmax[1] ||
                            org[3 * b + 2] + oz - 1 >= c.boxmax[2])
                            continue;
                        const int v = hwin[(size_t)b * nw + (oz * wy + oy) * wx + ox];
                        if (v >= head_idx && v < head_idx + num_octs) found++;
                    }
        if (found != num_octs) return fail(RH_ERR_MESH, "octs of the level missing from the device hash");
    }
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

int rh_set_unew(int head_idx, int num_octs)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridtot)
        return fail(RH_ERR_ARG, "rh_set_unew: oct range out of bounds");
    const size_t per = (size_t)s.nvar * s.ttd;
    RH_HIP_CHECK(hipMemcpyAsync(s.unew + per * (head_idx - 1), s.uold + per * (head_idx - 1),
                                per * num_octs * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    return 0;
}

int rh_set_uold(int head_idx, int num_octs)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridtot)
        return fail(RH_ERR_ARG, "rh_set_uold: oct range out of bounds");
    const size_t per = (size_t)s.nvar * s.ttd;
    RH_HIP_CHECK(hipMemcpyAsync(s.uold + per * (head_idx - 1), s.unew + per * (head_idx - 1),
                                per * num_octs * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    return 0;
}

int rh_cmpdt(int head_idx, int num_octs, double dx, double gamma, double smallr, double smallc2, double courant_factor,
             const double g[3], double *mass, double *ekin, double *eint, double *emag, double *dt)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridtot)
        return fail(RH_ERR_ARG, "rh_cmpdt: oct range out of bounds");
    if (!(dx > 0) || !(smallc2 > 0) || !(gamma > 1)) return fail(RH_ERR_ARG, "rh_cmpdt: bad physical parameters");
    const double gz[3] = {0, 0, 0};
    const double smallc = __builtin_sqrt(smallc2);
    RH_HIP_CHECK(launch_courant(s.uold, s.refined, head_idx, num_octs, s.ndim, dx, gamma, smallr, smallc,
                                courant_factor, g ? g : gz, s.red_part, s.red_cap, s.red_out, nullptr, s.stream));
    RH_HIP_CHECK(hipMemcpyAsync(s.h_red, s.red_out, 5 * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    if (mass) *mass = s.h_red[0];
    if (ekin) *ekin = s.h_red[1];
    if (eint) *eint = s.h_red[2];
    if (emag) *emag = s.h_red[3];
    if (dt) *dt = s.h_red[4];
    return 0;
}

int rh_cmpdt_noalloc(int head_idx, int num_octs, double dx, double gamma, double smallr, double smallc2,
                     double courant_factor, const double g[3], double *mass, double *ekin, double *eint,
                     double *emag, double *dt)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridtot)
        return fail(RH_ERR_ARG, "rh_cmpdt_noalloc: oct range out of bounds");
    if (!(dx > 0) || !(smallc2 > 0) || !(gamma > 1))
        return fail(RH_ERR_ARG, "rh_cmpdt_noalloc: bad physical parameters");
    double *sp = scratch_ptr();
    if (!sp) return RH_ERR_HIP;
    double *partials = sp, *out = sp + 4 * kNoAllocParts;
    const double gz[3] = {0, 0, 0};
    RH_HIP_CHECK(launch_courant(s.uold, s.refined, head_idx, num_octs, s.ndim, dx, gamma, smallr,
                                __builtin_sqrt(smallc2), courant_factor, g ? g : gz, partials, kNoAllocParts, out,
                                nullptr, s.stream));
    // results through the pinned host buffer allocated at rh_init
    RH_HIP_CHECK(hipMemcpyAsync(s.h_red, out, 5 * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    if (mass) *mass = s.h_red[0];
    if (ekin) *ekin = s.h_red[1];
    if (eint) *eint = s.h_red[2];
    if (emag) *emag = s.h_red[3];
    if (dt) *dt = s.h_red[4];
    return 0;
}

int rh_godunov(int head_idx, int num_subgrids, int ngridmax, int ilevel, int levelmin, int levelmax, double gamma,
               double smallr, double smallc2, double dt, double dx, double smax_cap, int slope, int riemann,
               const double g[3])
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    (void)ngridmax;
    (void)levelmax;
    if (num_subgrids == 0) return 0;
    if (riemann < 1 || riemann > 3) return fail(RH_ERR_ARG, "riemann must be 1 (llf), 2 (hll) or 3 (hllc)");
    if (slope < 0 || slope > 3) return fail(RH_ERR_ARG, "slope_type must be 0..3");
    if (!(dt > 0) || !(dx > 0)) return fail(RH_ERR_ARG, "dt and dx must be positive");
    const double gz[3] = {0, 0, 0};
    const double *gg = g ? g : gz;
    auto itn = s.nbors.find(ilevel);
    if (itn == s.nbors.end() || itn->second.head != head_idx || itn->second.noct != num_subgrids)
        return fail(RH_ERR_STATE, "rh_godunov: rh_build_nbor was not called for this level");
    const LevelNbor &lnb = itn->second;
    // ghost octs of the level take their values from the parent level now
    if (lnb.nghost > 0) {
        GhostArgs ga;
        ga.uold = s.uold;
        ga.refined = s.refined;
        const int gi0 = lnb.ghost_first - s.ngridmax - 1;
        ga.src = s.ghost_src + gi0;
        ga.src_ld = ghost_ld(s);
        ga.srckey = s.ghost_srckey + (size_t)gi0 * 21;
        ga.order = s.ghost_order + gi0;
        ga.gi0 = gi0;
        ga.first = lnb.ghost_first;
        ga.nghost = lnb.nghost;
        ga.ndim = s.ndim;
        ga.interpol_var = s.interpol_var;
        ga.interpol_type = s.interpol_type;
        ga.smallr = smallr;
        ga.nseg = 1;
        ga.seg_t[0] = 0;
        ga.seg_t[1] = lnb.nghost;
        ga.seg_o[0] = 0;
        for (int d = 0; d < 3; d++) {
            ga.seg_box[0][d] = d < s.ndim ? s.box_min[3 * (ilevel - 2) + d] : 0;
            ga.seg_box[0][3 + d] = d < s.ndim ? s.box_max[3 * (ilevel - 2) + d] : 1;
            ga.btype[d][0] = s.btype[d][0];
            ga.btype[d][1] = s.btype[d][1];
        }
        RH_HIP_CHECK(launch_refresh_ghosts(ga, s.stream));
    }
    const int reflux = ilevel > levelmin ? 1 : 0;
    if (s.difmag > 0.0 && s.ndim != 3) return fail(RH_ERR_ARG, "difmag > 0 is supported in 3D only");
    // deterministic refluxes: per-face slots of the level's ghosts, folded into
    // the parent cells in slot order after the launch
    double *racc = nullptr;
    if (reflux && s.deterministic && lnb.nghost > 0) {
        if (s.racc_det_cap < s.nghost_total) {
            dfree(s.racc_det);
            s.racc_det = nullptr;
            s.racc_det_cap = 0;
            const long n = 5l * kRefluxSlotsPerGhost * s.nghost_total;
            RH_HIP_CHECK(dmalloc((void **)&s.racc_det, (size_t)n * sizeof(double)));
            RH_HIP_CHECK(launch_zero(s.racc_det, n, s.stream));  // each fold leaves it zeroed
            s.racc_det_cap = s.nghost_total;
        }
        racc = s.racc_det;
    }
    if (s.ndim == 3) {
        auto it = s.bricks.find(ilevel);
        if (it == s.bricks.end() || it->second.head != head_idx || it->second.noct != num_subgrids)
            return fail(RH_ERR_STATE, "rh_godunov: rh_build_nbor was not called for this level");
        const LevelBricks &lb = it->second;
        GodArgs a;
        a.uold = s.uold;
        a.unew = s.unew;
        a.ubase = s.unew;
        a.refined = s.refined;
        a.win = lb.win;
        a.org = lb.org;
        a.zr = lb.zr;
        a.any_ref = lb.any_ref;
        a.bnd = lb.bnd;
        a.nbricks = lb.nbricks;
        a.head = head_idx;
        a.noct = num_subgrids;
        for (int d = 0; d < 3; d++) {
            a.boxmin[d] = lb.boxmin[d];
            a.boxmax[d] = lb.boxmax[d];
            a.btype[d][0] = s.btype[d][0];
            a.btype[d][1] = s.btype[d][1];
            a.g[d] = gg[d];
        }
        a.dt_dev = nullptr;
        a.dt = dt;
        a.dx = dx;
        a.inv_dx = 1.0 / dx;
        a.gamma = gamma;
        a.smallr = smallr;
        a.smallc2 = smallc2;
        a.courant_factor = 0.0;
        a.smax_cap = smax_cap;
        a.use_smax = 1;
        a.slope = slope;
        a.riemann = riemann;
        a.xcd_remap = 1;
        a.reflux = reflux;
        a.ngridmax = s.ngridmax;
        a.ghost_parent = s.ghost_parent;
        a.courant = 0;
        a.cpart = nullptr;
        a.racc = racc;
        a.rdet = racc != nullptr;
        RH_HIP_CHECK(launch_godunov3d(a, lb.bx, lb.by, lb.bz, s.precision, s.stream));
        RH_HIP_CHECK(enqueue_diffusion(s, ilevel, head_idx, num_subgrids, dx, smallr, gg, nullptr, dt, reflux, racc,
                                       racc != nullptr));
    } else {
        const LevelNbor &ln = lnb;
        God1Args a;
        a.uold = s.uold;
        a.unew = s.unew;
        a.ubase = s.unew;
        a.refined = s.refined;
        a.nbor = ln.nbor;
        a.head = head_idx;
        a.noct = num_subgrids;
        a.boxmin = ln.boxmin[0];
        a.boxmax = ln.boxmax[0];
        a.btype[0] = s.btype[0][0];
        a.btype[1] = s.btype[0][1];
        a.dt_dev = nullptr;
        a.dt = dt;
        a.dx = dx;
        a.inv_dx = 1.0 / dx;
        a.gamma = gamma;
        a.smallr = smallr;
        a.smallc2 = smallc2;
        a.courant_factor = 0.0;
        a.smax_cap = smax_cap;
        a.use_smax = 1;
        a.g = gg[0];
        a.slope = slope;
        a.riemann = riemann;
        a.reflux = reflux;
        a.ngridmax = s.ngridmax;
        a.ghost_parent = s.ghost_parent;
        a.racc = racc;
        a.rdet = racc != nullptr;
        RH_HIP_CHECK(launch_godunov1d(a, s.precision, s.stream));
    }
    if (racc)
        RH_HIP_CHECK(launch_reflux_fold(s.unew, racc, s.ghost_parent, s.ghost_order + (lnb.ghost_first - s.ngridmax - 1),
                                        lnb.nghost, s.ttd, s.stream, kRefluxSlotsPerGhost));
    return 0;
}

int rh_get_level_table(int ilevel, int which, int32_t *out, int64_t capacity, int64_t *len)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    if (!len) return fail(RH_ERR_ARG, "rh_get_level_table: null len");
    *len = 0;
    auto nb = s.nbors.find(ilevel);
    if (nb == s.nbors.end()) return fail(RH_ERR_STATE, "rh_get_level_table: rh_build_nbor was not called for this level");
    const LevelNbor &ln = nb->second;
    const int nslot = s.ndim == 3 ? 27 : 3;
    const int32_t *src = nullptr;
    int64_t n = 0;
    std::vector<int32_t> host;
    switch (which) {
    case 0: src = ln.nbor; n = (int64_t)ln.noct * nslot; break;
    case 1: src = ln.father; n = ln.father ? (int64_t)ln.noct * 2 : 0; break;
    case 2:
    case 3: {
        auto bk = s.bricks.find(ilevel);
        if (bk == s.bricks.end()) return fail(RH_ERR_STATE, "rh_get_level_table: the level has no bricks (1D)");
        const LevelBricks &lb = bk->second;
        const int64_t nw = (int64_t)(lb.bx + 2) * (lb.by + 2) * (lb.bz + 2);
        src = which == 2 ? lb.org : lb.win;
        n = which == 2 ? (int64_t)lb.nbricks * 3 : (int64_t)lb.nbricks * nw;
        break;
    }
    case 4:
    case 5: {
        const int g0 = ln.ghost_first > 0 ? ln.ghost_first - s.ngridmax - 1 : 0;
        n = (int64_t)ln.nghost * (which == 4 ? 4 : 2);
        if (capacity < n) break;
        if (n > 0 && !out) return fail(RH_ERR_ARG, "rh_get_level_table: null out");
        host.resize((size_t)ln.nghost * 3);
        if (ln.nghost > 0) {
            RH_HIP_CHECK(hipMemcpyAsync(host.data(), which == 4 ? s.ghost_key + 3 * (size_t)g0 : s.ghost_parent + 2 * (size_t)g0,
                                        (size_t)ln.nghost * (which == 4 ? 3 : 2) * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, s.stream));
            RH_HIP_CHECK(hipStreamSynchronize(s.stream));
        }
        for (int g = 0; g < ln.nghost; g++) {
            if (which == 4) {
                out[4 * g] = ln.ghost_first + g;
                for (int d = 0; d < 3; d++) out[4 * g + 1 + d] = host[3 * (size_t)g + d];
            } else {
                out[2 * g] = host[2 * (size_t)g];
                out[2 * g + 1] = host[2 * (size_t)g + 1];
            }
        }
        *len = n;
        return 0;
    }
    default: return fail(RH_ERR_ARG, "rh_get_level_table: which must be 0..5");
    }
    *len = n;
    if (capacity < n) return fail(RH_ERR_ARG, "rh_get_level_table: capacity too small");
    if (n > 0) {
        if (!out) return fail(RH_ERR_ARG, "rh_get_level_table: null out");
        RH_HIP_CHECK(hipMemcpyAsync(out, src, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream));
        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    }
    return 0;
}

int rh_brick_lattice(const int32_t *ckey, int noct, const int32_t box_min[3], const int32_t box_max[3], int bx,
                     int by, int bz, int32_t off[3], double *work)
{
    if (noct < 0 || (noct > 0 && !ckey) || !box_min || !box_max || !off || bx < 1 || by < 1 || bz < 1)
        return fail(RH_ERR_ARG, "rh_brick_lattice: bad arguments");
    for (int i = 0; i < 3 * noct; i++)
        if (ckey[i] < box_min[i % 3]) return fail(RH_ERR_ARG, "rh_brick_lattice: oct key below box_min");
    const int B[3] = {bx, by, bz};
    int o[3];
    const double w = choose_brick_offset(ckey, noct, box_min, box_max, B, o);
    for (int d = 0; d < 3; d++) off[d] = o[d];
    if (work) *work = w;
    return 0;
}

int rh_set_interpolation(int interpol_var, int interpol_type)
{
    if (interpol_var < 0 || interpol_var > 1 || interpol_type < 0 || interpol_type > 3)
        return fail(RH_ERR_ARG, "interpol_var must be 0/1 and interpol_type 0..3");
    st().interpol_var = interpol_var;
    st().interpol_type = interpol_type;
    return 0;
}

int rh_upload(int head_idx, int num_octs, int internal_energy, double gamma, double smallr, double smallc2)
{
    if (int r = ensure_mesh()) return r;
    State &s = st();
    (void)gamma;
    (void)smallc2;
    if (num_octs == 0) return 0;
    int ilevel = 0;
    if (int r = level_of_range(head_idx, num_octs, ilevel)) return r;
    auto it = s.nbors.find(ilevel);
    if (it == s.nbors.end() || it->second.head != head_idx || it->second.noct != num_octs || !it->second.father)
        return fail(RH_ERR_STATE, "rh_upload: rh_build_nbor was not called for the fine level");
    RH_HIP_CHECK(launch_upload(s.uold, it->second.father, head_idx, num_octs, s.ndim, internal_energy, smallr,
                               s.stream));
    return 0;
}

int rh_device_sync(void)
{
    if (int r = ensure_init()) return r;
    RH_HIP_CHECK(hipStreamSynchronize(st().stream));
    return 0;
}

}  // extern "C"
