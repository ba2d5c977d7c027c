// Ensemble serving extensions (include/ramses2_hip_ext.h): contexts that each
// hold one job's mesh, device-side initial conditions, and batched time steps
// that advance many jobs with shared launches (one work-list Godunov kernel
// over the bricks of every job, one per-job Courant reduction).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ramses2_hip.h"
#include "../../include/ramses2_hip_ext.h"
#include "rh_godunov.h"
#include "rh_state.h"

using namespace rh;

namespace {

int efail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

// Initial conditions of the ensemble jobs, one thread per cell of the level.
// kind 1, blast: ambient (rho0, 0, p0), pressure pb in the 2^3 cells whose
//   integer coordinates are all in [lo, lo+2) (the energy E deposited in them,
//   pb = (gamma-1) E / V computed by the caller);
// kind 2, planar two-state Riemann problem: (rl, 0, pl) where
//   (i_axis + 0.5)/ncell < 0.5, else (rr, 0, pr).
// Conservative variables as prim_to_cons: E = p/(gamma-1) + rho v^2/2.
__global__ void k_init_ic(double *uold, double *unew, const int32_t *__restrict__ ckey, int head, int noct,
                          int ncell, int kind, double a0, double a1, double a2, double a3, double a4, double gamma)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)noct * 8) return;
    const long o = (long)(head - 1) + i / 8;
    const int ind = (int)(i % 8);
    const int c[3] = {2 * ckey[3 * o] + (ind & 1), 2 * ckey[3 * o + 1] + ((ind >> 1) & 1),
                      2 * ckey[3 * o + 2] + ((ind >> 2) & 1)};
    double rho, p;
    if (kind == 1) {
        const int lo = (int)a3;
        const bool in = c[0] >= lo && c[0] < lo + 2 && c[1] >= lo && c[1] < lo + 2 && c[2] >= lo && c[2] < lo + 2;
        rho = a0;
        p = in ? a2 : a1;
    } else {
        const int axis = (int)a4;
        const bool left = ((double)c[axis] + 0.5) / (double)ncell < 0.5;
        rho = left ? a0 : a2;
        p = left ? a1 : a3;
    }
    const double ekin = 0.5 * rho * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
    const double u[5] = {rho, rho * 0.0, rho * 0.0, rho * 0.0, p / (gamma - 1.0) + ekin};
    double *po = uold + o * 40 + ind;
    double *pn = unew + o * 40 + ind;
#pragma unroll
    for (int n = 0; n < 5; n++) {
        po[8 * n] = u[n];
        pn[8 * n] = u[n];
    }
}

// Device and pinned staging buffers of batched steps (shared by all contexts):
// one block per step holding, 256-byte aligned, the job descriptors
// ([2][njobs] GodArgs: parity 0 uold->unew, parity 1 unew->uold), the per-job
// Courant records and the work items, uploaded with one copy.
struct Batch {
    size_t cap = 0;
    char *d_buf = nullptr;
    char *h_buf = nullptr;
    hipEvent_t staged = nullptr;   // the last upload from the staging buffer has completed
    bool pending = false;
};
Batch g_batch;

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int batch_reserve(size_t bytes, hipStream_t s)
{
    Batch &b = g_batch;
    if (!b.staged) {
        RH_HIP_CHECK(hipEventCreateWithFlags(&b.staged, hipEventDisableTiming));
    }
    if (b.pending) {
        RH_HIP_CHECK(hipEventSynchronize(b.staged));
        b.pending = false;
    }
    if (bytes > b.cap) {
        const size_t cap = std::max(bytes, 2 * b.cap);
        RH_HIP_CHECK(hipStreamSynchronize(s));
        (void)hipFree(b.d_buf);
        (void)hipHostFree(b.h_buf);
        b.d_buf = b.h_buf = nullptr;
        b.cap = 0;
        RH_HIP_CHECK(hipMalloc(&b.d_buf, cap));
        RH_HIP_CHECK(hipHostMalloc(&b.h_buf, cap, hipHostMallocDefault));
        b.cap = cap;
    }
    return 0;
}

// Fences of a pipelined request loop: an event on the library stream and the
// statuses of some jobs at that point, copied by the stream into pinned host
// memory (8 doubles per job: the 7 of rh_job_status and one spare).
struct Fence {
    hipEvent_t ev = nullptr;
    double *h = nullptr;
    int cap = 0, n = 0;
    bool busy = false;
};
constexpr int kMaxFences = 64;
std::vector<Fence> g_fences;

}  // namespace

extern "C" {

int rh_ctx_create(int32_t *id)
{
    if (!id) return efail(RH_ERR_ARG, "null id");
    State *s = new State();
    *id = ctx_add(s);
    return 0;
}

int rh_ctx_select(int id)
{
    State *s = ctx_get(id);
    if (!s) return efail(RH_ERR_ARG, "no context with id " + std::to_string(id));
    ctx_select(s);
    return 0;
}

int rh_ctx_current(int32_t *id)
{
    if (!id) return efail(RH_ERR_ARG, "null id");
    State *cur = &st();
    for (int i = 0; i < ctx_count(); i++)
        if (ctx_get(i) == cur) {
            *id = i;
            return 0;
        }
    return efail(RH_ERR_STATE, "current context not in the table");
}

int rh_ctx_destroy(int id)
{
    State *s = ctx_get(id);
    if (!s) return efail(RH_ERR_ARG, "no context with id " + std::to_string(id));
    State *prev = &st();
    ctx_select(s);
    (void)rh_finalize();
    if (id == 0) {
        ctx_select(prev);
        return 0;
    }
    ctx_set_slot(id, nullptr);
    delete s;
    ctx_select(prev == s ? ctx_get(0) : prev);
    return 0;
}

int rh_ctx_clone_mesh(int src_id)
{
    State *src = ctx_get(src_id);
    State &s = st();
    if (!src || src == &s) return efail(RH_ERR_ARG, "rh_ctx_clone_mesh: bad source context");
    if (!src->inited || !src->uold) return efail(RH_ERR_STATE, "the source context holds no mesh");
    if (s.uold) return efail(RH_ERR_STATE, "the current context already holds a mesh");
    if (int r = rh_init()) return r;
    // scalar and host-side mesh description
    s.ndim = src->ndim;
    s.nvar = src->nvar;
    s.ttd = src->ttd;
    s.ngridmax = src->ngridmax;
    s.ncachemax = src->ncachemax;
    s.ngridtot = src->ngridtot;
    s.hash_cap = src->hash_cap;
    s.nlevelmax = src->nlevelmax;
    s.ckey_max = src->ckey_max;
    s.key_off = src->key_off;
    s.box_min = src->box_min;
    s.box_max = src->box_max;
    for (int d = 0; d < 3; d++) {
        s.periodic[d] = src->periodic[d];
        s.btype[d][0] = src->btype[d][0];
        s.btype[d][1] = src->btype[d][1];
    }
    s.h_ckey = src->h_ckey;
    s.h_lev = src->h_lev;
    s.h_refany = src->h_refany;
    s.brick_bx = src->brick_bx;
    s.brick_by = src->brick_by;
    s.brick_bz = src->brick_bz;
    s.interpol_var = src->interpol_var;
    s.interpol_type = src->interpol_type;
    s.nghost_total = src->nghost_total;
    // device arrays: fresh state, copies of the mesh tables (all on the stream)
    const size_t nd = (size_t)s.ngridtot * s.nvar * s.ttd * sizeof(double);
    const size_t ng = (size_t)(s.ncachemax > 0 ? s.ncachemax : 1);
    auto dup = [&](void **dst, const void *from, size_t bytes) -> hipError_t {
        hipError_t e = dmalloc(dst, bytes);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(*dst, from, bytes, hipMemcpyDeviceToDevice, s.stream);
    };
    RH_HIP_CHECK(dmalloc((void **)&s.uold, nd));
    RH_HIP_CHECK(dmalloc((void **)&s.unew, nd));
    RH_HIP_CHECK(hipMemsetAsync(s.uold, 0, nd, s.stream));
    RH_HIP_CHECK(hipMemsetAsync(s.unew, 0, nd, s.stream));
    RH_HIP_CHECK(dup((void **)&s.ckey, src->ckey, (size_t)s.ngridtot * 3 * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.lev, src->lev, (size_t)s.ngridtot * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.refined, src->refined, (size_t)s.ngridtot * 8));
    RH_HIP_CHECK(dup((void **)&s.hkeys, src->hkeys, (size_t)s.hash_cap * sizeof(uint64_t)));
    RH_HIP_CHECK(dup((void **)&s.hvals, src->hvals, (size_t)s.hash_cap * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.ghost_parent, src->ghost_parent, ng * 2 * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.ghost_src, src->ghost_src, ng * 14 * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.ghost_srckey, src->ghost_srckey, ng * 21 * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.ghost_order, src->ghost_order, ng * sizeof(int32_t)));
    RH_HIP_CHECK(dup((void **)&s.ghost_key, src->ghost_key, ng * 3 * sizeof(int32_t)));
    const int nslot = s.ndim == 3 ? 27 : 3;
    for (const auto &kv : src->nbors) {
        LevelNbor ln = kv.second;
        RH_HIP_CHECK(dup((void **)&ln.nbor, kv.second.nbor, (size_t)ln.noct * nslot * sizeof(int32_t)));
        if (kv.second.father)
            RH_HIP_CHECK(dup((void **)&ln.father, kv.second.father, (size_t)ln.noct * 2 * sizeof(int32_t)));
        s.nbors[kv.first] = ln;
    }
    for (const auto &kv : src->bricks) {
        LevelBricks lb = kv.second;
        const size_t nw = (size_t)(lb.bx + 2) * (lb.by + 2) * (lb.bz + 2);
        RH_HIP_CHECK(dup((void **)&lb.win, kv.second.win, (size_t)lb.nbricks * nw * sizeof(int32_t)));
        RH_HIP_CHECK(dup((void **)&lb.org, kv.second.org, (size_t)lb.nbricks * 3 * sizeof(int32_t)));
        RH_HIP_CHECK(dup((void **)&lb.zr, kv.second.zr, (size_t)lb.nbricks * sizeof(int32_t)));
        s.bricks[kv.first] = lb;
    }
    return 0;
}

int rh_cache_release(int64_t *bytes_released)
{
    if (bytes_released) *bytes_released = cached_bytes();
    if (ctx_get(0)->inited) RH_HIP_CHECK(hipStreamSynchronize(ctx_get(0)->stream));
    release_cached();
    return 0;
}

int rh_ctx_footprint(int id, int64_t *bytes)
{
    State *s = ctx_get(id);
    if (!s) return efail(RH_ERR_ARG, "no context with id " + std::to_string(id));
    if (bytes) *bytes = s->dev_bytes;
    return 0;
}

int rh_estimate_uniform_bytes(int n_octs_per_dim, int ncachemax, int hash_size, int64_t *bytes)
{
    if (n_octs_per_dim < 1 || ncachemax < 0 || hash_size < 1 || !bytes)
        return efail(RH_ERR_ARG, "rh_estimate_uniform_bytes: bad arguments");
    const int64_t n = n_octs_per_dim, noct = n * n * n, ntot = noct + ncachemax;
    int64_t cap = 1;
    while (cap < hash_size) cap <<= 1;
    const int bx = st().brick_bx, by = st().brick_by, bz = st().brick_bz;
    const int64_t nb = ((n + bx - 1) / bx) * ((n + by - 1) / by) * ((n + bz - 1) / bz);
    int64_t b = 0;
    b += 2 * ntot * 40 * 8;                                   // uold, unew
    b += ntot * (12 + 4 + 8);                                 // ckey, lev, refined
    b += cap * (8 + 4);                                       // hash keys, values
    b += (int64_t)std::max(ncachemax, 1) * (2 + 14 + 21 + 1 + 3) * 4;  // ghost-oct tables
    b += 8 * 8 + 2 * 8 + 4 * 4 + 2 * 8;                       // reduction / dt / counters
    b += std::max<int64_t>(4096, nb) * 4 * 8;                 // Courant partials
    b += noct * 27 * 4 + noct * 2 * 4;                        // neighbour and father tables
    b += nb * ((int64_t)(bx + 2) * (by + 2) * (bz + 2) + 4) * 4;  // brick windows, origins, z ranges
    *bytes = b;
    return 0;
}

int rh_init_ic(int head_idx, int num_octs, int ncell, int kind, const double params[5], double gamma)
{
    State &s = st();
    if (!s.inited || !s.uold) return efail(RH_ERR_STATE, "library not initialised / mesh not allocated");
    if (s.ndim != 3) return efail(RH_ERR_ARG, "rh_init_ic builds 3D initial conditions");
    if (kind != 1 && kind != 2) return efail(RH_ERR_ARG, "kind must be 1 (blast) or 2 (Riemann)");
    if (!params || ncell < 2 || head_idx < 1 || num_octs < 0 || head_idx - 1 + num_octs > s.ngridtot)
        return efail(RH_ERR_ARG, "rh_init_ic: bad oct range or parameters");
    if (kind == 2 && (params[4] < 0 || params[4] > 2)) return efail(RH_ERR_ARG, "axis must be 0, 1 or 2");
    const long n = (long)num_octs * 8;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_init_ic, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, s.uold, s.unew, s.ckey,
                       head_idx, num_octs, ncell, kind, params[0], params[1], params[2], params[3], params[4], gamma);
    RH_HIP_CHECK(hipGetLastError());
    return 0;
}

int rh_job_prepare(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
                   double courant_factor, const double g[3], int slope, int riemann)
{
    State &s = st();
    if (!s.inited || !s.uold) return efail(RH_ERR_STATE, "library not initialised / mesh not allocated");
    if (s.ndim != 3) return efail(RH_ERR_ARG, "batched jobs are 3D");
    if (s.difmag > 0.0) return efail(RH_ERR_ARG, "difmag > 0 runs through rh_godunov and rh_level_run, not batched jobs");
    GodArgs a;
    const LevelBricks *lb = nullptr;
    if (int r = fill_args3(head_idx, num_octs, ilevel, dx, gamma, smallr, smallc2, courant_factor, g, slope,
                           riemann, s.dt_dev, 0.0, a, lb))
        return r;
    if (int r = ensure_partials(lb->nbricks)) return r;
    a.courant = 1;
    a.cpart = s.red_part;
    s.job = a;
    s.job_bx = lb->bx;
    s.job_by = lb->by;
    s.job_bz = lb->bz;
    s.job_dtmax = dx / __builtin_sqrt(smallc2);
    // dt of the first step: a standalone Courant pass over the initial state
    const double gz[3] = {0, 0, 0};
    RH_HIP_CHECK(hipMemsetAsync(s.t_dev, 0, sizeof(double), s.stream));
    RH_HIP_CHECK(launch_courant(s.uold, s.refined, head_idx, num_octs, 3, dx, gamma, smallr, __builtin_sqrt(smallc2),
                                courant_factor, g ? g : gz, s.red_part, s.red_cap, s.red_out, s.dt_dev, s.stream,
                                nullptr));
    s.job_ready = true;
    return 0;
}

int rh_batch_step(const int32_t *ids, int njobs, int precision, int nsteps)
{
    if (njobs < 0 || nsteps < 0 || (njobs > 0 && !ids)) return efail(RH_ERR_ARG, "rh_batch_step: bad arguments");
    if (precision != 0 && precision != 1) return efail(RH_ERR_ARG, "precision must be 0 or 1");
    if (njobs == 0 || nsteps == 0) return 0;
    std::vector<State *> js(njobs);
    long nitems = 0;
    bool gen = false;  // some job needs the general brick body
    for (int j = 0; j < njobs; j++) {
        State *s = ctx_get(ids[j]);
        if (!s || !s->job_ready)
            return efail(RH_ERR_STATE, "context " + std::to_string(ids[j]) + " has no prepared job");
        for (int k = 0; k < j; k++)
            if (js[k] == s) return efail(RH_ERR_ARG, "a context appears twice in the batch");
        js[j] = s;
        if (s->job_bx != js[0]->job_bx || s->job_by != js[0]->job_by || s->job_bz != js[0]->job_bz ||
            s->job.slope != js[0]->job.slope || s->job.riemann != js[0]->job.riemann)
            return efail(RH_ERR_ARG, "the jobs of one batch need the same brick shape, slope and Riemann solver");
        nitems += s->job.nbricks;
        gen = gen || s->job.bnd || s->job.any_ref || s->job.reflux;  // (batches add to uold: ubase == uold)
    }
    if (nitems > (1l << 30)) return efail(RH_ERR_ARG, "too many bricks in one batch");
    hipStream_t stream = js[0]->stream;
    const size_t off_c = align256(2 * (size_t)njobs * sizeof(GodArgs));
    const size_t off_i = off_c + align256((size_t)njobs * sizeof(CourantJob));
    const size_t bytes = off_i + (size_t)nitems * sizeof(int2);
    if (int r = batch_reserve(bytes, stream)) return r;
    Batch &b = g_batch;
    GodArgs *h_descs = (GodArgs *)b.h_buf;
    CourantJob *h_cjobs = (CourantJob *)(b.h_buf + off_c);
    int2 *h_items = (int2 *)(b.h_buf + off_i);
    int pos = 0;
    for (int j = 0; j < njobs; j++) {
        State &s = *js[j];
        GodArgs a = s.job;
        a.uold = s.uold;
        a.unew = s.unew;
        a.ubase = s.uold;
        h_descs[j] = a;
        a.uold = s.unew;
        a.unew = s.uold;
        a.ubase = s.unew;
        h_descs[njobs + j] = a;
        h_cjobs[j] = CourantJob{s.red_part, s.job.nbricks, s.job_dtmax, s.red_out, s.dt_dev, s.t_dev};
        for (int i = 0; i < s.job.nbricks; i++) h_items[pos++] = make_int2(j, i);
    }
    RH_HIP_CHECK(hipMemcpyAsync(b.d_buf, b.h_buf, bytes, hipMemcpyHostToDevice, stream));
    RH_HIP_CHECK(hipEventRecord(b.staged, stream));
    b.pending = true;
    const GodArgs *d_descs = (const GodArgs *)b.d_buf;
    const CourantJob *d_cjobs = (const CourantJob *)(b.d_buf + off_c);
    WorkList w;
    w.items = (const int2 *)(b.d_buf + off_i);
    for (int r = 0; r <= 8; r++) w.range[r] = (int)((nitems * r) / 8);
    for (int k = 0; k < nsteps; k++) {
        w.descs = d_descs + (k & 1) * njobs;
        RH_HIP_CHECK(launch_godunov3d_wl(w, js[0]->job_bx, js[0]->job_by, js[0]->job_bz, js[0]->job.slope,
                                         js[0]->job.riemann, precision, stream, gen));
        RH_HIP_CHECK(launch_courant_reduce_batch(d_cjobs, njobs, stream));
    }
    if (nsteps & 1)
        for (State *s : js) std::swap(s->uold, s->unew);
    return 0;
}

int rh_job_status(int id, double out[7])
{
    State *s = ctx_get(id);
    if (!s || !s->inited) return efail(RH_ERR_ARG, "no initialised context with id " + std::to_string(id));
    if (!out) return efail(RH_ERR_ARG, "null output");
    double v[8];
    RH_HIP_CHECK(hipMemcpyAsync(v, s->t_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    RH_HIP_CHECK(hipMemcpyAsync(v + 1, s->dt_dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    RH_HIP_CHECK(hipMemcpyAsync(v + 2, s->red_out, 5 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    RH_HIP_CHECK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < 7; i++) out[i] = v[i];
    return 0;
}

int rh_fence_record(const int32_t *ids, int n, int32_t *fence)
{
    if (n < 0 || (n > 0 && !ids) || !fence) return efail(RH_ERR_ARG, "rh_fence_record: bad arguments");
    std::vector<State *> js(n);
    for (int j = 0; j < n; j++) {
        js[j] = ctx_get(ids[j]);
        if (!js[j] || !js[j]->inited)
            return efail(RH_ERR_ARG, "no initialised context with id " + std::to_string(ids[j]));
    }
    hipStream_t stream;
    if (n > 0) {
        stream = js[0]->stream;  // every context enqueues on the one library stream
    } else {
        if (!st().inited) return efail(RH_ERR_STATE, "rh_init has not been called");
        stream = st().stream;
    }
    int f = 0;
    while (f < (int)g_fences.size() && g_fences[f].busy) f++;
    if (f == kMaxFences) return efail(RH_ERR_STATE, "more than 64 fences outstanding");
    if (f == (int)g_fences.size()) g_fences.emplace_back();
    Fence &fc = g_fences[f];
    if (!fc.ev) RH_HIP_CHECK(hipEventCreateWithFlags(&fc.ev, hipEventDisableTiming));
    if (n > fc.cap) {
        // a free fence: no copy into its buffer is pending
        const int cap = std::max(n, std::max(16, 2 * fc.cap));
        (void)hipHostFree(fc.h);
        fc.h = nullptr;
        fc.cap = 0;
        RH_HIP_CHECK(hipHostMalloc(&fc.h, (size_t)cap * 8 * sizeof(double), hipHostMallocDefault));
        fc.cap = cap;
    }
    for (int j = 0; j < n; j++) {
        const State &s = *js[j];
        double *v = fc.h + 8 * j;
        RH_HIP_CHECK(hipMemcpyAsync(v, s.t_dev, sizeof(double), hipMemcpyDeviceToHost, stream));
        RH_HIP_CHECK(hipMemcpyAsync(v + 1, s.dt_dev, sizeof(double), hipMemcpyDeviceToHost, stream));
        RH_HIP_CHECK(hipMemcpyAsync(v + 2, s.red_out, 5 * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    RH_HIP_CHECK(hipEventRecord(fc.ev, stream));
    fc.n = n;
    fc.busy = true;
    *fence = f;
    return 0;
}

int rh_fence_wait(int32_t fence, double *out)
{
    if (fence < 0 || fence >= (int)g_fences.size() || !g_fences[fence].busy)
        return efail(RH_ERR_ARG, "no outstanding fence " + std::to_string(fence));
    Fence &fc = g_fences[fence];
    fc.busy = false;
    RH_HIP_CHECK(hipEventSynchronize(fc.ev));
    if (out)
        for (int j = 0; j < fc.n; j++)
            for (int i = 0; i < 7; i++) out[7 * j + i] = fc.h[8 * j + i];
    return 0;
}

}  // extern "C"
