// Extensions beyond the drop-in ABI (include/ramses2_hip_ext.h): precision
// switch, brick shape, event timer, and device-resident time stepping of a
// single-level mesh with the Courant dt kept on the device (optionally replayed
// as a captured hipGraph).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/ramses2_hip.h"
#include "../../include/ramses2_hip_ext.h"
#include "rh_godunov.h"
#include "rh_state.h"

using namespace rh;

namespace {

int efail(int code, const char *msg)
{
    set_error(msg);
    return code;
}

// The level must be the whole mesh: every allocated oct outside the range is unused.
bool level_is_whole_mesh(int head_idx, int num_octs)
{
    State &s = st();
    for (int i = 0; i < s.ngridtot; i++) {
        const bool in = i >= head_idx - 1 && i < head_idx - 1 + num_octs;
        if (!in && s.h_lev[i] != 0) return false;
    }
    return true;
}

}  // namespace

namespace rh {

// Fill the Godunov arguments of a 3D level in fused mode (unew = uold + dF).
int fill_args3(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
               double courant_factor, const double g[3], int slope, int riemann, const double *dt_dev, double dt,
               GodArgs &a, const LevelBricks *&lbo)
{
    State &s = st();
    auto it = s.bricks.find(ilevel);
    if (it == s.bricks.end() || it->second.head != head_idx || it->second.noct != num_octs) {
        set_error("rh_build_nbor was not called for this level");
        return RH_ERR_STATE;
    }
    const LevelBricks &lb = it->second;
    lbo = &lb;
    a.uold = s.uold;
    a.unew = s.unew;
    a.ubase = s.uold;
    a.refined = s.refined;
    a.win = lb.win;
    a.org = lb.org;
    a.zr = lb.zr;
    a.any_ref = lb.any_ref;
    a.bnd = lb.bnd;
    a.nbricks = lb.nbricks;
    a.head = head_idx;
    a.noct = num_octs;
    for (int d = 0; d < 3; d++) {
        a.boxmin[d] = lb.boxmin[d];
        a.boxmax[d] = lb.boxmax[d];
        a.btype[d][0] = s.btype[d][0];
        a.btype[d][1] = s.btype[d][1];
        a.g[d] = g ? g[d] : 0.0;
    }
    a.dt_dev = dt_dev;
    a.dt = dt;
    a.dx = dx;
    a.inv_dx = 1.0 / dx;
    a.gamma = gamma;
    a.smallr = smallr;
    a.smallc2 = smallc2;
    a.courant_factor = courant_factor;
    a.smax_cap = 0.0;
    a.use_smax = 0;
    a.slope = slope;
    a.riemann = riemann;
    a.xcd_remap = 1;
    a.reflux = 0;
    a.ngridmax = s.ngridmax;
    a.ghost_parent = s.ghost_parent;
    a.courant = 0;
    a.cpart = nullptr;
    a.racc = nullptr;
    return 0;
}

int fill_args1(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
               double courant_factor, const double g[3], int slope, int riemann, const double *dt_dev, double dt,
               God1Args &a)
{
    State &s = st();
    auto it = s.nbors.find(ilevel);
    if (it == s.nbors.end() || it->second.head != head_idx || it->second.noct != num_octs) {
        set_error("rh_build_nbor was not called for this level");
        return RH_ERR_STATE;
    }
    const LevelNbor &ln = it->second;
    a.uold = s.uold;
    a.unew = s.unew;
    a.ubase = s.uold;
    a.refined = s.refined;
    a.nbor = ln.nbor;
    a.head = head_idx;
    a.noct = num_octs;
    a.boxmin = ln.boxmin[0];
    a.boxmax = ln.boxmax[0];
    a.btype[0] = s.btype[0][0];
    a.btype[1] = s.btype[0][1];
    a.dt_dev = dt_dev;
    a.dt = dt;
    a.dx = dx;
    a.inv_dx = 1.0 / dx;
    a.gamma = gamma;
    a.smallr = smallr;
    a.smallc2 = smallc2;
    a.courant_factor = courant_factor;
    a.smax_cap = 0.0;
    a.use_smax = 0;
    a.g = g ? g[0] : 0.0;
    a.slope = slope;
    a.riemann = riemann;
    a.reflux = 0;
    a.ngridmax = s.ngridmax;
    a.ghost_parent = s.ghost_parent;
    return 0;
}

static void drop_graph_entry(State &s, int i)
{
    if (s.level_exec[i]) (void)hipGraphExecDestroy(s.level_exec[i]);
    if (s.level_graph[i]) (void)hipGraphDestroy(s.level_graph[i]);
    s.level_exec[i] = nullptr;
    s.level_graph[i] = nullptr;
    s.graph_key[i].clear();
}

void drop_level_graph(State &s)
{
    drop_graph_entry(s, 0);
    drop_graph_entry(s, 1);
}

// Grow the partial-sum buffer to n entries of 4 doubles.
int ensure_partials(int n)
{
    State &s = st();
    if (n <= s.red_cap) return 0;
    drop_level_graph(s);
    dfree(s.red_part);
    s.red_part = nullptr;
    RH_HIP_CHECK(dmalloc((void **)&s.red_part, (size_t)n * 4 * sizeof(double)));
    s.red_cap = n;
    return 0;
}

hipError_t enqueue_diffusion(State &s, int ilevel, int head_idx, int num_octs, double dx, double smallr,
                             const double g[3], const double *dt_dev, double dt, int reflux, double *racc, int rdet)
{
    if (!(s.difmag > 0.0) || num_octs <= 0) return hipSuccess;
    auto it = s.nbors.find(ilevel);
    if (it == s.nbors.end()) return hipErrorInvalidValue;
    const LevelNbor &ln = it->second;
    DiffArgs d;
    d.uold = s.uold;
    d.unew = s.unew;
    d.refined = s.refined;
    d.nbor = ln.nbor;
    d.ckey = s.ckey;
    d.head = head_idx;
    d.noct = num_octs;
    d.ndim = s.ndim;
    d.refl = 0;
    for (int k = 0; k < 3; k++) {
        d.boxmin[k] = ln.boxmin[k];
        d.boxmax[k] = ln.boxmax[k];
        for (int sd = 0; sd < 2; sd++) d.refl |= (s.btype[k][sd] == 1) << (2 * k + sd);
        d.g[k] = g ? g[k] : 0.0;
    }
    d.dt_dev = dt_dev;
    d.dt = dt;
    d.dx = dx;
    d.difmag = s.difmag;
    d.smallr = smallr;
    d.reflux = reflux;
    d.ngridmax = s.ngridmax;
    d.ghost_parent = s.ghost_parent;
    d.racc = racc;
    d.rdet = rdet;
    return launch_diffusion3d(d, s.stream);
}

// One step.  3D with a Courant dt and fusion on: the Godunov kernel also
// reduces the Courant condition of the state it writes (dt of the next step),
// followed by a one-block final reduction; otherwise courant then godunov.
}  // namespace rh

namespace {

// Whether the Courant condition of the next step is reduced by the Godunov
// kernel (the FP64 kernel takes each cell's new state as it stores it, the
// FP32 kernel re-reads its brick's cells) instead of a separate pass over the
// state.  Measured at 256^3 turbulence, 8x6x16 bricks: FP64 step 1.15 -> 1.06
// ms, and a little faster in mixed precision, hence the automatic choice.
bool courant_fused()
{
    const State &s = st();
    // the diffusion pass changes the state after the Godunov kernel reduced it
    if (s.difmag > 0.0) return false;
    return s.fused_courant < 0 ? true : s.fused_courant != 0;
}

hipError_t enqueue_step(int head_idx, int num_octs, double dx, double gamma, double smallr, double smallc2,
                        double courant_factor, const double g[3], double dt_fixed, GodArgs *a3, God1Args *a1,
                        const LevelBricks *lb, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    State &s = st();
    const double gz[3] = {0, 0, 0};
    if (a3 && dt_fixed <= 0.0 && courant_fused()) {
        a3->uold = s.uold;
        a3->unew = s.unew;
        a3->ubase = s.uold;
        a3->courant = 1;
        a3->cpart = s.red_part;
        int nparts = 0;
        hipError_t e = launch_godunov3d(*a3, lb->bx, lb->by, lb->bz, s.precision, s.stream, e0, e1, &nparts);
        if (e != hipSuccess) return e;
        return launch_courant_reduce(s.red_part, nparts, dx / __builtin_sqrt(smallc2), s.red_out, s.dt_dev,
                                     s.t_dev, s.stream);
    }
    if (dt_fixed <= 0.0) {
        hipError_t e = launch_courant(s.uold, s.refined, head_idx, num_octs, s.ndim, dx, gamma, smallr,
                                      __builtin_sqrt(smallc2), courant_factor, g ? g : gz, s.red_part, s.red_cap,
                                      s.red_out, s.dt_dev, s.stream, s.t_dev);
        if (e != hipSuccess) return e;
    }
    hipError_t e;
    if (a3) {
        a3->uold = s.uold;
        a3->unew = s.unew;
        a3->ubase = s.uold;
        e = launch_godunov3d(*a3, lb->bx, lb->by, lb->bz, s.precision, s.stream, e0, e1);
        if (e == hipSuccess)
            e = enqueue_diffusion(s, lb->ilevel, head_idx, num_octs, dx, smallr, g, a3->dt_dev, a3->dt, 0, nullptr, 0);
    } else {
        a1->uold = s.uold;
        a1->unew = s.unew;
        a1->ubase = s.uold;
        if (e0) (void)hipEventRecord(e0, s.stream);
        e = launch_godunov1d(*a1, s.precision, s.stream);
        if (e1) (void)hipEventRecord(e1, s.stream);
    }
    return e;
}

__global__ void k_copy_probe(const double4 *__restrict__ a, double4 *__restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

}  // namespace

extern "C" {

int rh_set_precision(int precision)
{
    if (precision != 0 && precision != 1) return efail(RH_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed)");
    st().precision = precision;
    return 0;
}

int rh_set_fused_courant(int mode)
{
    if (mode < -1 || mode > 1) return efail(RH_ERR_ARG, "fused Courant mode must be -1, 0 or 1");
    st().fused_courant = mode;
    return 0;
}

int rh_set_brick_shape(int bx, int by, int bz)
{
    if (!brick_shape_supported(bx, by, bz)) return efail(RH_ERR_ARG, "brick shape has no compiled kernel");
    st().brick_bx = bx;
    st().brick_by = by;
    st().brick_bz = bz;
    return 0;
}

int rh_set_difmag(double difmag)
{
    if (!(difmag >= 0.0)) return efail(RH_ERR_ARG, "rh_set_difmag: difmag must be >= 0");
    st().difmag = difmag;
    drop_level_graph(st());
    return 0;
}

int rh_set_deterministic(int on)
{
    if (on != 0 && on != 1) return efail(RH_ERR_ARG, "rh_set_deterministic: 0 or 1");
    st().deterministic = on;
    return 0;
}

int rh_set_specialized(int on)
{
    if (on != 0 && on != 1) return efail(RH_ERR_ARG, "rh_set_specialized: 0 or 1");
    g_specialized = on;
    return 0;
}

int rh_get_brick_shape(int32_t *bx, int32_t *by, int32_t *bz)
{
    if (bx) *bx = st().brick_bx;
    if (by) *by = st().brick_by;
    if (bz) *bz = st().brick_bz;
    return 0;
}

int rh_timer_start(void)
{
    State &s = st();
    if (!s.inited) return efail(RH_ERR_STATE, "rh_init has not been called");
    RH_HIP_CHECK(hipEventRecord(s.ev0, s.stream));
    return 0;
}

int rh_timer_stop(float *ms)
{
    State &s = st();
    if (!s.inited) return efail(RH_ERR_STATE, "rh_init has not been called");
    RH_HIP_CHECK(hipEventRecord(s.ev1, s.stream));
    RH_HIP_CHECK(hipEventSynchronize(s.ev1));
    float t = 0;
    RH_HIP_CHECK(hipEventElapsedTime(&t, s.ev0, s.ev1));
    if (ms) *ms = t;
    return 0;
}

int rh_level_run(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
                 double courant_factor, const double g[3], int slope, int riemann, double dt_fixed, int nsteps,
                 int use_graph, double *dt_last, double *t_total, double sums[4], float *kernel_ms)
{
    State &s = st();
    if (!s.inited || !s.uold) return efail(RH_ERR_STATE, "library not initialised / mesh not allocated");
    if (nsteps < 0) return efail(RH_ERR_ARG, "nsteps must be >= 0");
    if (!level_is_whole_mesh(head_idx, num_octs))
        return efail(RH_ERR_STATE, "rh_level_run needs a single-level mesh (use the drop-in calls for AMR)");
    if (s.difmag > 0.0 && s.ndim != 3) return efail(RH_ERR_ARG, "difmag > 0 is supported in 3D only");
    GodArgs a3;
    God1Args a1;
    const LevelBricks *lb = nullptr;
    const double *dtp = dt_fixed > 0.0 ? nullptr : s.dt_dev;
    int r = s.ndim == 3 ? fill_args3(head_idx, num_octs, ilevel, dx, gamma, smallr, smallc2, courant_factor, g, slope,
                                     riemann, dtp, dt_fixed, a3, lb)
                        : fill_args1(head_idx, num_octs, ilevel, dx, gamma, smallr, smallc2, courant_factor, g,
                                     slope, riemann, dtp, dt_fixed, a1);
    if (r) return r;
    RH_HIP_CHECK(hipMemsetAsync(s.t_dev, 0, sizeof(double), s.stream));
    if (s.ndim == 3 && dt_fixed <= 0.0) {
        // dt of the first step from a standalone Courant pass; each step's
        // kernel then produces the dt of the next one
        if (int r2 = ensure_partials(2 * lb->nbricks + 8)) return r2;  // bricks and split halves
        const double gz[3] = {0, 0, 0};
        RH_HIP_CHECK(launch_courant(s.uold, s.refined, head_idx, num_octs, s.ndim, dx, gamma, smallr,
                                    __builtin_sqrt(smallc2), courant_factor, g ? g : gz, s.red_part, s.red_cap,
                                    s.red_out, s.dt_dev, s.stream, nullptr));
    }
    if (use_graph && nsteps >= 2) {
        // the pair of steps uold->unew, unew->uold as one graph, captured once
        // for each order of the two arrays and kept for later calls with the
        // same arguments, arrays and tables
        auto make_key = [&]() {
            std::vector<uint64_t> key;
            auto put = [&](double v) {
                uint64_t b;
                memcpy(&b, &v, sizeof b);
                key.push_back(b);
            };
            for (double v : {(double)head_idx, (double)num_octs, (double)ilevel, dx, gamma, smallr, smallc2,
                             courant_factor, (double)slope, (double)riemann, dt_fixed, (double)s.precision,
                             (double)courant_fused()})
                put(v);
            for (int d = 0; d < 3; d++) put(g ? g[d] : 0.0);
            for (int d = 0; d < 3; d++) put(s.btype[d][0] * 4.0 + s.btype[d][1]);
            if (lb)
                for (int v : {lb->nbricks, lb->bx, lb->by, lb->bz, lb->any_ref, lb->bnd, lb->head, lb->noct}) put(v);
            for (const void *q : {(const void *)s.uold, (const void *)s.unew, (const void *)s.refined,
                                  (const void *)s.red_part, (const void *)s.red_out, (const void *)s.dt_dev,
                                  (const void *)s.t_dev, lb ? (const void *)lb->win : nullptr,
                                  lb ? (const void *)lb->org : nullptr, lb ? (const void *)lb->zr : nullptr})
                key.push_back((uint64_t)(uintptr_t)q);
            return key;
        };
        // capture the step pair in the arrays' current order into entry ge
        auto capture = [&](int ge) -> int {
            drop_graph_entry(s, ge);
            const std::vector<uint64_t> key = make_key();
            hipGraph_t graph = nullptr;
            RH_HIP_CHECK(hipStreamSynchronize(s.stream));
            RH_HIP_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < 2; k++) {
                hipError_t e = enqueue_step(head_idx, num_octs, dx, gamma, smallr, smallc2, courant_factor, g,
                                            dt_fixed, s.ndim == 3 ? &a3 : nullptr, s.ndim == 3 ? nullptr : &a1, lb);
                if (e != hipSuccess) {
                    // the arrays stay where they were (a swap only follows a captured step)
                    if (k == 1) std::swap(s.uold, s.unew);
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(s.stream, &g2);
                    if (g2) (void)hipGraphDestroy(g2);
                    set_error(hipGetErrorString(e));
                    return RH_ERR_HIP;
                }
                std::swap(s.uold, s.unew);
            }
            RH_HIP_CHECK(hipStreamEndCapture(s.stream, &graph));
            hipGraphExec_t exec = nullptr;
            hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(graph);
                set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                return RH_ERR_HIP;
            }
            s.level_graph[ge] = graph;
            s.level_exec[ge] = exec;
            s.graph_key[ge] = key;
            return 0;
        };
        const std::vector<uint64_t> key = make_key();
        int ge = -1;
        for (int i = 0; i < 2; i++)
            if (s.level_exec[i] && key == s.graph_key[i]) ge = i;
        if (ge < 0) {
            // both orders at once, so that a later call after an odd number
            // of steps replays without capturing
            if (int r = capture(0)) return r;
            std::swap(s.uold, s.unew);
            const int r1 = capture(1);
            std::swap(s.uold, s.unew);
            if (r1) return r1;
            ge = 0;
        }
        const int npairs = nsteps / 2;
        for (int i = 0; i < npairs; i++) RH_HIP_CHECK(hipGraphLaunch(s.level_exec[ge], s.stream));
        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
        // after an even number of swaps the roles are back where they started;
        // the state of the last step is in uold (the captured pair ended on a swap)
        if (nsteps % 2) {
            RH_HIP_CHECK(enqueue_step(head_idx, num_octs, dx, gamma, smallr, smallc2, courant_factor, g, dt_fixed,
                                      s.ndim == 3 ? &a3 : nullptr, s.ndim == 3 ? nullptr : &a1, lb));
            std::swap(s.uold, s.unew);
        }
    } else {
        // per-step events around the Godunov launch (kernel time inside the timed steps)
        std::vector<hipEvent_t> ev;
        if (kernel_ms) {
            ev.resize(2 * (size_t)nsteps);
            for (auto &e : ev) RH_HIP_CHECK(hipEventCreate(&e));
        }
        for (int k = 0; k < nsteps; k++) {
            RH_HIP_CHECK(enqueue_step(head_idx, num_octs, dx, gamma, smallr, smallc2, courant_factor, g, dt_fixed,
                                      s.ndim == 3 ? &a3 : nullptr, s.ndim == 3 ? nullptr : &a1, lb,
                                      kernel_ms ? ev[2 * k] : nullptr, kernel_ms ? ev[2 * k + 1] : nullptr));
            std::swap(s.uold, s.unew);
        }
        if (kernel_ms) {
            RH_HIP_CHECK(hipStreamSynchronize(s.stream));
            double tot = 0;
            for (int k = 0; k < nsteps; k++) {
                float t = 0;
                RH_HIP_CHECK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
                tot += t;
            }
            kernel_ms[0] = nsteps ? (float)(tot / nsteps) : 0.0f;
            for (auto &e : ev) (void)hipEventDestroy(e);
        }
    }
    // unew mirrors uold as after set_uold
    const size_t per = (size_t)s.nvar * s.ttd;
    RH_HIP_CHECK(hipMemcpyAsync(s.unew + per * (head_idx - 1), s.uold + per * (head_idx - 1),
                                per * num_octs * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    double h[2] = {0, 0};
    RH_HIP_CHECK(hipMemcpyAsync(s.h_red, s.red_out, 5 * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipMemcpyAsync(h, s.t_dev, sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    if (dt_last) *dt_last = dt_fixed > 0.0 ? dt_fixed : s.h_red[4];
    if (t_total) *t_total = dt_fixed > 0.0 ? dt_fixed * nsteps : h[0];
    if (sums)
        for (int i = 0; i < 4; i++) sums[i] = s.h_red[i];
    return 0;
}

int rh_copy_probe(int64_t nbytes, float *ms)
{
    State &s = st();
    if (!s.inited) return efail(RH_ERR_STATE, "rh_init has not been called");
    if (nbytes <= 0 || nbytes % 32) return efail(RH_ERR_ARG, "nbytes must be a positive multiple of 32");
    void *a = nullptr, *b = nullptr;
    RH_HIP_CHECK(hipMalloc(&a, (size_t)nbytes));
    RH_HIP_CHECK(hipMalloc(&b, (size_t)nbytes));
    RH_HIP_CHECK(hipMemsetAsync(a, 0, (size_t)nbytes, s.stream));
    const long n = (long)(nbytes / 32);
    RH_HIP_CHECK(hipEventRecord(s.ev0, s.stream));
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, s.stream, (const double4 *)a, (double4 *)b, n);
    RH_HIP_CHECK(hipGetLastError());
    RH_HIP_CHECK(hipEventRecord(s.ev1, s.stream));
    RH_HIP_CHECK(hipEventSynchronize(s.ev1));
    float t = 0;
    RH_HIP_CHECK(hipEventElapsedTime(&t, s.ev0, s.ev1));
    if (ms) *ms = t;
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

int rh_time_godunov(int head_idx, int num_octs, int ilevel, double dx, double gamma, double smallr, double smallc2,
                    double courant_factor, const double g[3], int slope, int riemann, double dt, int nlaunch,
                    float *ms_per_launch)
{
    State &s = st();
    if (!s.inited || !s.uold) return efail(RH_ERR_STATE, "library not initialised / mesh not allocated");
    if (nlaunch < 1 || !(dt > 0)) return efail(RH_ERR_ARG, "nlaunch >= 1 and dt > 0 required");
    if (!level_is_whole_mesh(head_idx, num_octs)) return efail(RH_ERR_STATE, "single-level mesh required");
    GodArgs a3;
    God1Args a1;
    const LevelBricks *lb = nullptr;
    int r = s.ndim == 3 ? fill_args3(head_idx, num_octs, ilevel, dx, gamma, smallr, smallc2, courant_factor, g, slope,
                                     riemann, nullptr, dt, a3, lb)
                        : fill_args1(head_idx, num_octs, ilevel, dx, gamma, smallr, smallc2, courant_factor, g,
                                     slope, riemann, nullptr, dt, a1);
    if (r) return r;
    RH_HIP_CHECK(hipEventRecord(s.ev0, s.stream));
    for (int k = 0; k < nlaunch; k++) {
        if (s.ndim == 3) {
            a3.uold = s.uold;
            a3.unew = s.unew;
            a3.ubase = s.uold;
            RH_HIP_CHECK(launch_godunov3d(a3, lb->bx, lb->by, lb->bz, s.precision, s.stream));
        } else {
            a1.uold = s.uold;
            a1.unew = s.unew;
            a1.ubase = s.uold;
            RH_HIP_CHECK(launch_godunov1d(a1, s.precision, s.stream));
        }
        std::swap(s.uold, s.unew);
    }
    RH_HIP_CHECK(hipEventRecord(s.ev1, s.stream));
    RH_HIP_CHECK(hipEventSynchronize(s.ev1));
    float t = 0;
    RH_HIP_CHECK(hipEventElapsedTime(&t, s.ev0, s.ev1));
    if (ms_per_launch) *ms_per_launch = t / nlaunch;
    const size_t per = (size_t)s.nvar * s.ttd;
    RH_HIP_CHECK(hipMemcpyAsync(s.unew + per * (head_idx - 1), s.uold + per * (head_idx - 1),
                                per * num_octs * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

}  // extern "C"
