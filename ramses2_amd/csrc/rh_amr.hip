// Batched AMR time steps (include/ramses2_hip_ext.h, rh_amr_prepare /
// rh_amr_run): all levels of a multi-level mesh advance together with one
// shared dt (amr_step with nsubcycle=1 on every level, amr/amr_step.f90:588).
//
// One step on the device:
//   1. interpol_hydro into the ghost octs of every level above levelmin;
//   2. one work-list Godunov launch over the bricks of all levels
//      (unew = uold + dF, set_unew/set_uold folded in); coarse-fine refluxes
//      go to a per-ghost-oct accumulator, because the parent cell's own update
//      runs in the same launch;
//   3. the accumulated refluxes are added to the parent cells;
//   4. upload_fine, finest level first;
//   5. courant_fine over the leaf cells of all levels in one pass: the dt of
//      the next step and the conservation sums, kept on the device.
// The arrays then swap roles.  A pair of steps can be captured as a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
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

// interpol_hydro into the ghost octs of the levels above levelmin, all of
// them in one launch of up to kMaxGhostSegs levels: the levels read only
// their parent levels' state, which no refresh writes.
hipError_t refresh_ghosts(State &s, const std::vector<int> &levels, int levelmin, double *u)
{
    GhostArgs ga;
    ga.uold = u;
    ga.refined = s.refined;
    ga.src = s.ghost_src;
    ga.src_ld = ghost_ld(s);
    ga.srckey = s.ghost_srckey;
    ga.order = s.ghost_order;
    ga.gi0 = 0;
    ga.first = s.ngridmax + 1;
    ga.ndim = 3;
    ga.interpol_var = s.interpol_var;
    ga.interpol_type = s.interpol_type;
    ga.smallr = s.amr.smallr;
    for (int d = 0; d < 3; d++) {
        ga.btype[d][0] = s.btype[d][0];
        ga.btype[d][1] = s.btype[d][1];
    }
    ga.nseg = 0;
    ga.nghost = 0;
    for (int l : levels) {
        if (l <= levelmin) continue;
        auto it = s.nbors.find(l);
        if (it == s.nbors.end() || it->second.nghost <= 0) continue;
        const LevelNbor &lnb = it->second;
        if (ga.nseg == kMaxGhostSegs) {
            hipError_t e = launch_refresh_ghosts(ga, s.stream);
            if (e != hipSuccess) return e;
            ga.nseg = 0;
            ga.nghost = 0;
        }
        const int k = ga.nseg++;
        ga.seg_t[k] = ga.nghost;
        ga.seg_o[k] = lnb.ghost_first - s.ngridmax - 1;
        for (int d = 0; d < 3; d++) {
            ga.seg_box[k][d] = s.box_min[3 * (l - 2) + d];
            ga.seg_box[k][3 + d] = s.box_max[3 * (l - 2) + d];
        }
        ga.nghost += lnb.nghost;
        ga.seg_t[k + 1] = ga.nghost;
    }
    return ga.nseg > 0 ? launch_refresh_ghosts(ga, s.stream) : hipSuccess;
}

// One batched step reading the array of parity p (0: uold, 1: unew).
hipError_t enqueue_amr_step(State &s, int parity, int precision)
{
    AmrPlan &m = s.amr;
    double *ucur = parity ? s.unew : s.uold;
    double *unext = parity ? s.uold : s.unew;
    hipError_t e;
    if ((e = refresh_ghosts(s, m.levels, m.levelmin, ucur)) != hipSuccess) return e;
    for (const auto &gr : m.groups) {
        WorkList w;
        w.descs = m.d_descs + parity * m.ndesc;
        w.items = m.d_items + gr.first;
        for (int r = 0; r <= 8; r++) w.range[r] = gr.range[r];
        if ((e = launch_godunov3d_wl(w, gr.bx, gr.by, gr.bz, m.slope, m.riemann, precision, s.stream)) != hipSuccess)
            return e;
    }
    if ((e = launch_reflux_fold(unext, m.racc, s.ghost_parent, s.ghost_order, s.nghost_total, 8, s.stream,
                                m.nslot)) != hipSuccess)
        return e;
    for (int i = (int)m.levels.size() - 1; i >= 1; i--) {
        const int lf = m.levels[i];
        if (m.levels[i - 1] != lf - 1) continue;
        const LevelNbor &fn = s.nbors.at(lf);
        if ((e = launch_upload(unext, fn.father, fn.head, fn.noct, 3, s.interpol_var == 1, m.smallr, s.stream)) !=
            hipSuccess)
            return e;
    }
    return launch_courant_ml(unext, s.refined, s.lev, m.head0, m.noct_all, m.boxlen, m.dx_finest, m.gamma, m.smallr,
                             __builtin_sqrt(m.smallc2), m.courant, m.g, s.red_part, s.red_cap, s.red_out, s.dt_dev,
                             s.stream, s.t_dev, precision == 1);
}

}  // namespace

extern "C" {

int rh_amr_prepare(int levelmin, int levelmax, const int32_t *head, const int32_t *noct, double boxlen, double gamma,
                   double smallr, double smallc2, double courant_factor, const double g[3], int slope, int riemann)
{
    State &s = st();
    if (!s.inited || !s.uold) return efail(RH_ERR_STATE, "library not initialised / mesh not allocated");
    if (s.ndim != 3) return efail(RH_ERR_ARG, "batched AMR stepping is 3D");
    if (s.difmag > 0.0)
        return efail(RH_ERR_ARG, "difmag > 0 runs through rh_godunov and rh_level_run, not the batched AMR step");
    if (levelmin < 1 || levelmax < levelmin || levelmax > s.nlevelmax || !head || !noct)
        return efail(RH_ERR_ARG, "rh_amr_prepare: bad level range");
    if (riemann < 1 || riemann > 3 || slope < 0 || slope > 3 || !(boxlen > 0))
        return efail(RH_ERR_ARG, "rh_amr_prepare: bad physics parameters");
    AmrPlan &m = s.amr;
    m.ready = false;
    m.levels.clear();
    int end = -1;
    std::vector<GodArgs> descs;
    std::vector<int2> items;
    std::vector<std::array<int, 3>> shapes;  // brick shape of each level's descriptor
    std::vector<const LevelBricks *> lbs;    // bricks of each level's descriptor
    const double gz[3] = {0, 0, 0};
    const double *gg = g ? g : gz;
    for (int l = levelmin; l <= levelmax; l++) {
        const int h = head[l - 1], n = noct[l - 1];
        if (n <= 0) continue;
        if (end >= 0 && h != end) return efail(RH_ERR_MESH, "the levels' octs must be contiguous, coarse to fine");
        if (end < 0) m.head0 = h;
        end = h + n;
        const double dx = boxlen / (double)(1ll << l);
        GodArgs a;
        const LevelBricks *lb = nullptr;
        if (int r = fill_args3(h, n, l, dx, gamma, smallr, smallc2, courant_factor, gg, slope, riemann, s.dt_dev, 0.0,
                               a, lb))
            return r;
        a.reflux = l > levelmin ? 1 : 0;
        shapes.push_back({lb->bx, lb->by, lb->bz});
        lbs.push_back(lb);
        descs.push_back(a);
        m.levels.push_back(l);
        m.dx_finest = dx;
    }
    if (m.levels.empty()) return efail(RH_ERR_ARG, "no octs in the level range");
    // items grouped by brick shape: one work-list launch per shape
    m.groups.clear();
    for (size_t d = 0; d < descs.size(); d++) {
        bool seen = false;
        for (size_t e = 0; e < d; e++) seen = seen || shapes[e] == shapes[d];
        if (seen) continue;
        AmrPlan::Group gr{shapes[d][0], shapes[d][1], shapes[d][2], (int)items.size(), 0};
        for (size_t e = d; e < descs.size(); e++)
            if (shapes[e] == shapes[d])
                for (int b = 0; b < descs[e].nbricks; b++) items.push_back(make_int2((int)e, b));
        gr.count = (int)items.size() - gr.first;
        // one range of the items per XCD, of equal modelled work: the bricks
        // come in z-slab order, and a thin shell's bricks march fewer planes
        // where it lies flat in z, so equal counts leave the XCDs unequal
        std::vector<double> pre((size_t)gr.count + 1, 0.0);
        for (int i = 0; i < gr.count; i++) {
            const int2 it = items[(size_t)gr.first + i];
            const LevelBricks &lb = *lbs[it.x];
            const int zr = lb.h_zr.size() > (size_t)it.y ? lb.h_zr[it.y] : ((lb.bz - 1) << 16);
            pre[i + 1] = pre[i] + brick_work(lb.bx, lb.by, 2.0 * ((zr >> 16) - (zr & 0xffff) + 1));
        }
        gr.range[0] = 0;
        gr.range[8] = gr.count;
        for (int r = 1; r < 8; r++) {
            const int q = (int)(std::lower_bound(pre.begin(), pre.end(), pre.back() * r / 8.0) - pre.begin());
            gr.range[r] = std::min(std::max(q, gr.range[r - 1]), gr.count);
        }
        m.groups.push_back(gr);
    }
    m.noct_all = end - m.head0;
    m.levelmin = levelmin;
    m.levelmax = levelmax;
    m.boxlen = boxlen;
    m.gamma = gamma;
    m.smallr = smallr;
    m.smallc2 = smallc2;
    m.courant = courant_factor;
    for (int d = 0; d < 3; d++) m.g[d] = gg[d];
    m.slope = slope;
    m.riemann = riemann;
    dfree(m.racc);
    dfree(m.d_descs);
    dfree(m.d_items);
    m.racc = nullptr;
    m.d_descs = nullptr;
    m.d_items = nullptr;
    // one accumulator per ghost oct, or one per face of each ghost oct in
    // deterministic mode (rh_set_deterministic)
    m.nslot = s.deterministic ? kRefluxSlotsPerGhost : 1;
    const long nacc = 5l * m.nslot * std::max(s.nghost_total, 1);
    RH_HIP_CHECK(dmalloc((void **)&m.racc, sizeof(double) * (size_t)nacc));
    // zeroed once: each step's fold leaves it zeroed again
    RH_HIP_CHECK(launch_zero(m.racc, nacc, s.stream));
    for (auto &a : descs) {
        a.racc = m.racc;
        a.rdet = s.deterministic;
    }
    m.ndesc = (int)descs.size();
    std::vector<GodArgs> both(2 * descs.size());
    for (size_t i = 0; i < descs.size(); i++) {
        GodArgs a = descs[i];
        a.uold = s.uold;
        a.unew = s.unew;
        a.ubase = s.uold;
        both[i] = a;
        a.uold = s.unew;
        a.unew = s.uold;
        a.ubase = s.unew;
        both[descs.size() + i] = a;
    }
    m.nitems = (int)items.size();
    m.h_descs = both;
    RH_HIP_CHECK(dmalloc((void **)&m.d_descs, both.size() * sizeof(GodArgs)));
    RH_HIP_CHECK(dmalloc((void **)&m.d_items, std::max<size_t>(items.size(), 1) * sizeof(int2)));
    RH_HIP_CHECK(hipMemcpyAsync(m.d_descs, both.data(), both.size() * sizeof(GodArgs), hipMemcpyHostToDevice, s.stream));
    RH_HIP_CHECK(hipMemcpyAsync(m.d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, s.stream));
    // the courant pass writes one partial per block, up to red_cap blocks
    if (int r = ensure_partials(1024)) return r;
    // dt of the first step from the current state; time starts at 0
    RH_HIP_CHECK(hipMemsetAsync(s.t_dev, 0, sizeof(double), s.stream));
    RH_HIP_CHECK(launch_courant_ml(s.uold, s.refined, s.lev, m.head0, m.noct_all, boxlen, m.dx_finest, gamma, smallr,
                                   __builtin_sqrt(smallc2), courant_factor, gg, s.red_part, s.red_cap, s.red_out,
                                   s.dt_dev, s.stream, nullptr));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    m.ready = true;
    return 0;
}

int rh_amr_run(int nsteps, int precision, int use_graph, double *dt_last, double *t_total, double sums[4],
               float *ms_per_step)
{
    State &s = st();
    AmrPlan &m = s.amr;
    if (!s.inited || !m.ready) return efail(RH_ERR_STATE, "rh_amr_prepare has not been called");
    if (nsteps < 0 || (precision != 0 && precision != 1)) return efail(RH_ERR_ARG, "rh_amr_run: bad arguments");
    RH_HIP_CHECK(hipEventRecord(s.ev0, s.stream));
    int done = 0;
    if (use_graph && nsteps >= 2) {
        hipGraph_t graph;
        hipGraphExec_t exec;
        RH_HIP_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_amr_step(s, 0, precision);
        if (e == hipSuccess) e = enqueue_amr_step(s, 1, precision);
        hipGraph_t g2 = nullptr;
        hipError_t e2 = hipStreamEndCapture(s.stream, &g2);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g2) (void)hipGraphDestroy(g2);
            return efail(RH_ERR_HIP, std::string("graph capture of the AMR step failed: ") +
                                         hipGetErrorString(e != hipSuccess ? e : e2));
        }
        graph = g2;
        RH_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        for (int i = 0; i < nsteps / 2; i++) RH_HIP_CHECK(hipGraphLaunch(exec, s.stream));
        RH_HIP_CHECK(hipStreamSynchronize(s.stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        done = 2 * (nsteps / 2);
    }
    // after an even number of steps the new state is back in uold
    int par = 0;
    for (int k = done; k < nsteps; k++) {
        RH_HIP_CHECK(enqueue_amr_step(s, par, precision));
        par ^= 1;
    }
    if (par) {
        // odd count: the state is in unew; swap the arrays and the two
        // descriptor halves so that parity 0 reads uold again
        std::swap(s.uold, s.unew);
        for (int i = 0; i < m.ndesc; i++) std::swap(m.h_descs[i], m.h_descs[m.ndesc + i]);
        RH_HIP_CHECK(hipMemcpyAsync(m.d_descs, m.h_descs.data(), m.h_descs.size() * sizeof(GodArgs),
                                    hipMemcpyHostToDevice, s.stream));
    }
    RH_HIP_CHECK(hipEventRecord(s.ev1, s.stream));
    RH_HIP_CHECK(hipEventSynchronize(s.ev1));
    float ms = 0;
    RH_HIP_CHECK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    if (ms_per_step) *ms_per_step = nsteps ? ms / nsteps : 0.0f;
    double v[7];
    RH_HIP_CHECK(hipMemcpy(v, s.red_out, 5 * sizeof(double), hipMemcpyDeviceToHost));
    RH_HIP_CHECK(hipMemcpy(v + 5, s.dt_dev, sizeof(double), hipMemcpyDeviceToHost));
    RH_HIP_CHECK(hipMemcpy(v + 6, s.t_dev, sizeof(double), hipMemcpyDeviceToHost));
    if (sums)
        for (int i = 0; i < 4; i++) sums[i] = v[i];
    if (dt_last) *dt_last = v[5];
    if (t_total) *t_total = v[6];
    return 0;
}

}  // extern "C"
