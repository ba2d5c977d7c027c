// Refinement flags of a level from its current state: hydro_flag
// (hydro/hydro_flag.f90:7-159) with the density and pressure gradient
// criteria of hydro_refine (hydro/godunov_utils.f90:6-227), one thread per
// cell, into a device-resident flag1 array, and the flag1 helpers of the
// reference's Metal bridge (metal/metal_interface.f90:225-241 and around).
//
// For each cell and each of its 2*ndim face neighbours, the neighbour cell
// (key at level l+1, periodic directions wrapped) is read from the level-l
// oct that holds it, or, when that oct is not one of the level's own, from
// the cell of level l-1 that holds it; keys outside the box read the
// boundary oct built from the reference oct inside (init_bound_refine).
// The three values of each direction (left, cell, right) go through
// hydro_refine's arithmetic in its order, with no FMA contraction, so that
// the flags agree with the reference's cell for cell.
#include <hip/hip_runtime.h>

#include "../../include/ramses2_hip.h"
#include "rh_gather.h"
#include "rh_mesh.h"
#include "rh_state.h"

namespace rh {

struct FlagArgs {
    const double *uold;
    const uint8_t *refined;
    int32_t *flag1;      // [ngridmax][8]
    const int32_t *ckey; // [ngrid][3]
    int head, noct;
    ResolveCtx fine, coarse;  // levels l and l-1 (coarse.nx == 0: none)
    int cmin[3], cmax[3];     // box at level l+1 (cell keys)
    int periodic[3];
    int refl;                 // reflexive faces (bit 2*dim+side)
    double gamma, smallr, err_d, err_p, floor_d, floor_p;
    int *counter;             // [0] cells newly flagged, [1] a neighbour's cell missing
};

// hydro_refine's density and pressure tests for one direction (ug, um, ud:
// left, cell, right conservative states), in the reference's order.
__device__ bool refine_dp(const FlagArgs &a, double ug[5], double um[5], double ud[5])
{
#pragma clang fp contract(off)
    ug[0] = ug[0] > a.smallr ? ug[0] : a.smallr;
    um[0] = um[0] > a.smallr ? um[0] : a.smallr;
    ud[0] = ud[0] > a.smallr ? ud[0] : a.smallr;
    for (int d = 1; d <= 3; d++) {
        ug[d] = ug[d] / ug[0];
        um[d] = um[d] / um[0];
        ud[d] = ud[d] / ud[0];
    }
    double eking = 0.0, ekinm = 0.0, ekind = 0.0;
    for (int d = 1; d <= 3; d++) {
        eking = eking + 0.5 * ug[0] * (ug[d] * ug[d]);
        ekinm = ekinm + 0.5 * um[0] * (um[d] * um[d]);
        ekind = ekind + 0.5 * ud[0] * (ud[d] * ud[d]);
    }
    const double pg = (a.gamma - 1.0) * (ug[4] - eking);
    const double pm = (a.gamma - 1.0) * (um[4] - ekinm);
    const double pd = (a.gamma - 1.0) * (ud[4] - ekind);
    bool ok = false;
    if (a.err_d >= 0.0) {
        const double x = __builtin_fabs((ud[0] - um[0]) / (ud[0] + um[0] + a.floor_d));
        const double y = __builtin_fabs((um[0] - ug[0]) / (um[0] + ug[0] + a.floor_d));
        ok = ok || 2.0 * (x > y ? x : y) > a.err_d;
    }
    if (a.err_p >= 0.0) {
        const double x = __builtin_fabs((pd - pm) / (pd + pm + a.floor_p));
        const double y = __builtin_fabs((pm - pg) / (pm + pg + a.floor_p));
        ok = ok || 2.0 * (x > y ? x : y) > a.err_p;
    }
    return ok;
}

__global__ void __launch_bounds__(256) k_hydro_flag(const FlagArgs a)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)a.noct * 8) return;
    const int o = a.head - 1 + (int)(i >> 3), ind = (int)(i & 7);
    int hk[3];
    for (int d = 0; d < 3; d++) hk[d] = 2 * a.ckey[3 * o + d] + ((ind >> d) & 1);
    double un[6][5];
    for (int nb = 0; nb < 6; nb++) {
        int kn[3] = {hk[0], hk[1], hk[2]};
        const int d0 = nb >> 1;
        kn[d0] += (nb & 1) ? 1 : -1;
        if (a.periodic[d0]) {
            if (kn[d0] < a.cmin[d0]) kn[d0] = a.cmax[d0] - 1;
            if (kn[d0] >= a.cmax[d0]) kn[d0] = a.cmin[d0];
        }
        int kp[3], c[3];
        for (int d = 0; d < 3; d++) {
            kp[d] = kn[d] >> 1;
            c[d] = kn[d] & 1;
        }
        int v = resolve_oct(a.fine, kp[0], kp[1], kp[2]);
        const int *bmin = a.fine.boxmin, *bmax = a.fine.boxmax;
        if (v > a.fine.ngridmax || v == 0) {
            // not one of the level's octs: the level l-1 cell holding the neighbour
            int kq[3];
            for (int d = 0; d < 3; d++) {
                c[d] = kp[d] & 1;
                kq[d] = kp[d] >> 1;
            }
            v = a.coarse.nx ? resolve_oct(a.coarse, kq[0], kq[1], kq[2]) : 0;
            if (v > a.coarse.ngridmax || v == 0) {
                atomicOr(&a.counter[1], 1);
                return;
            }
            for (int d = 0; d < 3; d++) kp[d] = kq[d];
            bmin = a.coarse.boxmin;
            bmax = a.coarse.boxmax;
        }
        const CellRef r = cell_ref3(a.uold, a.refined, v, kp, bmin, bmax, a.refl, c);
        uint8_t rf;
        cell_load3(r, un[nb], rf);
        cell_flip(un[nb], r.flip);
    }
    double um0[5];
    for (int n = 0; n < 5; n++) um0[n] = a.uold[(size_t)o * 40 + n * 8 + ind];
    bool ok = false;
    for (int d = 0; d < 3; d++) {
        double ug[5], um[5], ud[5];
        for (int n = 0; n < 5; n++) {
            ug[n] = un[2 * d][n];
            um[n] = um0[n];
            ud[n] = un[2 * d + 1][n];
        }
        ok = refine_dp(a, ug, um, ud) || ok;
    }
    if (ok) {
        int32_t *f = a.flag1 + (size_t)o * 8 + ind;
        if (*f == 0) atomicAdd(&a.counter[0], 1);
        *f = 1;
    }
}

__global__ void k_fill_flag1(int32_t *flag1, long first, long n, int32_t v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag1[first + i] = v;
}

__global__ void k_count_flag1(const int32_t *flag1, long first, long n, int *count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int v = i < n && flag1[first + i] != 0 ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(count, v);
}

}  // namespace rh

using namespace rh;

namespace {

int efail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int ensure_flag1()
{
    State &s = st();
    if (!s.inited || !s.uold) return efail(RH_ERR_STATE, "library not initialised / mesh not allocated");
    if (s.ndim != 3) return efail(RH_ERR_ARG, "refinement flags are computed for 3D meshes");
    if (!s.flag1) {
        const size_t n = (size_t)s.ngridmax * 8;
        RH_HIP_CHECK(dmalloc((void **)&s.flag1, n * sizeof(int32_t)));
        hipLaunchKernelGGL(k_fill_flag1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, s.flag1, 0l,
                           (long)n, 0);
        RH_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

int oct_range(int head_idx, int num_octs)
{
    const State &s = st();
    if (head_idx < 1 || num_octs < 0 || (int64_t)head_idx - 1 + num_octs > s.ngridmax)
        return efail(RH_ERR_ARG, "oct range out of bounds (flags cover octs 1..ngridmax)");
    return 0;
}

ResolveCtx level_ctx(int l)
{
    const State &s = st();
    ResolveCtx c;
    c.keys = s.hkeys;
    c.vals = s.hvals;
    c.cap = (uint64_t)s.hash_cap;
    c.koff = s.key_off[l - 1];
    c.nx = s.ckey_max[l - 1];
    c.ndim = s.ndim;
    c.ngridmax = s.ngridmax;
    for (int d = 0; d < 3; d++) {
        c.boxmin[d] = s.box_min[3 * (l - 1) + d];
        c.boxmax[d] = s.box_max[3 * (l - 1) + d];
        c.periodic[d] = s.periodic[d];
    }
    return c;
}

}  // namespace

extern "C" {

int rh_reset_flag1(int head_idx, int num_octs)
{
    if (int r = ensure_flag1()) return r;
    if (int r = oct_range(head_idx, num_octs)) return r;
    State &s = st();
    if (num_octs == 0) return 0;
    const long n = (long)num_octs * 8;
    hipLaunchKernelGGL(k_fill_flag1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, s.flag1,
                       (long)(head_idx - 1) * 8, n, 0);
    RH_HIP_CHECK(hipGetLastError());
    return 0;
}

int rh_hydro_flag(int head_idx, int num_octs, double gamma, double smallr, double smallc2, double err_grad_d,
                  double err_grad_p, double floor_d, double floor_p)
{
    (void)smallc2;
    if (int r = ensure_flag1()) return r;
    if (int r = oct_range(head_idx, num_octs)) return r;
    State &s = st();
    if (num_octs == 0) return 0;
    if (err_grad_d == -1.0 && err_grad_p == -1.0) return 0;  // no gradient criterion (hydro_flag.f90:49-62)
    if (!(gamma > 1.0)) return efail(RH_ERR_ARG, "rh_hydro_flag: gamma must exceed 1");
    const int l = s.h_lev[head_idx - 1];
    if (l < 1 || l > s.nlevelmax) return efail(RH_ERR_MESH, "oct level outside the level parameters");
    for (int i = head_idx - 1; i < head_idx - 1 + num_octs; i++)
        if (s.h_lev[i] != l) return efail(RH_ERR_MESH, "octs of the range are not all on one level");
    FlagArgs a;
    a.uold = s.uold;
    a.refined = s.refined;
    a.flag1 = s.flag1;
    a.ckey = s.ckey;
    a.head = head_idx;
    a.noct = num_octs;
    a.fine = level_ctx(l);
    if (l > 1)
        a.coarse = level_ctx(l - 1);
    else
        a.coarse.nx = 0;
    for (int d = 0; d < 3; d++) {
        a.cmin[d] = s.box_min[3 * l + d];  // box of level l+1
        a.cmax[d] = s.box_max[3 * l + d];
        a.periodic[d] = s.periodic[d];
    }
    a.refl = 0;
    for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) a.refl |= (s.btype[d][sd] == 1) << (2 * d + sd);
    a.gamma = gamma;
    a.smallr = smallr;
    a.err_d = err_grad_d;
    a.err_p = err_grad_p;
    a.floor_d = floor_d;
    a.floor_p = floor_p;
    a.counter = s.d_counter;
    RH_HIP_CHECK(hipMemsetAsync(s.d_counter, 0, 2 * sizeof(int), s.stream));
    const long n = (long)num_octs * 8;
    hipLaunchKernelGGL(k_hydro_flag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, a);
    RH_HIP_CHECK(hipGetLastError());
    int hc[2] = {0, 0};
    RH_HIP_CHECK(hipMemcpyAsync(hc, s.d_counter, 2 * sizeof(int), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    if (hc[1]) return efail(RH_ERR_MESH, "hydro_flag: a neighbour cell has no parent cell on the coarser level");
    return 0;
}

int rh_count_flag1(int head_idx, int num_octs, int32_t *count)
{
    if (int r = ensure_flag1()) return r;
    if (int r = oct_range(head_idx, num_octs)) return r;
    if (!count) return efail(RH_ERR_ARG, "rh_count_flag1: null count");
    State &s = st();
    *count = 0;
    if (num_octs == 0) return 0;
    const long n = (long)num_octs * 8;
    RH_HIP_CHECK(hipMemsetAsync(s.d_counter, 0, sizeof(int), s.stream));
    hipLaunchKernelGGL(k_count_flag1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s.stream, s.flag1,
                       (long)(head_idx - 1) * 8, n, s.d_counter);
    RH_HIP_CHECK(hipGetLastError());
    int c = 0;
    RH_HIP_CHECK(hipMemcpyAsync(&c, s.d_counter, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    *count = c;
    return 0;
}

int rh_upload_flag1(const int32_t *flag1, int ngridmax)
{
    if (int r = ensure_flag1()) return r;
    State &s = st();
    if (!flag1 || ngridmax != s.ngridmax) return efail(RH_ERR_ARG, "rh_upload_flag1: flag1(8, ngridmax) expected");
    RH_HIP_CHECK(hipMemcpyAsync(s.flag1, flag1, (size_t)ngridmax * 8 * sizeof(int32_t), hipMemcpyHostToDevice,
                                s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

int rh_transfer_flag1_host(int32_t *flag1, int ngridmax)
{
    if (int r = ensure_flag1()) return r;
    State &s = st();
    if (!flag1 || ngridmax != s.ngridmax) return efail(RH_ERR_ARG, "rh_transfer_flag1_host: flag1(8, ngridmax) expected");
    RH_HIP_CHECK(hipMemcpyAsync(flag1, s.flag1, (size_t)ngridmax * 8 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                s.stream));
    RH_HIP_CHECK(hipStreamSynchronize(s.stream));
    return 0;
}

}  // extern "C"
