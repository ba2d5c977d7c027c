// gfx950 kernels of the fine-level Godunov hydro step.
//
// 3D: one workgroup per oct brick (BX*BY*BZ octs = 2BX x 2BY x 2BZ cells).
// The workgroup marches through the brick along z, one cell plane per
// iteration.  One thread owns one cell column of the (2BX+2) x (2BY+2) trace
// region (the brick's cells plus one cell of halo).  Per plane:
//   P: gather plane t of the (2BX+4) x (2BY+4) primitive region from uold
//      through the brick window (hash-resolved oct indices, boundary ghosts
//      synthesised on the fly), convert to primitives in FP64, keep a ring of
//      two planes in LDS (three for the 27-point slope), the next plane's
//      values already in flight (prefetched a plane ahead);
//   T: slopes (uslope) and Hancock predictor (trace3d) of plane t-1, once per
//      cell; left states of x/y faces go to LDS, z states stay in registers;
//   F: each column solves the Riemann problem on its own low x, y and z face
//      (cmpflxm), so every face is solved once; refined faces get zero flux;
//   U: conservative update u + (dt/dx)(F(i) - F(i+1)), dimension by dimension
//      in the order of godfine1 (hydro/godunov_fine.f90:737-816), written to
//      unew; the courant_fine sums and dt of the new state are reduced from
//      registers as it is stored.
// No MFMA: the work is divides, square roots and compares, not a contraction.
#include <hip/hip_ext.h>

#include <map>
#include <type_traits>

#include "rh_gather.h"
#include "rh_godunov.h"

namespace rh {

struct CourantK {
    double dx, vol, gamma, smallr, smallc, courant_factor, smallp, gsum_dx;
    int ndim;
};

// The gravity factor 1 + sqrt(1 + 2 C max(g dx/c_tot^2, 1e-4)) of cmpdt when
// g = 0: a constant, formed once per thread with the expression cell_dt uses.
__device__ __forceinline__ double grav_factor0(const CourantK &c)
{
    return 1.0 + __builtin_sqrt(1.0 + 2.0 * c.courant_factor * 0.0001);
}

// cmpdt, hydro/godunov_utils.f90:228-337, for one cell (no MHD).  ir = 1/r
// (the Newton-refined reciprocal) replaces the reference's three divisions by
// r, and with no magnetic field c_f = sqrt(c2 + sqrt(c2^2)) with c2 = a^2/2 is
// exactly sqrt(a^2); gravity's correction is the constant gf0
// (grav_factor0) when g = 0.
__device__ __forceinline__ double cell_dt(double uu0, double uu1, double uu2, double uu3, double uu4, const CourantK &c,
                                          double gf0)
{
    double r = uu0 > c.smallr ? uu0 : c.smallr;
    const double ir = rcp_nr(r);
    double v[3] = {uu1 * ir, uu2 * ir, uu3 * ir};
    double e = uu4;
    e = e - 0.5 * r * (v[0] * v[0]);
    e = e - 0.5 * r * (v[1] * v[1]);
    e = e - 0.5 * r * (v[2] * v[2]);
    double pe = (c.gamma - 1.0) * e, pf = r * c.smallp;
    double p = pe > pf ? pe : pf;
    double a2 = c.gamma * p;
    a2 = a2 * ir;
    // a2 >= gamma smallp = smallc^2: the square root of the Riemann solvers
    // (floored where smallc^2 < 2^-900)
    double cf = sqrt_pos(a2 > 0x1p-900 ? a2 : 0x1p-900);
    double ctot = 0.0;
    for (int d = 0; d < c.ndim; d++) ctot = ctot + __builtin_fabs(v[d]) + cf;
    double gf = gf0;
    if (c.gsum_dx != 0.0) {
        double grav = c.gsum_dx / (ctot * ctot);
        grav = grav > 0.0001 ? grav : 0.0001;
        gf = 1.0 + __builtin_sqrt(1.0 + 2.0 * c.courant_factor * grav);
    }
    double dtmax = c.courant_factor * c.dx / c.smallc;
    double dtcell = 2.0 * c.courant_factor * c.dx / (ctot * gf);
    return dtmax < dtcell ? dtmax : dtcell;
}

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ double wave_min(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double y = __shfl_xor(x, o, 64);
        x = y < x ? y : x;
    }
    return x;
}
// The same for a wave of which only lanes [0, nvalid) exist (the last wave of
// a workgroup whose size is not a multiple of 64): partners at or beyond
// nvalid are skipped.  Lane 0's partners in the round of offset o are lanes
// o..2o-1, whose partial results only gather lanes above them, so lane 0
// ends with the reduction over exactly the existing lanes.
__device__ __forceinline__ double wave_sum_n(double x, int nvalid)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(x, o, 64);
        if ((lane ^ o) < nvalid) x += y;
    }
    return x;
}
__device__ __forceinline__ double wave_min_n(double x, int nvalid)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(x, o, 64);
        if ((lane ^ o) < nvalid && y < x) x = y;
    }
    return x;
}

__device__ __forceinline__ CourantK make_courant_k(double dx, int ndim, double gamma, double smallr, double smallc2,
                                                   double courant_factor, const double g[3])
{
    CourantK c;
    c.dx = dx;
    c.vol = ndim == 1 ? dx : dx * dx * dx;
    c.gamma = gamma;
    c.smallr = smallr;
    c.smallc = __builtin_sqrt(smallc2);
    c.courant_factor = courant_factor;
    c.smallp = smallc2 / gamma;
    double gs = 0.0;
    for (int d = 0; d < ndim; d++) gs = gs + __builtin_fabs(g[d]);
    c.gsum_dx = gs * dx;
    c.ndim = ndim;
    return c;
}

// One leaf cell of courant_fine (hydro/courant_fine.f90:120-160).
__device__ __forceinline__ void courant_cell(const double u[5], const CourantK &c, double gf0, double &mass,
                                             double &ekin, double &eint, double &dtmin)
{
    mass = mass + u[0] * c.vol;
    ekin = ekin + u[4] * c.vol;
    eint = eint + u[4] * c.vol;
    const double rr = u[0] > c.smallr ? u[0] : c.smallr;
    const double irr = rcp_nr(rr);
    eint = eint - 0.5 * (u[1] * u[1]) * irr * c.vol;
    eint = eint - 0.5 * (u[2] * u[2]) * irr * c.vol;
    eint = eint - 0.5 * (u[3] * u[3]) * irr * c.vol;
    const double d = cell_dt(u[0], u[1], u[2], u[3], u[4], c, gf0);
    dtmin = d < dtmin ? d : dtmin;
}

// courant_fine of one cell for the mixed-precision kernel: the conservation
// sums and the pressure recovery p=(gamma-1)(E-rho v^2/2) in FP64, the signal
// speed and the cell's dt in FP32 (cmpdt's c_f = sqrt(c2+sqrt(c2^2)) with
// c2 = a^2/2 and no magnetic field is a), the dt rounded to FP64 for the
// min-reduction.
__device__ __forceinline__ void courant_cell_mixed(const double u[5], const CourantK &c, double &mass, double &ekin,
                                                   double &eint, double &dtmin)
{
    mass = mass + u[0] * c.vol;
    ekin = ekin + u[4] * c.vol;
    const double rr = u[0] > c.smallr ? u[0] : c.smallr;
    const double ir = rcp_nr(rr);
    const double ke = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * ir;
    eint = eint + (u[4] - ke) * c.vol;
    const double pe = (c.gamma - 1.0) * (u[4] - ke), pf = rr * c.smallp;
    const double p = pe > pf ? pe : pf;
    const float irf = (float)ir;
    const float cf = __builtin_amdgcn_sqrtf((float)(c.gamma * p) * irf);
    const float ctot = (__builtin_fabsf((float)u[1] * irf) + cf) + (__builtin_fabsf((float)u[2] * irf) + cf) +
                       (__builtin_fabsf((float)u[3] * irf) + cf);
    const float cfl = (float)c.courant_factor;
    float grav = (float)c.gsum_dx * __builtin_amdgcn_rcpf(ctot * ctot);
    grav = grav > 0.0001f ? grav : 0.0001f;
    const double dtmax = c.courant_factor * c.dx / c.smallc;
    const float dtcell = 2.0f * cfl * (float)c.dx *
                         __builtin_amdgcn_rcpf(ctot * (1.0f + __builtin_amdgcn_sqrtf(1.0f + 2.0f * cfl * grav)));
    const double d = dtmax < (double)dtcell ? dtmax : (double)dtcell;
    dtmin = d < dtmin ? d : dtmin;
}

// Gather the conservative state of cell (i2,j2,k2) of the oct at window slot
// value v, whose oct key is okey.  Boundary ghosts (v<0) follow
// init_bound_refine (amr/boundaries.f90:39-165): zero gradient copies the
// adjacent interior cell, reflexive mirrors it and flips the normal momentum.
// refl: bit 2*d+side set where that face of the domain is reflexive
// (bound_type 1), else zero gradient (bound_type 2).
__device__ __forceinline__ void gather_cell_m(const double *__restrict__ uold, const uint8_t *__restrict__ refined,
                                              int v, const int okey[3], int ndim, const int boxmin[3],
                                              const int boxmax[3], int refl, int c[3], double u[5], uint8_t &rf)
{
    if (v > 0) {
        const int ind = c[0] + 2 * c[1] + 4 * c[2];
        const double *p = uold + (size_t)(v - 1) * 40 + ind;
        if (ndim == 1) p = uold + (size_t)(v - 1) * 10 + ind;
        const int stride = ndim == 1 ? 2 : 8;
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = p[n * stride];
        rf = refined[(size_t)(v - 1) * 8 + ind];
        return;
    }
    if (v == 0) {
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = 0.0;
        rf = 0;
        return;
    }
    // boundary ghost: first direction in which the oct key leaves the domain
    int dir = 0, side = 0;
    for (int d = ndim - 1; d >= 0; d--) {
        if (okey[d] < boxmin[d]) { dir = d; side = 0; }
        else if (okey[d] >= boxmax[d]) { dir = d; side = 1; }
    }
    const int type = (refl >> (2 * dir + side)) & 1 ? 1 : 2;
    int s[3] = {c[0], c[1], c[2]};
    uint8_t rflag_src_valid;
    int rs[3] = {c[0], c[1], c[2]};
    if (side == 0) {
        // reference oct lies at +1: refined(ind) = ref cell with c_dir-1 (ref_right)
        rflag_src_valid = c[dir] == 1;
        rs[dir] = 0;
    } else {
        rflag_src_valid = c[dir] == 0;
        rs[dir] = 1;
    }
    if (type == 1)
        s[dir] = 1 - c[dir];
    else
        s[dir] = side == 0 ? 0 : 1;
    const int ref = -v;
    const int ind = s[0] + 2 * s[1] + 4 * s[2];
    const int stride = ndim == 1 ? 2 : 8;
    const double *p = uold + (size_t)(ref - 1) * 5 * stride + ind;
#pragma unroll
    for (int n = 0; n < 5; n++) u[n] = p[n * stride];
    if (type == 1) u[1 + dir] = -u[1 + dir];
    rf = rflag_src_valid ? refined[(size_t)(ref - 1) * 8 + rs[0] + 2 * rs[1] + 4 * rs[2]] : 0;
}

__device__ __forceinline__ void gather_cell(const double *__restrict__ uold, const uint8_t *__restrict__ refined,
                                            int v, const int okey[3], int ndim, const int boxmin[3],
                                            const int boxmax[3], const int btype[3][2], int c[3], double u[5],
                                            uint8_t &rf)
{
    gather_cell_m(uold, refined, v, okey, ndim, boxmin, boxmax, reflexive_mask(btype), c, u, rf);
}

template <typename T>
__device__ __forceinline__ HydroK<T> make_k(const double dt, double dx, double gamma, double smallr, double smallc2,
                                            double courant_factor, double smax_cap, int use_smax, int slope,
                                            int riemann)
{
    HydroK<T> k;
    k.gamma = T(gamma);
    k.smallr = T(smallr);
    k.smallc2 = T(smallc2 > 0x1p-900 ? smallc2 : 0x1p-900);
    k.smallp_trace = T(smallr * smallc2);
    k.smallp_riem = T(smallc2 / gamma);
    k.entho = T(1.0 / (gamma - 1.0));
    k.dtdx = T(dt / dx);
    k.smax_cap = T(use_smax ? smax_cap : courant_factor * dx / dt);
    k.slope_type = slope;
    k.riemann = riemann;
    return k;
}

// Positivity-preserving 27-point slope (slope_type 3, umuscl.f90:2207-2270)
// for variable n of the cell at ring index (s0, pc), returned halved like
// slope_half: the centred differences are formed as (q+ - q-)/4 = dfx/2, so
// the sum dff and the products slop*dfx/2 are the reference's values halved
// exactly.
template <typename T, int PX, int NP>
__device__ __forceinline__ void slope27(const T (*sq)[5][NP], int sm, int s0, int sp, int n, int pc, T dq3[3])

{
    const T qc = sq[s0][n][pc];
    T vmin = T(0), vmax = T(0);
    bool first = true;
    const int sl[3] = {sm, s0, sp};
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
        for (int jj = -1; jj <= 1; jj++)
#pragma unroll
            for (int ii = -1; ii <= 1; ii++) {
                T df = sq[sl[kk]][n][pc + ii + jj * PX] - qc;
                if (first) {
                    vmin = vmax = df;
                    first = false;
                } else {
                    vmin = tmin(vmin, df);
                    vmax = tmax(vmax, df);
                }
            }
    T dfx = T(0.25) * (sq[s0][n][pc + 1] - sq[s0][n][pc - 1]);
    T dfy = T(0.25) * (sq[s0][n][pc + PX] - sq[s0][n][pc - PX]);
    T dfz = T(0.25) * (sq[sp][n][pc] - sq[sm][n][pc]);
    T dff = fabs_(dfx) + fabs_(dfy) + fabs_(dfz);
    T slop = dff > T(0) ? tmin(T(1), tdiv(tmin(fabs_(vmin), fabs_(vmax)), dff)) : T(1);
    dq3[0] = slop * dfx;
    dq3[1] = slop * dfy;
    dq3[2] = slop * dfz;
}

// Coarse-fine refluxing, godfine1 (hydro/godunov_fine.f90:894-1052): a face
// between a cell of the level and an interpolated ghost cell sends -flux/2^ndim
// (ghost on the low side) or +flux/2^ndim (ghost on the high side) to the
// ghost's parent cell.  Three ways:
// - racc null: FP64 atomics into the parent cell of unew;
// - racc, rdet 0: FP64 atomics into the ghost's accumulator racc[5 g];
// either way the order of the contributions to one parent cell is not fixed
// (last-bit differences from run to run and from the serial reference);
// - racc, rdet 1 (deterministic mode): a plain store into the face's own slot
//   racc[5 (kReluxSlots g + slot)], slot = 4 side + sub (side: the ghost oct's
//   face, -x +x -y +y -z +z; sub: which of its 2^(ndim-1) cells), one writer
//   per slot; k_reflux_fold then adds the slots in slot order.
constexpr int kRefluxSlots = kRefluxSlotsPerGhost;
__device__ __forceinline__ void reflux_add(double *unew, const int32_t *ghost_parent, int ngridmax, int vghost,
                                           double sign, const double F[5], int ttd, double *racc = nullptr,
                                           int rdet = 0, int slot = 0, bool accumulate = false)
{
    const int gi = vghost - ngridmax - 1;
    if (racc) {
        const double w = 1.0 / (double)ttd;
        if (rdet) {
            // accumulate: the diffusion pass adds to the slot the Godunov launch wrote
            double *q = racc + 5 * ((size_t)gi * kRefluxSlots + slot);
#pragma unroll
            for (int n = 0; n < 5; n++) q[n] = accumulate ? q[n] + sign * (F[n] * w) : sign * (F[n] * w);
            return;
        }
#pragma unroll
        for (int n = 0; n < 5; n++) atomicAdd(racc + 5 * (size_t)gi + n, sign * (F[n] * w));
        return;
    }
    const int igrid = ghost_parent[2 * gi], icell = ghost_parent[2 * gi + 1];
    if (igrid <= 0) return;  // parent missing: rh_build_nbor reported it
    const double w = 1.0 / (double)ttd;
    double *p = unew + (size_t)(igrid - 1) * 5 * ttd + (icell - 1);
#pragma unroll
    for (int n = 0; n < 5; n++) atomicAdd(p + n * ttd, sign * (F[n] * w));
}

// The summed refluxes q (already weighted by 1/2^ndim, with their signs) of
// one ghost oct: into its accumulator, or into its parent cell of unew.
__device__ __forceinline__ void reflux_flush(double *unew, const int32_t *ghost_parent, int ngridmax, int vghost,
                                             const double q[5], int ttd, double *racc)
{
    const int gi = vghost - ngridmax - 1;
    if (racc) {
#pragma unroll
        for (int n = 0; n < 5; n++) atomicAdd(racc + 5 * (size_t)gi + n, q[n]);
        return;
    }
    const int igrid = ghost_parent[2 * gi], icell = ghost_parent[2 * gi + 1];
    if (igrid <= 0) return;
    double *p = unew + (size_t)(igrid - 1) * 5 * ttd + (icell - 1);
#pragma unroll
    for (int n = 0; n < 5; n++) atomicAdd(p + n * ttd, q[n]);
}

// Loads of the next plane's P-phase cells issued a plane ahead, per precision
// (FP32: the per-level kernel only; it runs at two waves per SIMD with them,
// measured 4% faster at 256^3 than at three without, while the work-list
// kernel keeps four waves per SIMD).
// The z march unrolled by two planes.
#ifndef RH_MARCH_UNROLL2
#define RH_MARCH_UNROLL2 0
#endif
// Levels periodic in every dimension gather through a path without the
// boundary-ghost synthesis (a uniform branch per launch).
#ifndef RH_FAST_FETCH
#define RH_FAST_FETCH 1
#endif
#ifndef RH_PREFETCH_F64
#define RH_PREFETCH_F64 1
#endif
#ifndef RH_PREFETCH_F32
#define RH_PREFETCH_F32 1
#endif

// unew stored one oct layer (two z-planes) at a time, in whole 64-byte lines
// (pair_writes).
#ifndef RH_PAIR_WRITES
#define RH_PAIR_WRITES 1
#endif
// Whole-line stores: the FP64 kernels (writes 1.30x -> 0.97x of the
// algorithmic bytes, the same time), the FP32 per-level kernel and the FP32
// specialised work-list kernel (serving): the mixed kernel 0.624 -> 0.581 ms
// at 256^3, serving's mixed capacity +7%; the FP32 general work-list kernel
// (AMR) measured 7.6% slower with them and stores each half as it comes.
template <typename T, bool WL, bool GEN> constexpr bool pair_writes()
{
    return RH_PAIR_WRITES && (sizeof(T) == 8 || !WL || !GEN);
}
// Refluxes summed in LDS per window oct before one atomic per ghost oct and
// variable (instead of one per fine face), for bricks whose window has at
// most this many octs (4x4x4 bricks: 216 window octs, 8.6 KB of FP64 sums).
#ifndef RH_REFLUX_LDS_MAX
#define RH_REFLUX_LDS_MAX 512
#endif
// The update's base values (the cell's uold) taken from the P phase's loads
// through LDS instead of a second global load at the update, where the
// update adds to uold (device-resident steps).
#ifndef RH_UB_LDS
#define RH_UB_LDS 1
#endif
// The FP32 per-level kernel's Courant reduction in the update phase (as the
// FP64 kernel does) instead of re-reading the brick after the march: it costs
// the kernel a wave per SIMD, which the uniform level step recovers (step
// 0.93 -> 0.87 ms at 256^3) and the work-list kernel of the AMR and serving
// steps does not (AMR mixed step 5% slower), so the latter keeps the re-read.
#ifndef RH_COURANT_IN_UPDATE_F32
#define RH_COURANT_IN_UPDATE_F32 1
#endif

// The FP32 per-level kernel of 8x6x16 bricks runs at three waves per SIMD:
// it fits in 168 VGPRs without spilling when the update's base values are
// loaded at the update (1.5% faster than two waves with them loaded before the
// slopes, and 4.6% faster than two waves with them loaded at the update).  The
// other FP32 shapes spill at three waves and keep two, with the early loads.
template <int BX, int BY, int BZ, typename T, bool WL>
constexpr bool f32_three_waves()
{
    return sizeof(T) == 4 && !WL && BX == 8 && BY == 6 && BZ == 16;
}

// Minimum waves per SIMD requested from the register allocator for the brick
// kernels (occupancy vs. register spills; chosen by measurement).
#ifndef RH_BRICK_MIN_WAVES_F32
#define RH_BRICK_MIN_WAVES_F32 2
#endif
#ifndef RH_BRICK_MIN_WAVES
#define RH_BRICK_MIN_WAVES 2
#endif

// Waves per SIMD a brick kernel runs at (its launch bound): the FP32
// work-list kernel at four, at most 128 VGPRs (the AMR 4x4x4 kernel spills 4
// of its 131 for it, and the mixed AMR step is 9% faster: 0.670 -> 0.612 ms).
template <int BX, int BY, int BZ, typename T, bool WL> constexpr int brick_waves()
{
    return sizeof(T) == 8 ? RH_BRICK_MIN_WAVES
                          : (WL ? 4 : (f32_three_waves<BX, BY, BZ, T, WL>() ? 3 : RH_BRICK_MIN_WAVES_F32));
}

// Halo-wave thread order: the 4 BX BY interior columns fill whole waves and
// the halo columns fit one more (8x6 bricks, 192 + 60 threads; 4x4, 64 + 36).
template <int BX, int BY, int RS, int NS> constexpr bool brick_hw()
{
    return (4 * BX * BY) % 64 == 0 && (2 * BX + 2) * (2 * BY + 2) - 4 * BX * BY <= 64 && RS == 3 && NS == 2;
}

// The next-plane prefetch: the HLLC variants of the FP64 kernels and of the
// FP32 per-level kernel, with a thread for every cell of the P-phase ring.
template <int BX, int BY, typename T, int RS, int NS, bool WL> constexpr bool brick_prefetch()
{
    constexpr int CX = 2 * BX + 2, CY = 2 * BY + 2, NT = CX * CY, NI = 4 * BX * BY;
    constexpr bool HW = brick_hw<BX, BY, RS, NS>();
    constexpr int NRING = HW ? 4 * (BX + BY) : 2 * (CX + CY);
    return (sizeof(T) == 8 ? RH_PREFETCH_F64 : RH_PREFETCH_F32 && !WL) && RS == 3 && NS == 2 &&
           (HW ? NRING <= NT - NI : NRING <= NT);
}

// LDS of one brick workgroup.  NS primitive planes are kept: 2 when the slopes
// need only face neighbours (the z neighbours of a column sit in registers),
// 3 for the 27-point slope of slope_type 3.
// SEP: face fluxes get their own buffer (sff) instead of overwriting the
// face states in sf, which saves one barrier per plane.  Measured: no gain
// (256^3 FP64 and mixed within 1%), so off by default; the barriers are not
// what limits the kernel.
template <int BX, int BY, int BZ, typename T, int NS, bool SEP, bool PWT> struct BrickSmemCore {
    static constexpr int CX = 2 * BX + 2, CY = 2 * BY + 2, NT = CX * CY;
    static constexpr int PX = 2 * BX + 4, PY = 2 * BY + 4, NP = PX * PY;
    static constexpr int WX = BX + 2, WY = BY + 2, WZ = BZ + 2, NW = WX * WY * WZ;
    int32_t swin[NW];
    T sq[NS][5][NP];
    uint8_t sref[NS][NP];
    T sf[2][5][NT];
    T sff[SEP ? 2 : 1][5][SEP ? NT : 1];
    // new state of the lower z-plane of the oct layer, per interior column
    // (RH_PAIR_WRITES)
    static constexpr bool PW = PWT;
    double sun[PW ? 5 : 1][PW ? 4 * BX * BY : 1];
};
// The optional buffers are added only while the workgroups the kernel's
// waves per SIMD call for (WPS) still fit the 160 KB of LDS of a CU, so that
// they never cost occupancy:
// - sacc (RA): the coarse-fine refluxes of the brick summed per window oct,
//   flushed once per ghost oct at the end (windows of at most
//   RH_REFLUX_LDS_MAX octs);
// - sub (UBL, kernels with the next-plane prefetch, UBOK): the conserved
//   state of each interior column's cell on the last two planes, kept from
//   the P phase's loads as the update's base values.
template <int BX, int BY, int BZ, typename T, int NS, bool SEP, int WPS, bool UBOK, bool PWT>
struct BrickSmem : BrickSmemCore<BX, BY, BZ, T, NS, SEP, PWT> {
    using Core = BrickSmemCore<BX, BY, BZ, T, NS, SEP, PWT>;
    static constexpr int NW = Core::NW, NT = Core::NT;
    static constexpr int WGS = (4 * WPS) / ((NT + 63) / 64) > 0 ? (4 * WPS) / ((NT + 63) / 64) : 1;
    static constexpr size_t LIM = (size_t)160 * 1024 / WGS;
    static constexpr bool RA = NW <= RH_REFLUX_LDS_MAX && sizeof(Core) + (size_t)NW * 40 <= LIM;
    static constexpr size_t RAB = RA ? (size_t)NW * 40 : 40;
    static constexpr bool UBL =
        UBOK && RH_UB_LDS && NT <= 256 && sizeof(Core) + RAB + (size_t)80 * 4 * BX * BY <= LIM;
    double sacc[RA ? NW : 1][5];
    double sub[UBL ? 2 : 1][5][UBL ? 4 * BX * BY : 1];
};

// Separate flux buffer where the LDS allows it: the two-plane ring (the
// 27-point slope's three planes are the big ones) within the 160 KB of a CU.
#ifndef RH_SEP_FLUX
#define RH_SEP_FLUX 0
#endif
template <int BX, int BY, int BZ, typename T, int NS> constexpr bool sep_flux()
{
    return RH_SEP_FLUX && NS == 2 && sizeof(BrickSmemCore<BX, BY, BZ, T, NS, true, false>) <= 160 * 1024;
}

// The LDS of the per-level (WL = false) or work-list kernel.  The base
// values through LDS (UBL) pay in the FP32 kernel, whose three waves per SIMD
// leave no registers to load them early (a load at the update waits, by the
// in-order memory counter, for the next-plane prefetch issued before it): at
// 256^3 the mixed kernel 0.666 -> 0.618 ms.  The general FP64 body loads them
// early into registers (through LDS behind a runtime test it measured 4%
// slower in the AMR work-list step: the compiler then waits for the early
// loads of the other path); the specialised FP64 body (GEN = false, launched
// only where the update adds to uold) takes them through LDS.
template <int BX, int BY, int BZ, typename T, int RS, int NS, bool WL, bool GEN>
using BrickLds = BrickSmem<BX, BY, BZ, T, NS, sep_flux<BX, BY, BZ, T, NS>(), brick_waves<BX, BY, BZ, T, WL>(),
                           brick_prefetch<BX, BY, T, RS, NS, WL>() && (sizeof(T) == 4 || !GEN),
                           pair_writes<T, WL, GEN>()>;

__device__ __forceinline__ int ring_slot2(int t) { return t & 1; }
__device__ __forceinline__ int ring_slot3(int t) { return (t + 3) % 3; }

// One brick: the body shared by the per-level kernel and the work-list kernel.
// RS = 3 compiles HLLC in; NS is the depth of the primitive ring.
// A is GodArgs (a kernel argument) or GodArgs in the constant address space
// (a work-list descriptor): both are read-only for the kernel's lifetime, so
// the compiler reloads their fields with scalar loads instead of holding them.
// GEN = false compiles the body for a level that is periodic in every
// dimension, has no refined cell and no coarse-fine faces (every uniform
// level): without the boundary-ghost gather, the refined-face zeroing and the
// reflux code, the FP64 kernel needs 224 instead of 245 VGPRs and no SGPR
// spills (39 with them), and runs 2-3% faster.  The launch checks the level.
template <int BX, int BY, int BZ, typename T, int RS, int NS, bool WL, bool GEN = true, typename A>
__device__ __forceinline__ void brick_body(A &a, const int bid, BrickLds<BX, BY, BZ, T, RS, NS, WL, GEN> &sm,
                                           const int part = -1, const int pslot = -1)
{
    constexpr int CX = 2 * BX + 2, CY = 2 * BY + 2, NT = CX * CY;
    constexpr int PX = 2 * BX + 4, PY = 2 * BY + 4, NP = PX * PY;
    constexpr int WX = BX + 2, WY = BY + 2, WZ = BZ + 2, NW = WX * WY * WZ;
    int32_t *swin = sm.swin;
    auto &sq = sm.sq;
    auto &sref = sm.sref;
    auto &sf = sm.sf;
    constexpr bool SEP = sep_flux<BX, BY, BZ, T, NS>();
    T(*fb)[5][NT];
    if constexpr (SEP)
        fb = sm.sff;
    else
        fb = sm.sf;
    (void)CY;
    const int tid = threadIdx.x;
    for (int i = tid; i < NW; i += NT) swin[i] = a.win[(size_t)bid * NW + i];
    // refluxes summed in LDS (atomic mode only: the deterministic mode keeps
    // one slot per fine face)
    constexpr bool RA = std::remove_reference_t<decltype(sm)>::RA;
    const bool rlds = RA && GEN && a.reflux && !a.rdet;
    if (rlds)
        for (int i = tid; i < NW * 5; i += NT) (&sm.sacc[0][0])[i] = 0.0;
    const int ox0 = a.org[3 * bid + 0], oy0 = a.org[3 * bid + 1], oz0 = a.org[3 * bid + 2];
    // cell planes [tlo, thi] of the brick hold octs of the level; the march
    // covers them plus two halo planes on each side
    const int zr = a.zr ? a.zr[bid] : ((BZ - 1) << 16);
    int zlo = zr & 0xffff, zhi = zr >> 16;
    if (part >= 0) {
        // half of the brick's z layers (the lower half holds the middle one)
        const int mid = (zlo + zhi) >> 1;
        if (part == 0)
            zhi = mid;
        else
            zlo = mid + 1;
    }
    const int tlo = 2 * zlo, thi = 2 * zhi + 1;

    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const HydroK<T> k = make_k<T>(dt, a.dx, a.gamma, a.smallr, a.smallc2, a.courant_factor, a.smax_cap, a.use_smax,
                                  a.slope, a.riemann);
    const double smalle = a.smallc2 / a.gamma / (a.gamma - 1.0);
    const double dtxhalf = dt * 0.5;
    const double gdt[3] = {a.g[0] * dtxhalf, a.g[1] * dtxhalf, a.g[2] * dtxhalf};
    const double inv_dx = a.inv_dx;
    // The face fluxes stay unscaled (in LDS and in registers); the update
    // adds dt/dx times each difference of the two faces of a dimension with
    // one FMA, u + (dt/dx)(F(i) - F(i+1)), both cells of a face reading the
    // same flux, so the update conserves.  The reference stores (F dt)/dx per
    // face (hydro/umuscl.f90:135-220); that product, in its order for FP64,
    // is formed only for the coarse reflux.
    const double dtdx64 = dt * inv_dx;
    auto fscale = [&](double f) -> double { return sizeof(T) == 4 ? f * dtdx64 : (f * dt) * inv_dx; };
    const int boxmin[3] = {a.boxmin[0], a.boxmin[1], a.boxmin[2]};
    const int boxmax[3] = {a.boxmax[0], a.boxmax[1], a.boxmax[2]};
    const bool anyref = GEN && a.any_ref != 0;  // refined flags are all 0 when not
    int refl = 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int sd = 0; sd < 2; sd++) refl |= (a.btype[d][sd] == 1) << (2 * d + sd);
    const double gvec[3] = {a.g[0], a.g[1], a.g[2]};
    __syncthreads();

    // Column of this thread.  Row-major over the (2BX+2)(2BY+2) columns, or,
    // when the 4 BX BY interior columns fill whole waves and the halo columns
    // fit one more (HW: 8x6 bricks, 192 + 60 threads; 4x4, 64 + 36), interior columns
    // first and the halo ring last, so that the halo wave, which needs at most
    // one face per column and no update, runs one solver pass instead of three
    // and skips the update; it takes the P-phase ring instead.
    constexpr int NI = 4 * BX * BY;
    constexpr bool HW = brick_hw<BX, BY, RS, NS>();
    int cx, cy;
    if constexpr (HW) {
        if (tid < NI) {
            cx = tid % (2 * BX);
            cy = tid / (2 * BX);
        } else {
            const int h = tid - NI;
            if (h < CX) {
                cx = h - 1;
                cy = -1;
            } else if (h < 2 * CX) {
                cx = h - CX - 1;
                cy = 2 * BY;
            } else if (h < 2 * CX + 2 * BY) {
                cx = -1;
                cy = h - 2 * CX;
            } else {
                cx = 2 * BX;
                cy = h - 2 * CX - 2 * BY;
            }
        }
    } else {
        cx = tid % CX - 1;
        cy = tid / CX - 1;
    }
    // index of this column in the face-state and flux buffers (row-major)
    const int col = (cy + 1) * CX + (cx + 1);
    const int pc = (cy + 2) * PX + (cx + 2);
    const bool interior = cx >= 0 && cx < 2 * BX && cy >= 0 && cy < 2 * BY;
    const bool do_fx = cx >= 0 && cx <= 2 * BX && cy >= 0 && cy < 2 * BY;
    const bool do_fy = cy >= 0 && cy <= 2 * BY && cx >= 0 && cx < 2 * BX;

    T qmz_prev[5] = {T(0), T(0), T(0), T(0), T(0)};
    uint8_t ref_prev = 0;
    T fz_prev[5] = {T(0), T(0), T(0), T(0), T(0)};  // z-face fluxes in the solver's precision
    double part_prev[5] = {0, 0, 0, 0, 0};
    // own column primitives of planes tt-1 and tt (used when NS == 2)
    T qcol_m[5] = {T(0), T(0), T(0), T(0), T(0)}, qcol_0[5] = {T(0), T(0), T(0), T(0), T(0)};


    // P-phase cells of this thread: its own column's cell, and for the first
    // NRING threads one cell of the ring around the column region.  Their
    // loads are issued one plane ahead (u_o, u_r) and converted the next plane.
    // The ring leaves out its four corner cells, which no slope reads (the
    // 27-point stencil reaches them only from the corner columns, whose
    // states feed no face): 2(CX+CY) cells, one wave for the 7x7 brick.
    // With HW the ring is the halo wave's and leaves out the two cells next to
    // each corner as well, which only the corner columns read with the
    // one-dimensional slopes of this variant: 4(BX+BY) cells.
    constexpr int NRING = HW ? 4 * (BX + BY) : 2 * (CX + CY);
    const int r = HW ? tid - NI : tid;
    const bool ring = r >= 0 && r < NRING;
    int ring_i = 0;
    if (ring) {
        int rx, ry;
        if constexpr (HW) {
            if (r < 2 * BX) {
                rx = r;
                ry = -2;
            } else if (r < 4 * BX) {
                rx = r - 2 * BX;
                ry = 2 * BY + 1;
            } else if (r < 4 * BX + 2 * BY) {
                rx = -2;
                ry = r - 4 * BX;
            } else {
                rx = 2 * BX + 1;
                ry = r - 4 * BX - 2 * BY;
            }
        } else if (r < CX) {
            rx = r - 1;
            ry = -2;
        } else if (r < 2 * CX) {
            rx = r - CX - 1;
            ry = 2 * BY + 1;
        } else if (r < 2 * CX + CY) {
            rx = -2;
            ry = r - 2 * CX - 1;
        } else {
            rx = 2 * BX + 1;
            ry = r - 2 * CX - CY - 1;
        }
        ring_i = (ry + 2) * PX + (rx + 2);
    }
    const bool bnd = GEN && (RH_FAST_FETCH ? a.bnd != 0 : true);
    auto fetch = [&](int i, int tz, double u[5], uint8_t &rf, int &fl) {
        const int px = i % PX - 2, py = i / PX - 2;
        const int ox = (px + 2) >> 1, oy = (py + 2) >> 1, oz = (tz + 2) >> 1;
        const int v = swin[(oz * WY + oy) * WX + ox];
        const int c[3] = {px & 1, py & 1, tz & 1};
        if (!bnd) {
            // no boundary ghosts: an oct (v > 0) or none (v == 0)
            const int ind = c[0] + 2 * c[1] + 4 * c[2];
            CellRef r;
            r.p = v > 0 ? a.uold + (size_t)(v - 1) * 40 + ind : nullptr;
            r.rf = anyref && v > 0 ? a.refined + (size_t)(v - 1) * 8 + ind : nullptr;
            cell_load3(r, u, rf);
            fl = -1;
            return;
        }
        const int okey[3] = {ox0 + ox - 1, oy0 + oy - 1, oz0 + oz - 1};
        CellRef r = cell_ref3(a.uold, a.refined, v, okey, boxmin, boxmax, refl, c);
        if (!anyref) r.rf = nullptr;
        cell_load3(r, u, rf);
        fl = r.flip;
    };
    // prefetch a plane ahead (about 45 more VGPRs, load latency hidden) in the
    // HLLC variants: the FP64 kernels, and the FP32 per-level kernel, which
    // drops to two waves per SIMD with it and is faster for it; the FP32
    // work-list kernel and the other variants load at use, as they would
    // spill or lose waves (it also needs a thread for every ring cell)
    constexpr bool PF = brick_prefetch<BX, BY, T, RS, NS, WL>();
    double u_o[5], u_r[5];
    uint8_t rf_o, rf_r = 0;
    int fl_o, fl_r = -1;
    // the update's base values from the P phase (through LDS) when the update
    // adds to uold; the own column's cell of plane t is u_o in the P phase
    constexpr bool UBL = std::remove_reference_t<decltype(sm)>::UBL && PF;
    const bool ubl = UBL && (!GEN || a.ubase == a.uold);  // the specialised body is launched with ubase == uold
    const int icol = interior ? cy * (2 * BX) + cx : 0;
    if (PF) {
        fetch(pc, tlo - 2, u_o, rf_o, fl_o);
        if (ring) fetch(ring_i, tlo - 2, u_r, rf_r, fl_r);
    }

    // ---- P: primitives of plane t into ring slot st, own column to qcol_p ----
    auto p_phase = [&](const int t, const int st, T qcol_p[5]) {
        if constexpr (!PF) {
            for (int i = tid; i < NP; i += NT) {
                const int px = i % PX - 2, py = i / PX - 2;
                const int ox = (px + 2) >> 1, oy = (py + 2) >> 1, oz = (t + 2) >> 1;
                const int v = swin[(oz * WY + oy) * WX + ox];
                int okey[3] = {ox0 + ox - 1, oy0 + oy - 1, oz0 + oz - 1};
                int c[3] = {px & 1, py & 1, t & 1};
                double u[5], q[5];
                uint8_t rf;
                gather_cell_m(a.uold, a.refined, v, okey, 3, boxmin, boxmax, refl, c, u, rf);
                prim_from_cons<sizeof(T) == 4>(u, q, a.smallr, smalle, a.gamma, gdt);
#pragma unroll
                for (int n = 0; n < 5; n++) sq[st][n][i] = T(q[n]);
                sref[st][i] = rf;
            }
            __syncthreads();
#pragma unroll
            for (int n = 0; n < 5; n++) qcol_p[n] = sq[st][n][pc];
        } else {
            double q[5];
            if (bnd && fl_o >= 0) cell_flip(u_o, fl_o);  // reflexive ghosts only: skipped by waves without any
            if (ubl && interior)
#pragma unroll
                for (int n = 0; n < 5; n++) sm.sub[UBL ? (t & 1) : 0][n][UBL ? icol : 0] = u_o[n];
            prim_from_cons<sizeof(T) == 4>(u_o, q, a.smallr, smalle, a.gamma, gdt);
#pragma unroll
            for (int n = 0; n < 5; n++) {
                qcol_p[n] = T(q[n]);
                sq[st][n][pc] = qcol_p[n];
            }
            sref[st][pc] = rf_o;
            if (ring) {
                if (bnd && fl_r >= 0) cell_flip(u_r, fl_r);
                prim_from_cons<sizeof(T) == 4>(u_r, q, a.smallr, smalle, a.gamma, gdt);
#pragma unroll
                for (int n = 0; n < 5; n++) sq[st][n][ring_i] = T(q[n]);
                sref[st][ring_i] = rf_r;
            }
            // unconditional (the last plane re-reads itself): a load under a
            // branch makes the compiler copy the loop-carried registers at the
            // loop latch, which waits for the loads (and the update's stores)
            const int tn = t < thi + 2 ? t + 1 : t;
            // keep the next plane's loads below the conversions that consume
            // the current values, so that both use the same registers
            __builtin_amdgcn_sched_barrier(0);
            fetch(pc, tn, u_o, rf_o, fl_o);
            if (ring) fetch(ring_i, tn, u_r, rf_r, fl_r);
            __syncthreads();
        }
    };
    auto rotate = [&](const T qcol_p[5]) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
            qcol_m[n] = qcol_0[n];
            qcol_0[n] = qcol_p[n];
        }
    };
    // the two planes below the first output plane only fill the ring; peeled,
    // so that the march has one loop body and latch (a skipped body makes the
    // compiler copy every loop-carried register at the latch, waiting for the
    // prefetch loads and the update's stores)
    for (int t = tlo - 2; t < tlo; ++t) {
        T qcol_p[5];
        p_phase(t, NS == 3 ? ring_slot3(t) : ring_slot2(t), qcol_p);
        rotate(qcol_p);
    }
    // courant_fine of the state written (the dt of the next step), when asked
    // for: the FP64 kernel and the FP32 per-level kernel take each leaf
    // cell's new state as they store it; the FP32 work-list kernel re-reads
    // the brick's cells after the march (RH_COURANT_IN_UPDATE_F32)
    constexpr bool CU = sizeof(T) == 8 || (RH_COURANT_IN_UPDATE_F32 && !WL);
    const CourantK ck = make_courant_k(a.dx, 3, a.gamma, a.smallr, a.smallc2, a.courant_factor, gvec);
    const double gf0 = grav_factor0(ck);
    double c_mass = 0.0, c_ekin = 0.0, c_eint = 0.0, c_dt = a.dx / ck.smallc;
    // one plane of the march; par = t & 1 (the ring slot of plane t when NS == 2)
    auto march = [&](const int t, const int par) __attribute__((always_inline)) {
        const int st = NS == 3 ? ring_slot3(t) : par;
        T qcol_p[5];
        p_phase(t, st, qcol_p);
        // ---- T: slopes and predictor of plane tt = t-1 ----
        const int tt = t - 1;
        const bool plane_out = tt >= tlo && tt <= thi;
        // the values the update of plane tt adds to: with the prefetch (same
        // register budget), loaded here so that the T and F phases cover
        // their latency, else (and in the three-wave FP32 kernel) at the update
        constexpr bool UB_EARLY = PF && !f32_three_waves<BX, BY, BZ, T, WL>();
        double ub[5];
        if constexpr (UB_EARLY) {
            if (!ubl && interior && plane_out) {
                const int ox = (cx >> 1) + 1, oy = (cy >> 1) + 1, oz = (tt >> 1) + 1;
                const int v = swin[(oz * WY + oy) * WX + ox];
                const int ind = (cx & 1) + 2 * (cy & 1) + 4 * (tt & 1);
                const double *pb = a.ubase + (size_t)(v > 0 ? v - 1 : 0) * 40 + ind;
#pragma unroll
                for (int n = 0; n < 5; n++) ub[n] = pb[8 * n];
            }
        }
        const int s0 = NS == 3 ? ring_slot3(t - 1) : par ^ 1;
        const int smm = NS == 3 ? ring_slot3(t - 2) : 0;
        T q0[5], dq[3][5];
#pragma unroll
        for (int n = 0; n < 5; n++) q0[n] = qcol_0[n];
        if (NS == 3 && k.slope_type == 3) {
#pragma unroll
            for (int n = 0; n < 5; n++) {
                T d3[3];
                slope27<T, PX, NP>(sq, smm, s0, st, n, pc, d3);
                dq[0][n] = d3[0];
                dq[1][n] = d3[1];
                dq[2][n] = d3[2];
            }
        } else {
            // one uniform branch on the slope type for all 15 slopes
            auto slopes = [&](const int sl) {
#pragma unroll
                for (int n = 0; n < 5; n++) {
                    dq[0][n] = slope_half(sl, sq[s0][n][pc - 1], q0[n], sq[s0][n][pc + 1]);
                    dq[1][n] = slope_half(sl, sq[s0][n][pc - PX], q0[n], sq[s0][n][pc + PX]);
                    dq[2][n] = slope_half(sl, qcol_m[n], q0[n], qcol_p[n]);
                }
            };
            if (k.slope_type == 1)
                slopes(1);
            else if (k.slope_type != 0)
                slopes(2);
            else
#pragma unroll
                for (int n = 0; n < 5; n++) dq[0][n] = dq[1][n] = dq[2][n] = T(0);
        }
        const T inv_r0 = trcp(q0[0]);
        if (k.slope_type > 0) {
            dq[0][1] = vel_bound_half(dq[0][1], sq[s0][0][pc - 1], sq[s0][1][pc - 1], inv_r0, q0[1],
                                      sq[s0][0][pc + 1], sq[s0][1][pc + 1]);
            dq[1][2] = vel_bound_half(dq[1][2], sq[s0][0][pc - PX], sq[s0][2][pc - PX], inv_r0, q0[2],
                                      sq[s0][0][pc + PX], sq[s0][2][pc + PX]);
            dq[2][3] = vel_bound_half(dq[2][3], qcol_m[0], qcol_m[3], inv_r0, q0[3], qcol_p[0], qcol_p[3]);
        }
        T qp[3][5], qm[3][5];
        trace3(q0, dq, qp, qm, k, inv_r0);
        const uint8_t myref = anyref ? sref[s0][pc] : 0;
#pragma unroll
        for (int n = 0; n < 5; n++) {
            sf[0][n][col] = qm[0][n];
            sf[1][n][col] = qm[1][n];
        }
        __syncthreads();
        // ---- F: Riemann problems on the low x, y, z faces of this column ----
        T fx[5], fy[5];
        T fz[5] = {T(0), T(0), T(0), T(0), T(0)};
        const bool wfx = do_fx && plane_out, wfy = do_fy && plane_out;
        if (HW && tid >= NI) {
            // halo wave: the right column's x face or the top row's y face,
            // one solver pass with the direction chosen per lane
            if (wfx || wfy) {
                const bool dy = wfy;
                const int nb = dy ? col - CX : col - 1;
                T ql[5], qr[5], a5[5], b5[5], g[5];
#pragma unroll
                for (int n = 0; n < 5; n++) {
                    ql[n] = dy ? sf[1][n][nb] : sf[0][n][nb];
                    qr[n] = dy ? qp[1][n] : qp[0][n];
                }
                a5[0] = ql[0], a5[1] = dy ? ql[2] : ql[1], a5[2] = dy ? ql[1] : ql[2], a5[3] = ql[3], a5[4] = ql[4];
                b5[0] = qr[0], b5[1] = dy ? qr[2] : qr[1], b5[2] = dy ? qr[1] : qr[2], b5[3] = qr[3], b5[4] = qr[4];
                riemann_hllc(a5, b5, g, k);
                T f[5] = {g[0], dy ? g[2] : g[1], dy ? g[1] : g[2], g[3], g[4]};
                if (anyref) {
                    const bool z = myref || sref[s0][dy ? pc - PX : pc - 1];
#pragma unroll
                    for (int n = 0; n < 5; n++) f[n] = z ? T(0) : f[n];
                }
#pragma unroll
                for (int n = 0; n < 5; n++) {
                    fx[n] = f[n];
                    fy[n] = f[n];
                }
            }
        } else {
            if (wfx) {
                T ql[5];
#pragma unroll
                for (int n = 0; n < 5; n++) ql[n] = sf[0][n][col - 1];
                face_flux<RS>(0, ql, qp[0], fx, k);
                if (anyref) {
                    const bool z = myref || sref[s0][pc - 1];
#pragma unroll
                    for (int n = 0; n < 5; n++) fx[n] = z ? T(0) : fx[n];
                }
            }
            if (wfy) {
                T ql[5];
#pragma unroll
                for (int n = 0; n < 5; n++) ql[n] = sf[1][n][col - CX];
                face_flux<RS>(1, ql, qp[1], fy, k);
                if (anyref) {
                    const bool z = myref || sref[s0][pc - PX];
#pragma unroll
                    for (int n = 0; n < 5; n++) fy[n] = z ? T(0) : fy[n];
                }
            }
        }
        if (interior && tt >= tlo && tt <= thi + 1) {
            T f[5];
            face_flux<RS>(2, qmz_prev, qp[2], f, k);
#pragma unroll
            for (int n = 0; n < 5; n++) fz[n] = f[n];
            if (anyref) {
                const bool z = myref || ref_prev;
#pragma unroll
                for (int n = 0; n < 5; n++) fz[n] = z ? T(0) : fz[n];
            }
        }
        if (GEN && a.reflux) {
            // window slot of a cell of the trace region
            auto wi = [&](int x, int y, int z) { return (((z + 2) >> 1) * WY + ((y + 2) >> 1)) * WX + ((x + 2) >> 1); };
            // one fine face's reflux into the ghost oct of window slot i:
            // summed in LDS, or straight to the accumulator / parent cell
            auto radd = [&](int i, int v, double sign, const double F[5], int slot) {
                if (RA && rlds) {
#pragma unroll
                    for (int n = 0; n < 5; n++) atomicAdd(&sm.sacc[RA ? i : 0][n], sign * (F[n] * 0.125));
                } else {
                    reflux_add(a.unew, a.ghost_parent, a.ngridmax, v, sign, F, 8, a.racc, a.rdet, slot);
                }
            };
            auto owned = [&](int x, int y, int z, int v) {
                return x >= 0 && x < 2 * BX && y >= 0 && y < 2 * BY && z >= 0 && z < 2 * BZ && v >= a.head &&
                       v < a.head + a.noct && ox0 + (x >> 1) < boxmax[0] && oy0 + (y >> 1) < boxmax[1] &&
                       oz0 + (z >> 1) < boxmax[2];
            };
            double F[5];
            if (wfx && !(myref || sref[s0][pc - 1])) {
                const int il = wi(cx - 1, cy, tt), ih = wi(cx, cy, tt), vl = swin[il], vh = swin[ih];
#pragma unroll
                for (int n = 0; n < 5; n++) F[n] = fscale(double(fx[n]));
                const int sub = (cy & 1) + 2 * (tt & 1);
                if (vl > a.ngridmax && owned(cx, cy, tt, vh)) radd(il, vl, -1.0, F, 4 + sub);
                if (vh > a.ngridmax && owned(cx - 1, cy, tt, vl)) radd(ih, vh, 1.0, F, sub);
            }
            if (wfy && !(myref || sref[s0][pc - PX])) {
                const int il = wi(cx, cy - 1, tt), ih = wi(cx, cy, tt), vl = swin[il], vh = swin[ih];
#pragma unroll
                for (int n = 0; n < 5; n++) F[n] = fscale(double(fy[n]));
                const int sub = (cx & 1) + 2 * (tt & 1);
                if (vl > a.ngridmax && owned(cx, cy, tt, vh)) radd(il, vl, -1.0, F, 12 + sub);
                if (vh > a.ngridmax && owned(cx, cy - 1, tt, vl)) radd(ih, vh, 1.0, F, 8 + sub);
            }
            if (interior && tt >= tlo && tt <= thi + 1 && !(myref || ref_prev)) {
                const int il = wi(cx, cy, tt - 1), ih = wi(cx, cy, tt), vl = swin[il], vh = swin[ih];
#pragma unroll
                for (int n = 0; n < 5; n++) F[n] = fscale(double(fz[n]));
                const int sub = (cx & 1) + 2 * (cy & 1);
                if (vl > a.ngridmax && owned(cx, cy, tt, vh)) radd(il, vl, -1.0, F, 20 + sub);
                if (vh > a.ngridmax && owned(cx, cy, tt - 1, vl)) radd(ih, vh, 1.0, F, 16 + sub);
            }
        }
        if constexpr (!SEP) __syncthreads();  // the fluxes overwrite the face states in sf
        if (wfx)
#pragma unroll
            for (int n = 0; n < 5; n++) fb[0][n][col] = fx[n];
        if (wfy)
#pragma unroll
            for (int n = 0; n < 5; n++) fb[1][n][col] = fy[n];
        __syncthreads();
        // ---- U: conservative update ----
        if (interior) {
            double part[5] = {0, 0, 0, 0, 0};
            if (plane_out) {
                if (ubl) {
#pragma unroll
                    for (int n = 0; n < 5; n++) ub[n] = sm.sub[UBL ? (tt & 1) : 0][n][UBL ? icol : 0];
                } else if constexpr (!UB_EARLY) {
                    const int ox = (cx >> 1) + 1, oy = (cy >> 1) + 1, oz = (tt >> 1) + 1;
                    const int v = swin[(oz * WY + oy) * WX + ox];
                    const int ind = (cx & 1) + 2 * (cy & 1) + 4 * (tt & 1);
                    const double *pb = a.ubase + (size_t)(v > 0 ? v - 1 : 0) * 40 + ind;
#pragma unroll
                    for (int n = 0; n < 5; n++) ub[n] = pb[8 * n];
                }
#pragma unroll
                for (int n = 0; n < 5; n++) {
                    const double fl = fb[0][n][col], fr = fb[0][n][col + 1];
                    const double gl = fb[1][n][col], gr = fb[1][n][col + CX];
                    part[n] = __builtin_fma(dtdx64, gl - gr, __builtin_fma(dtdx64, fl - fr, ub[n]));
                }
            }
            const int tf = tt - 1;
            if (tf >= tlo && tf <= thi) {
                const int ox = (cx >> 1) + 1, oy = (cy >> 1) + 1, oz = (tf >> 1) + 1;
                const int v = swin[(oz * WY + oy) * WX + ox];
                const bool own = v >= a.head && v < a.head + a.noct && ox0 + ox - 1 < boxmax[0] &&
                                 oy0 + oy - 1 < boxmax[1] && oz0 + oz - 1 < boxmax[2];
                if (own) {
                    const int ind = (cx & 1) + 2 * (cy & 1) + 4 * (tf & 1);
                    double *po = a.unew + (size_t)(v - 1) * 40 + ind;
                    double un[5];
#pragma unroll
                    for (int n = 0; n < 5; n++)
                        un[n] = __builtin_fma(dtdx64, double(fz_prev[n]) - double(fz[n]), part_prev[n]);
                    if constexpr (std::remove_reference_t<decltype(sm)>::PW) {
                        // the two z-planes of an oct fill each 64-byte line of
                        // unew: the lower plane's values wait in LDS (the
                        // thread's own slot) and both halves are stored
                        // together with the upper plane
                        const int ic = cy * (2 * BX) + cx;
                        if (tf & 1) {
#pragma unroll
                            for (int n = 0; n < 5; n++) {
                                po[8 * n - 4] = sm.sun[n][ic];
                                po[8 * n] = un[n];
                            }
                        } else {
#pragma unroll
                            for (int n = 0; n < 5; n++) sm.sun[n][ic] = un[n];
                        }
                    } else {
#pragma unroll
                        for (int n = 0; n < 5; n++) po[8 * n] = un[n];
                    }
                    if (CU && a.courant && !ref_prev) {
                        if constexpr (sizeof(T) == 4)
                            courant_cell_mixed(un, ck, c_mass, c_ekin, c_eint, c_dt);
                        else
                            courant_cell(un, ck, gf0, c_mass, c_ekin, c_eint, c_dt);
                    }
                }
            }
#pragma unroll
            for (int n = 0; n < 5; n++) {
                part_prev[n] = part[n];
                fz_prev[n] = fz[n];
            }
        }
#pragma unroll
        for (int n = 0; n < 5; n++) qmz_prev[n] = qm[2][n];
        ref_prev = myref;
        rotate(qcol_p);
    };
    if constexpr (RH_MARCH_UNROLL2 && NS == 2) {
        // two planes per trip (the trip count 2(zhi - zlo + 2) is even): the
        // ring slots are constants and the registers carried from plane to
        // plane are renamed instead of copied
        for (int t = tlo; t <= thi + 2; t += 2) {
            march(t, 0);
            march(t + 1, 1);
        }
    } else {
        for (int t = tlo; t <= thi + 2; ++t) march(t, t & 1);
    }
    if (RA && rlds) {
        // one atomic per ghost oct of the window and conserved variable
        __syncthreads();
        for (int i = tid; i < NW; i += NT) {
            const int v = swin[i];
            if (v <= a.ngridmax) continue;
            double q[5];
            bool any = false;
#pragma unroll
            for (int n = 0; n < 5; n++) {
                q[n] = sm.sacc[RA ? i : 0][n];
                any = any || q[n] != 0.0;
            }
            if (any) reflux_flush(a.unew, a.ghost_parent, a.ngridmax, v, q, 8, a.racc);
        }
    }
    if (a.courant) {
        // the brick's partials of courant_fine in a fixed reduction order
        // (FP64); the FP32 kernel first re-reads its own leaf cells from unew
        // while still in cache
        __syncthreads();
        for (int i = tid; !CU && i < BX * BY * BZ * 8; i += NT) {
            const int oc = i >> 3, ind = i & 7;
            const int ox = oc % BX + 1, oy = (oc / BX) % BY + 1, oz = oc / (BX * BY) + 1;
            const int v = swin[(oz * WY + oy) * WX + ox];
            const bool own = v >= a.head && v < a.head + a.noct && ox0 + ox - 1 < boxmax[0] &&
                             oy0 + oy - 1 < boxmax[1] && oz0 + oz - 1 < boxmax[2];
            if (!own || a.refined[(size_t)(v - 1) * 8 + ind]) continue;
            const double *pu = a.unew + (size_t)(v - 1) * 40 + ind;
            const double un[5] = {pu[0], pu[8], pu[16], pu[24], pu[32]};
            courant_cell_mixed(un, ck, c_mass, c_ekin, c_eint, c_dt);
        }
        // lanes of this wave that exist (NT need not be a multiple of 64)
        const int nv = NT - (tid & ~63) < 64 ? NT - (tid & ~63) : 64;
        c_mass = wave_sum_n(c_mass, nv);
        c_ekin = wave_sum_n(c_ekin, nv);
        c_eint = wave_sum_n(c_eint, nv);
        c_dt = wave_min_n(c_dt, nv);
        __syncthreads();
        double *red = (double *)&sf[0][0][0];
        const int w = tid >> 6;
        if ((tid & 63) == 0) {
            red[4 * w + 0] = c_mass;
            red[4 * w + 1] = c_ekin;
            red[4 * w + 2] = c_eint;
            red[4 * w + 3] = c_dt;
        }
        __syncthreads();
        if (tid == 0) {
            double m = 0, e = 0, ei = 0, d = red[3];
            for (int i = 0; i < (NT + 63) / 64; i++) {
                m += red[4 * i];
                e += red[4 * i + 1];
                ei += red[4 * i + 2];
                d = red[4 * i + 3] < d ? red[4 * i + 3] : d;
            }
            const int ps = pslot >= 0 ? pslot : bid;
            a.cpart[4 * ps + 0] = m;
            a.cpart[4 * ps + 1] = e;
            a.cpart[4 * ps + 2] = ei;
            a.cpart[4 * ps + 3] = d;
        }
    }
}

template <int BX, int BY, int BZ, typename T, int RS, int NS, bool GEN = true>
__global__ void __launch_bounds__((2 * BX + 2) * (2 * BY + 2),
                                  (sizeof(T) == 8                              ? RH_BRICK_MIN_WAVES
                                   : f32_three_waves<BX, BY, BZ, T, false>() ? 3
                                                                             : RH_BRICK_MIN_WAVES_F32))
    k_godunov3d(const GodArgs a)
{
    __shared__ BrickLds<BX, BY, BZ, T, RS, NS, false, GEN> sm;
    // XCD-aware brick assignment: consecutive bricks to the same XCD so that
    // their shared halo octs meet in that XCD's L2.
    int bid = blockIdx.x;
    int part = -1, pslot = bid;
    if (a.xcd_remap && a.nsplit == 0) {
        const int nb = gridDim.x, x = bid & 7, l = bid >> 3;
        const int q = nb >> 3, r = nb & 7;
        bid = x * q + (x < r ? x : r) + l;
        pslot = bid;
    } else if (a.xcd_remap) {
        // the last nsplit bricks of each XCD's range run as two work items,
        // one per half of the brick's z layers, so that the final round of
        // workgroups on the XCD is half as long
        const int x = blockIdx.x & 7, l = blockIdx.x >> 3;
        const int q = a.nbricks >> 3, r = a.nbricks & 7, m = a.nsplit;
        const int start = x * q + (x < r ? x : r), count = q + (x < r ? 1 : 0);
        if (l >= count + m) return;
        if (l < count - m) {
            bid = start + l;
            pslot = bid;
        } else {
            const int j = l - (count - m);
            bid = start + count - m + (j >> 1);
            part = j & 1;
            // Courant partial of the upper half: after the nbricks whole-brick slots
            pslot = part == 0 ? bid : a.nbricks + x * m + (j >> 1);
        }
    }
    if (bid >= a.nbricks) return;
    brick_body<BX, BY, BZ, T, RS, NS, false, GEN>(a, bid, sm, part, pslot);
}

// Work-list kernel: work item i = (descriptor, brick), where a descriptor
// carries the arrays and parameters of one level of one mesh, so bricks of
// several levels and of several jobs share one launch.  The items are split
// into 8 contiguous ranges, one per XCD: workgroups are placed on the XCDs
// round-robin by workgroup id, so workgroup b runs item range[b % 8] + b / 8
// on XCD b % 8, and neighbouring bricks (which share halo octs) meet in that
// XCD's L2.  One workgroup per item: the hardware dispatcher is the queue.  (A
// persistent loop over an atomic counter holds its loop state and the
// reloaded descriptor in registers and measured 160 instead of 125 VGPRs for
// the FP32 HLLC kernel, one wave per SIMD less.)
typedef const __attribute__((address_space(4))) GodArgs ConstGodArgs;

template <int BX, int BY, int BZ, typename T, int RS, int NS, bool GEN = true>
__global__ void __launch_bounds__((2 * BX + 2) * (2 * BY + 2), (brick_waves<BX, BY, BZ, T, true>()))
    k_godunov3d_wl(const WorkList w)
{
    __shared__ BrickLds<BX, BY, BZ, T, RS, NS, true, GEN> sm;
    const int x = blockIdx.x & 7, l = blockIdx.x >> 3;
    const int i = w.range[x] + l;
    if (i >= w.range[x + 1]) return;
    const int2 it = w.items[i];
    const int di = __builtin_amdgcn_readfirstlane(it.x), bi = __builtin_amdgcn_readfirstlane(it.y);
    brick_body<BX, BY, BZ, T, RS, NS, true, GEN>(((ConstGodArgs *)w.descs)[di], bi, sm);
}


// ---------------------------------------------------------------------------
// 1D: one thread per cell; stencil of 5 cells from the 3-slot neighbour table.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_godunov1d(const God1Args a)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2 * a.noct) return;
    const int o = a.head + (gid >> 1);  // 1-based oct
    const int i2 = gid & 1;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const HydroK<T> k = make_k<T>(dt, a.dx, a.gamma, a.smallr, a.smallc2, a.courant_factor, a.smax_cap, a.use_smax,
                                  a.slope < 2 ? a.slope : 2, a.riemann);
    const double smalle = a.smallc2 / a.gamma / (a.gamma - 1.0);
    const double gdt[3] = {a.g * (dt * 0.5), 0.0, 0.0};
    const int32_t *nb = a.nbor + (size_t)(o - a.head) * 3;
    const int boxmin[3] = {a.boxmin, 0, 0}, boxmax[3] = {a.boxmax, 1, 1};
    const int btype[3][2] = {{a.btype[0], a.btype[1]}, {2, 2}, {2, 2}};
    T q[5][5];
    uint8_t rf[5];
    for (int d = -2; d <= 2; d++) {
        const int p = i2 + d;
        const int slot = (p + 2) >> 1;
        const int v = nb[slot];
        // key of that oct relative to the domain: self oct key is unknown here,
        // the slot tells the side (slot 0 = low, 2 = high)
        int okey[3] = {slot == 0 ? a.boxmin - 1 : (slot == 2 ? a.boxmax : a.boxmin), 0, 0};
        int c[3] = {p & 1, 0, 0};
        double u[5], qq[5];
        gather_cell(a.uold, a.refined, v, okey, 1, boxmin, boxmax, btype, c, u, rf[d + 2]);
        prim_from_cons(u, qq, a.smallr, smalle, a.gamma, gdt);
#pragma unroll
        for (int n = 0; n < 5; n++) q[d + 2][n] = T(qq[n]);
    }
    T qp[3][5], qm[3][5];
#pragma unroll
    for (int c = 1; c <= 3; c++) {
        T dq[5];
#pragma unroll
        for (int n = 0; n < 5; n++) dq[n] = k.slope_type ? slope_half(k.slope_type, q[c - 1][n], q[c][n], q[c + 1][n]) : T(0);
        const T inv_r = trcp(q[c][0]);
        if (k.slope_type > 0)
            dq[1] = vel_bound_half(dq[1], q[c - 1][0], q[c - 1][1], inv_r, q[c][1], q[c + 1][0], q[c + 1][1]);
        trace1(q[c], dq, qp[c - 1], qm[c - 1], k, inv_r);
    }
    T fl[5], fr[5];
    face_flux(0, qm[0], qp[1], fl, k);
    face_flux(0, qm[1], qp[2], fr, k);
    const bool zl = rf[1] || rf[2], zr = rf[2] || rf[3];
    const int ind = i2;
    const double *pb = a.ubase + (size_t)(o - 1) * 10 + ind;
    double *po = a.unew + (size_t)(o - 1) * 10 + ind;
    double FLs[5], FRs[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
        FLs[n] = zl ? 0.0 : (double(fl[n]) * dt) * a.inv_dx;
        FRs[n] = zr ? 0.0 : (double(fr[n]) * dt) * a.inv_dx;
        po[2 * n] = pb[2 * n] + (FLs[n] - FRs[n]);
    }
    if (a.reflux) {
        // the left neighbour cell lies in this oct (i2=1) or in slot 0; the right one in this oct or slot 2
        const int vl = i2 == 1 ? o : nb[0], vr = i2 == 0 ? o : nb[2];
        if (!zl && vl > a.ngridmax) reflux_add(a.unew, a.ghost_parent, a.ngridmax, vl, -1.0, FLs, 2, a.racc, a.rdet, 4);
        if (!zr && vr > a.ngridmax) reflux_add(a.unew, a.ghost_parent, a.ngridmax, vr, 1.0, FRs, 2, a.racc, a.rdet, 0);
    }
}

// ---------------------------------------------------------------------------
// Courant time step and conservation sums (courant_fine + cmpdt).
// ---------------------------------------------------------------------------
constexpr int kCourantThreads = 256;

template <int TTD>
__global__ void __launch_bounds__(kCourantThreads) k_courant(const double *__restrict__ uold,
                                                             const uint8_t *__restrict__ refined, int head, int noct,
                                                             const CourantK c, double *__restrict__ part)
{
    double mass = 0, ekin = 0, eint = 0, dtmin = c.dx / c.smallc;
    const double gf0 = grav_factor0(c);
    constexpr int ttd = TTD, stride = 5 * TTD;
    const long ncell = (long)noct * ttd;
    const long step = (long)gridDim.x * blockDim.x;
    // four cells per thread per round, their loads issued together
    constexpr int U = 4;
    for (long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x; i0 < ncell; i0 += U * step) {
        double u[U][5];
        bool leaf[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const long i = i0 + k * step;
            leaf[k] = false;
            if (i < ncell) {
                const long o = (long)(head - 1) + i / ttd;
                const int ind = (int)(i % ttd);
                const double *p = uold + o * stride + ind;
#pragma unroll
                for (int n = 0; n < 5; n++) u[k][n] = p[n * ttd];
                leaf[k] = !refined[o * 8 + ind];
            }
        }
#pragma unroll
        for (int k = 0; k < U; k++)
            if (leaf[k]) courant_cell(u[k], c, gf0, mass, ekin, eint, dtmin);
    }
    __shared__ double red[kCourantThreads / 64][4];
    mass = wave_sum(mass);
    ekin = wave_sum(ekin);
    eint = wave_sum(eint);
    dtmin = wave_min(dtmin);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        red[w][0] = mass;
        red[w][1] = ekin;
        red[w][2] = eint;
        red[w][3] = dtmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0, m = red[0][3];
        for (int i = 0; i < kCourantThreads / 64; i++) {
            s0 += red[i][0];
            s1 += red[i][1];
            s2 += red[i][2];
            m = red[i][3] < m ? red[i][3] : m;
        }
        part[4 * blockIdx.x + 0] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = m;
    }
}

// courant_fine over octs of several levels: c0 carries the level-independent
// constants (c0.dx = finest dx, c0.gsum_dx = sum |g|), dx of an oct is
// boxlen*2^-lev (the same value the host computes as boxlen/2^lev).  MIXED:
// the cell's signal speed and dt in FP32, as the mixed-precision kernels'
// fused reduction (courant_cell_mixed).
template <bool MIXED>
__global__ void __launch_bounds__(kCourantThreads) k_courant_ml(const double *__restrict__ uold,
                                                                const uint8_t *__restrict__ refined,
                                                                const int32_t *__restrict__ lev, int head, int noct,
                                                                const CourantK c0, double boxlen,
                                                                double *__restrict__ part)
{
    double mass = 0, ekin = 0, eint = 0, dtmin = c0.dx / c0.smallc;
    const double gf0 = grav_factor0(c0);
    const long ncell = (long)noct * 8;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += (long)gridDim.x * blockDim.x) {
        const long o = (long)(head - 1) + i / 8;
        const int ind = (int)(i % 8);
        if (refined[o * 8 + ind]) continue;
        CourantK c = c0;
        const double dx = __builtin_ldexp(boxlen, -lev[o]);
        c.dx = dx;
        c.vol = dx * dx * dx;
        c.gsum_dx = c0.gsum_dx * dx;
        const double *p = uold + o * 40 + ind;
        const double u[5] = {p[0], p[8], p[16], p[24], p[32]};
        if constexpr (MIXED)
            courant_cell_mixed(u, c, mass, ekin, eint, dtmin);
        else
            courant_cell(u, c, gf0, mass, ekin, eint, dtmin);
    }
    __shared__ double red[kCourantThreads / 64][4];
    mass = wave_sum(mass);
    ekin = wave_sum(ekin);
    eint = wave_sum(eint);
    dtmin = wave_min(dtmin);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        red[w][0] = mass;
        red[w][1] = ekin;
        red[w][2] = eint;
        red[w][3] = dtmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0, m = red[0][3];
        for (int i = 0; i < kCourantThreads / 64; i++) {
            s0 += red[i][0];
            s1 += red[i][1];
            s2 += red[i][2];
            m = red[i][3] < m ? red[i][3] : m;
        }
        part[4 * blockIdx.x + 0] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = m;
    }
}

// Block b of a grid of nb blocks renumbered so that the blocks the hardware
// sends to one XCD (b mod 8) take one contiguous range of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int x = b & 7, l = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + l;
}

__global__ void k_zero(double *p, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

// Coarse-fine reflux fold: each ghost oct stands for one parent cell.  The
// accumulator is left zeroed for the next step (it starts zeroed); ghosts
// that received no flux leave their parent cell untouched, as the direct
// reflux of the per-level calls does.
// With nslot = kRefluxSlots (deterministic mode) each ghost has one slot per
// face of its oct, added to the parent cell one after another in slot order:
// the same sums on every run.
__global__ void k_reflux_fold(double *unew, double *__restrict__ racc, const int32_t *__restrict__ ghost_parent,
                              const int32_t *__restrict__ order, int nghost, int ttd, int nslot)
{
    // in parent-cell order (see k_refresh_ghosts)
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= nghost) return;
    const int g = order[t];
    double *rg = racc + 5 * (size_t)g * nslot;
    const int igrid = ghost_parent[2 * g], icell = ghost_parent[2 * g + 1];
    double *p = igrid > 0 ? unew + (size_t)(igrid - 1) * 5 * ttd + (icell - 1) : nullptr;
    double acc[5];
    bool touched = false;
    for (int k = 0; k < nslot; k++) {
        double f[5];
        bool any = false;
#pragma unroll
        for (int n = 0; n < 5; n++) {
            f[n] = rg[5 * k + n];
            any = any || f[n] != 0.0;
        }
        if (!any) continue;
#pragma unroll
        for (int n = 0; n < 5; n++) rg[5 * k + n] = 0.0;
        if (!p) continue;  // parent in a boundary oct: the reference drops it with the cache oct
        if (!touched) {
#pragma unroll
            for (int n = 0; n < 5; n++) acc[n] = p[n * ttd];
            touched = true;
        }
#pragma unroll
        for (int n = 0; n < 5; n++) acc[n] = acc[n] + f[n];
    }
    if (touched)
#pragma unroll
        for (int n = 0; n < 5; n++) p[n * ttd] = acc[n];
}

// Fixed-order reduction of the per-block partials by one 256-thread block:
// out = {mass, ekin, eint, emag, dt}; dt_out (optional) receives dt for the
// next Godunov launch; t_acc (optional) accumulates the time (the old dt_out
// when acc_old, else the new dt).
__device__ __forceinline__ void courant_final_block(const double *__restrict__ part, int nblocks, double dtmax,
                                                    double *__restrict__ out, double *__restrict__ dt_out,
                                                    double *__restrict__ t_acc, int acc_old)
{
    __shared__ double red[4][4];
    double s0 = 0, s1 = 0, s2 = 0, m = dtmax;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
        s0 += part[4 * i + 0];
        s1 += part[4 * i + 1];
        s2 += part[4 * i + 2];
        m = part[4 * i + 3] < m ? part[4 * i + 3] : m;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    m = wave_min(m);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        red[w][0] = s0;
        red[w][1] = s1;
        red[w][2] = s2;
        red[w][3] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = 0, a1 = 0, a2 = 0, mm = red[0][3];
        for (int i = 0; i < 4; i++) {
            a0 += red[i][0];
            a1 += red[i][1];
            a2 += red[i][2];
            mm = red[i][3] < mm ? red[i][3] : mm;
        }
        out[0] = a0;
        out[1] = a1;
        out[2] = a2;
        out[3] = 0.0;
        out[4] = mm;
        if (t_acc) *t_acc += (acc_old && dt_out) ? *dt_out : mm;
        if (dt_out) *dt_out = mm;
    }
}

__global__ void __launch_bounds__(256) k_courant_final(const double *__restrict__ part, int nblocks, double dtmax,
                                                       double *__restrict__ out, double *__restrict__ dt_out,
                                                       double *__restrict__ t_acc, int acc_old)
{
    courant_final_block(part, nblocks, dtmax, out, dt_out, t_acc, acc_old);
}

// One block per job of a batch: each job's own partials, in its fixed order.
__global__ void __launch_bounds__(256) k_courant_final_batch(const CourantJob *__restrict__ jobs)
{
    const CourantJob j = jobs[blockIdx.x];
    courant_final_block(j.part, j.nparts, j.dtmax, j.out, j.dt_inout, j.t_acc, 1);
}

// ---------------------------------------------------------------------------
// AMR: ghost octs (interpolated from the parent level) and restriction
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cell_xc(int ndim, int ind, double xc[3])
{
    const int iz = (ind - 1) / 4, iy = (ind - 1 - 4 * iz) / 2, ix = ind - 1 - 2 * iy - 4 * iz;
    xc[0] = (double)ix - 0.5;
    xc[1] = ndim > 1 ? (double)iy - 0.5 : 0.0;
    xc[2] = ndim > 2 ? (double)iz - 0.5 : 0.0;
}

// interpol_hydro (hydro/interpol_hydro.f90:5-113) with the limiters of
// :610-752; u1[j][ivar], j = 0 (central parent) and 1..2ndim (-x,+x,-y,+y,-z,+z).
// ND is a compile-time constant so that every loop unrolls and u1/u2 stay in
// registers (a runtime ndim put them in scratch memory).
template <int ND>
__device__ __forceinline__ void interpol_hydro_dev(double u1[7][5], double u2[8][5], int interpol_var, int interpol_type,
                                   double smallr)
{
    constexpr int ndim = ND, twondim = 2 * ND, ttd = 1 << ND;
    if (interpol_var == 1)
        #pragma unroll
        for (int j = 0; j <= twondim; j++) {
            double ekin = 0.0;
            #pragma unroll
            for (int d = 1; d <= 3; d++) ekin = ekin + 0.5 * (u1[j][d] * u1[j][d]) / fmax(u1[j][0], smallr);
            u1[j][4] = u1[j][4] - ekin - 0.0;
        }
#pragma unroll
    for (int iv = 0; iv < 5; iv++) {
        double a[7], w[3] = {0, 0, 0};
        #pragma unroll
        for (int j = 0; j <= twondim; j++) a[j] = u1[j][iv];
        if (interpol_type == 1) {
            #pragma unroll
            for (int d = 1; d <= ndim; d++) {
                const double dl = 0.5 * (a[2 * d] - a[0]), dr = 0.5 * (a[0] - a[2 * d - 1]);
                w[d - 1] = (dl * dr <= 0.0) ? 0.0 : fmin(fabs(dl), fabs(dr)) * dl / fabs(dl);
            }
        } else if (interpol_type == 2 || interpol_type == 3) {
            #pragma unroll
            for (int d = 1; d <= ndim; d++) w[d - 1] = 0.25 * (a[2 * d] - a[2 * d - 1]);
            if (interpol_type == 2) {
                double ac[8];
                #pragma unroll
                for (int c = 0; c < ttd; c++) ac[c] = a[0];
                #pragma unroll
                for (int d = 1; d <= ndim; d++)
                    #pragma unroll
                    for (int c = 1; c <= ttd; c++) {
                        double xc[3];
                        cell_xc(ndim, c, xc);
                        ac[c - 1] = ac[c - 1] + 2.0 * w[d - 1] * xc[d - 1];
                    }
                double cmax = ac[0], cmin = ac[0], kmax = a[1], kmin = a[1];
                #pragma unroll
                for (int c = 1; c < ttd; c++) {
                    cmax = fmax(cmax, ac[c]);
                    cmin = fmin(cmin, ac[c]);
                }
                #pragma unroll
                for (int j = 2; j <= twondim; j++) {
                    kmax = fmax(kmax, a[j]);
                    kmin = fmin(kmin, a[j]);
                }
                double dk = a[0] - kmax, dc = a[0] - cmax;
                const double lmax = (dk * dc > 0.0) ? fmin(1.0, dk / dc) : 0.0;
                dk = a[0] - kmin;
                dc = a[0] - cmin;
                const double lmin = (dk * dc > 0.0) ? fmin(1.0, dk / dc) : 0.0;
                const double lim = fmin(lmin, lmax);
                #pragma unroll
                for (int d = 0; d < ndim; d++) w[d] = w[d] * lim;
            }
        }
        #pragma unroll
        for (int c = 1; c <= ttd; c++) {
            double xc[3];
            cell_xc(ndim, c, xc);
            double v = a[0];
            #pragma unroll
            for (int d = 1; d <= ndim; d++) v = v + w[d - 1] * xc[d - 1];
            u2[c - 1][iv] = v;
        }
    }
    if (interpol_var == 1)
        #pragma unroll
        for (int c = 0; c < ttd; c++) {
            double ekin = 0.0;
            #pragma unroll
            for (int d = 1; d <= 3; d++) ekin = ekin + 0.5 * (u2[c][d] * u2[c][d]) / fmax(u2[c][0], smallr);
            u2[c][4] = u2[c][4] + ekin + 0.0;
        }
}

template <int ND>
__global__ void __launch_bounds__(64) k_refresh_ghosts(const GhostArgs a)
{
    // ghosts in parent-cell order, a contiguous range of that order per XCD,
    // so that the parent octs a wave gathers meet in one L2
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = t < a.nghost;
    if constexpr (ND == 1) {
        if (!valid) return;
    }
    constexpr int ttd = 1 << ND, nslot = 2 * ND + 1;
    const int stride = ttd * 5;
    double u2[8][5];
    int g = -1;
    if (valid) {
        int k = 0;
        for (int j = 1; j < a.nseg; j++) k = t >= a.seg_t[j] ? j : k;
        const int pos = a.seg_o[k] + (t - a.seg_t[k]);
        g = a.order[pos] - a.gi0;
        const int boxmin[3] = {a.seg_box[k][0], a.seg_box[k][1], a.seg_box[k][2]};
        const int boxmax[3] = {a.seg_box[k][3], a.seg_box[k][4], a.seg_box[k][5]};
        double u1[7][5];
#pragma unroll
        for (int j = 0; j < nslot; j++) {
            const int code = a.src[(size_t)(2 * j) * a.src_ld + pos], ind = a.src[(size_t)(2 * j + 1) * a.src_ld + pos];
            // the source's oct key locates boundary ghosts only (code < 0)
            int okey[3] = {0, 0, 0};
            if (code < 0) {
                okey[0] = a.srckey[(g * 7 + j) * 3];
                okey[1] = a.srckey[(g * 7 + j) * 3 + 1];
                okey[2] = a.srckey[(g * 7 + j) * 3 + 2];
            }
            int c[3] = {(ind - 1) & 1, ((ind - 1) >> 1) & 1, ((ind - 1) >> 2) & 1};
            uint8_t rf;
            if constexpr (ND == 3) {
                // select-based flip: no runtime index into u1 (which would put it in scratch)
                const CellRef r =
                    cell_ref3(a.uold, a.refined, code, okey, boxmin, boxmax, reflexive_mask(a.btype), c);
                cell_load3(r, u1[j], rf);
                cell_flip(u1[j], r.flip);
            } else {
                gather_cell(a.uold, a.refined, code, okey, ND, boxmin, boxmax, a.btype, c, u1[j], rf);
            }
        }
        interpol_hydro_dev<ND>(u1, u2, a.interpol_var, a.interpol_type, a.smallr);
    } else {
#pragma unroll
        for (int c = 0; c < 8; c++)
#pragma unroll
            for (int n = 0; n < 5; n++) u2[c][n] = 0.0;
    }
    if constexpr (ND == 3) {
        // Stores staged through LDS one variable at a time: a ghost's variable
        // is one aligned 64-byte line, written by four lanes with 16 bytes
        // each, so a store instruction of the wave covers 16 whole lines
        // instead of touching 64 lines with 16 bytes each.
        __shared__ double2 sm[64 * 4];
        const int lane = threadIdx.x;
        int gq[4];
#pragma unroll
        for (int q = 0; q < 4; q++) gq[q] = __shfl(g, (q * 64 + lane) >> 2);
#pragma unroll
        for (int n = 0; n < 5; n++) {
            if (n) __syncthreads();
#pragma unroll
            for (int c = 0; c < 4; c++) sm[lane * 4 + c] = make_double2(u2[2 * c][n], u2[2 * c + 1][n]);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (gq[q] >= 0) {
                    double2 *p2 = reinterpret_cast<double2 *>(a.uold + (size_t)(a.first - 1 + gq[q]) * stride);
                    p2[n * 4 + (lane & 3)] = sm[q * 64 + lane];
                }
        }
    } else {
        double *p = a.uold + (size_t)(a.first - 1 + g) * stride;
#pragma unroll
        for (int c = 0; c < ttd; c++)
#pragma unroll
            for (int n = 0; n < 5; n++) p[n * ttd + c] = u2[c][n];
    }
}

hipError_t launch_refresh_ghosts(const GhostArgs &a, hipStream_t s)
{
    if (a.nghost <= 0) return hipSuccess;
    if (a.ndim == 3)
        hipLaunchKernelGGL(k_refresh_ghosts<3>, dim3((a.nghost + 63) / 64), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_refresh_ghosts<1>, dim3((a.nghost + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_upload(double *uold, const int32_t *father, int head, int noct, int ttd,
                                                int internal_energy, double smallr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noct) return;
    const int o = head - 1 + i;
    const int igrid = father[2 * i], icell = father[2 * i + 1];
    if (igrid <= 0) return;
    const double *c = uold + (size_t)o * 5 * ttd;
    double *p = uold + (size_t)(igrid - 1) * 5 * ttd + (icell - 1);
    for (int n = 0; n < 5; n++) {
        double avg = 0.0;
        for (int k = 0; k < ttd; k++) avg = avg + c[n * ttd + k];
        p[n * ttd] = avg / (double)ttd;
    }
    if (internal_energy) {
        double avg = 0.0;
        for (int k = 0; k < ttd; k++) {
            double ekin = 0.0;
            for (int d = 1; d <= 3; d++) ekin = ekin + 0.5 * (c[d * ttd + k] * c[d * ttd + k]) / fmax(c[k], smallr);
            avg = avg + c[4 * ttd + k] - ekin - 0.0 - 0.0;
        }
        double ekin = 0.0;
        for (int d = 1; d <= 3; d++) ekin = ekin + 0.5 * (p[d * ttd] * p[d * ttd]) / fmax(p[0], smallr);
        p[4 * ttd] = avg / (double)ttd + ekin + 0.0 + 0.0;
    }
}

// upload_fine with interpol_var = 0: one thread per (fine oct, variable),
// each averaging one contiguous 64-byte line of the oct.
__global__ void __launch_bounds__(256) k_upload_var(double *uold, const int32_t *__restrict__ father, int head,
                                                    int noct)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)noct * 5) return;
    const int io = (int)(i / 5), n = (int)(i % 5);
    const int igrid = father[2 * io], icell = father[2 * io + 1];
    if (igrid <= 0) return;
    // the oct's variable is one aligned 64-byte line: four 16-byte loads
    const double2 *c = reinterpret_cast<const double2 *>(uold + ((size_t)(head - 1 + io) * 5 + n) * 8);
    double avg = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double2 v = c[k];
        avg = avg + v.x;
        avg = avg + v.y;
    }
    uold[((size_t)(igrid - 1) * 5 + n) * 8 + (icell - 1)] = avg / 8.0;
}

hipError_t launch_upload(double *uold, const int32_t *father, int head, int noct, int ndim, int internal_energy,
                         double smallr, hipStream_t s)
{
    if (noct <= 0) return hipSuccess;
    if (ndim == 3 && !internal_energy) {
        const long n = (long)noct * 5;
        hipLaunchKernelGGL(k_upload_var, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, uold, father, head, noct);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_upload, dim3((noct + 255) / 256), dim3(256), 0, s, uold, father, head, noct, 1 << ndim,
                       internal_energy, smallr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
#define RH_BRICK_SHAPES(X) \
    X(7, 7, 8)             \
    X(7, 7, 16)            \
    X(8, 6, 8)             \
    X(8, 6, 16)            \
    X(4, 4, 4)             \
    X(4, 4, 8)             \
    X(15, 7, 8)            \
    X(15, 7, 16)           \
    X(7, 7, 4)             \
    X(3, 3, 4)             \
    X(3, 3, 8)             \
    X(1, 1, 1)

bool brick_shape_supported(int bx, int by, int bz)
{
#define RH_CHECK_SHAPE(A, B, C) \
    if (bx == A && by == B && bz == C) return true;
    RH_BRICK_SHAPES(RH_CHECK_SHAPE)
#undef RH_CHECK_SHAPE
    return false;
}

// Kernel launch, optionally with start/stop events tracked by the dispatch
// itself (hipExtLaunchKernelGGL): timing without separate event packets on
// the stream between the kernels of a step.
template <typename K, typename... Args>
static void launch_ev(K kernel, dim3 g, dim3 b, hipStream_t s, hipEvent_t e0, hipEvent_t e1, Args... args)
{
    if (e0 || e1)
        hipExtLaunchKernelGGL(kernel, g, b, 0, s, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, g, b, 0, s, args...);
}

// Half-brick split of the final round (GodArgs::nsplit).  With equal work
// per brick, q bricks per XCD on `slots` concurrent workgroups per XCD run in
// ceil(q / slots) rounds; when the last round holds rem = q mod slots <=
// slots/2 bricks, running those rem bricks as 2 rem half-bricks ends the
// launch about half a round earlier (e.g. 352 bricks per XCD on 64 slots:
// 5.53 instead of 6 rounds, by a list-scheduling model).
#ifndef RH_TAIL_SPLIT
#define RH_TAIL_SPLIT 1
#endif
template <typename K> static int tail_split(K kernel, int nthreads, const GodArgs &a)
{
    if (!RH_TAIL_SPLIT || !a.xcd_remap || a.reflux) return 0;
    static std::map<const void *, int> occ_cache;
    static int cus_per_xcd = 0;
    if (!cus_per_xcd) {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        cus_per_xcd = ncu / 8 > 0 ? ncu / 8 : 1;
    }
    auto it = occ_cache.find((const void *)kernel);
    int occ;
    if (it == occ_cache.end()) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, nthreads, 0) != hipSuccess) occ = 0;
        occ_cache[(const void *)kernel] = occ;
    } else {
        occ = it->second;
    }
    const int slots = occ * cus_per_xcd, q = a.nbricks >> 3;
    if (slots <= 1 || q <= 0) return 0;
    const int rem = q % slots;
    return rem > 0 && 2 * rem <= slots ? rem : 0;
}

template <typename K>
static void launch_brick(K kernel, int nt, const GodArgs &a0, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                         int *nparts)
{
    GodArgs a = a0;
    a.nsplit = tail_split(kernel, nt, a);
    const int q = a.nbricks >> 3, r = a.nbricks & 7;
    const dim3 g(a.nsplit ? 8 * (q + (r ? 1 : 0) + a.nsplit) : a.nbricks), b(nt);
    if (nparts) *nparts = a.nbricks + 8 * a.nsplit;
    launch_ev(kernel, g, b, s, e0, e1, a);
}

int g_specialized = 1;

hipError_t launch_godunov3d(const GodArgs &a, int bx, int by, int bz, int prec, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1, int *nparts)
{
    if (nparts) *nparts = a.nbricks > 0 ? a.nbricks : 0;
    if (a.nbricks <= 0) return hipSuccess;
    // the specialised body: periodic in every dimension, no refined cells,
    // no refluxes
    const bool gen = a.bnd || a.any_ref || a.reflux || a.ubase != a.uold || !g_specialized;
#define RH_LAUNCH_SHAPE(A, B, C)                                                                         \
    if (bx == A && by == B && bz == C) {                                                                 \
        constexpr int nt = (2 * A + 2) * (2 * B + 2);                                                     \
        if (a.slope == 3) {                                                                              \
            if (prec == 0)                                                                               \
                launch_brick(k_godunov3d<A, B, C, double, 0, 3>, nt, a, s, e0, e1, nparts);              \
            else                                                                                         \
                launch_brick(k_godunov3d<A, B, C, float, 0, 3>, nt, a, s, e0, e1, nparts);               \
        } else if (a.riemann == 3 && !gen) {                                                             \
            if (prec == 0)                                                                               \
                launch_brick(k_godunov3d<A, B, C, double, 3, 2, false>, nt, a, s, e0, e1, nparts);       \
            else                                                                                         \
                launch_brick(k_godunov3d<A, B, C, float, 3, 2, false>, nt, a, s, e0, e1, nparts);        \
        } else if (a.riemann == 3) {                                                                     \
            if (prec == 0)                                                                               \
                launch_brick(k_godunov3d<A, B, C, double, 3, 2>, nt, a, s, e0, e1, nparts);              \
            else                                                                                         \
                launch_brick(k_godunov3d<A, B, C, float, 3, 2>, nt, a, s, e0, e1, nparts);               \
        } else {                                                                                         \
            if (prec == 0)                                                                               \
                launch_brick(k_godunov3d<A, B, C, double, 0, 2>, nt, a, s, e0, e1, nparts);              \
            else                                                                                         \
                launch_brick(k_godunov3d<A, B, C, float, 0, 2>, nt, a, s, e0, e1, nparts);               \
        }                                                                                                \
        return hipGetLastError();                                                                        \
    }
    RH_BRICK_SHAPES(RH_LAUNCH_SHAPE)
#undef RH_LAUNCH_SHAPE
    return hipErrorInvalidValue;
}

hipError_t launch_godunov3d_wl(const WorkList &w, int bx, int by, int bz, int slope, int riemann,
                               int prec, hipStream_t s, bool gen)
{
    if (w.range[8] <= 0) return hipSuccess;
    gen = gen || !g_specialized;
    int maxr = 0;
    for (int r = 0; r < 8; r++) maxr = w.range[r + 1] - w.range[r] > maxr ? w.range[r + 1] - w.range[r] : maxr;
    const int nblocks = 8 * maxr;
#define RH_LAUNCH_WL(A, B, C)                                                                            \
    if (bx == A && by == B && bz == C) {                                                                 \
        constexpr int nt = (2 * A + 2) * (2 * B + 2);                                                     \
        const dim3 g(nblocks), b(nt);                                                                    \
        if (slope == 3) {                                                                                \
            if (prec == 0)                                                                               \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, double, 0, 3>), g, b, 0, s, w);              \
            else                                                                                         \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, float, 0, 3>), g, b, 0, s, w);               \
        } else if (riemann == 3 && !gen) {                                                               \
            if (prec == 0)                                                                               \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, double, 3, 2, false>), g, b, 0, s, w);       \
            else                                                                                         \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, float, 3, 2, false>), g, b, 0, s, w);        \
        } else if (riemann == 3) {                                                                       \
            if (prec == 0)                                                                               \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, double, 3, 2>), g, b, 0, s, w);              \
            else                                                                                         \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, float, 3, 2>), g, b, 0, s, w);               \
        } else {                                                                                         \
            if (prec == 0)                                                                               \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, double, 0, 2>), g, b, 0, s, w);              \
            else                                                                                         \
                hipLaunchKernelGGL((k_godunov3d_wl<A, B, C, float, 0, 2>), g, b, 0, s, w);               \
        }                                                                                                \
        return hipGetLastError();                                                                        \
    }
    RH_BRICK_SHAPES(RH_LAUNCH_WL)
#undef RH_LAUNCH_WL
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Artificial diffusion (difmag > 0): cmpdivu + cmpdiff, hydro/umuscl.f90:
// 2392-2540, called by unsplit at :273-283.  The diffusive flux of a face,
// dt * difmag * min(0, mean of the velocity divergence at its 2^(ndim-1)
// corner nodes) * (u_R - u_L), is linear in the update, so it runs as a pass
// after the Godunov launch: unew += (D_low - D_high) per dimension, D zeroed
// where either cell is refined (as the fluxes), refluxed into the parent cell
// across coarse-fine faces (as godfine1 refluxes flux + D).  Each cell
// computes its own faces; the two cells of a face compute the same D from the
// same values in the same order.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_diffusion3d(const DiffArgs a)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)a.noct * 8) return;
    const int o = a.head - 1 + (int)(i >> 3), ind = (int)(i & 7);
    const int c0[3] = {ind & 1, (ind >> 1) & 1, (ind >> 2) & 1};
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double gdt[3] = {a.g[0] * (dt * 0.5), a.g[1] * (dt * 0.5), a.g[2] * (dt * 0.5)};
    const int32_t *nb = a.nbor + (size_t)(o - (a.head - 1)) * 27;
    // velocities (with ctoprim's gravity kick) of the 3x3x3 cells around the
    // cell, the conservative states of the cell and its 6 face neighbours,
    // the refined flags, and the oct index of the face neighbours
    double q[3][3][3][3];
    double uf[7][5];
    uint8_t rf[7];
    int vf[7];
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int cc[3] = {c0[0] + dx, c0[1] + dy, c0[2] + dz};
                int oo[3], cb[3];
                for (int d = 0; d < 3; d++) {
                    oo[d] = cc[d] >> 1;  // -1, 0 or 1
                    cb[d] = cc[d] & 1;
                }
                const int v = nb[(oo[0] + 1) + 3 * (oo[1] + 1) + 9 * (oo[2] + 1)];
                const int okey[3] = {a.ckey[3 * o] + oo[0], a.ckey[3 * o + 1] + oo[1], a.ckey[3 * o + 2] + oo[2]};
                const CellRef r = cell_ref3(a.uold, a.refined, v, okey, a.boxmin, a.boxmax, a.refl, cb);
                double u[5];
                uint8_t f;
                cell_load3(r, u, f);
                cell_flip(u, r.flip);
                const double rho = u[0] > a.smallr ? u[0] : a.smallr;
                const double ir = 1.0 / rho;
                for (int d = 0; d < 3; d++) q[dz + 1][dy + 1][dx + 1][d] = u[1 + d] * ir + gdt[d];
                const int nz = (dx != 0) + (dy != 0) + (dz != 0);
                if (nz <= 1) {
                    // 0: the cell; 1..6: -x +x -y +y -z +z
                    const int k = nz == 0 ? 0 : dx ? (dx < 0 ? 1 : 2) : dy ? (dy < 0 ? 3 : 4) : (dz < 0 ? 5 : 6);
                    for (int n = 0; n < 5; n++) uf[k][n] = u[n];
                    rf[k] = f;
                    vf[k] = v;
                }
            }
    // divergence at the node (x, y, z) in {0,1}^3 of the cell's corners,
    // cmpdivu's expression order (node (i,j,k) uses cells i-1..i, j-1..j, k-1..k)
    const double fact = 0.25 / a.dx;  // half**(ndim-1)/dx
    auto Q = [&](int x, int y, int z, int d) { return q[z + 1][y + 1][x + 1][d]; };
    double dv[2][2][2];
    for (int z = 0; z < 2; z++)
        for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++) {
                double ux = 0.0, vy = 0.0, wz = 0.0;
                ux = ux + fact * (Q(x, y, z, 0) - Q(x - 1, y, z, 0));
                ux = ux + fact * (Q(x, y - 1, z, 0) - Q(x - 1, y - 1, z, 0));
                vy = vy + fact * (Q(x, y, z, 1) - Q(x, y - 1, z, 1) + Q(x - 1, y, z, 1) - Q(x - 1, y - 1, z, 1));
                ux = ux + fact * (Q(x, y, z - 1, 0) - Q(x - 1, y, z - 1, 0) + Q(x, y - 1, z - 1, 0) -
                                  Q(x - 1, y - 1, z - 1, 0));
                vy = vy + fact * (Q(x, y, z - 1, 1) - Q(x, y - 1, z - 1, 1) + Q(x - 1, y, z - 1, 1) -
                                  Q(x - 1, y - 1, z - 1, 1));
                wz = wz + fact * (Q(x, y, z, 2) - Q(x, y, z - 1, 2) + Q(x, y - 1, z, 2) - Q(x, y - 1, z - 1, 2) +
                                  Q(x - 1, y, z, 2) - Q(x - 1, y, z - 1, 2) + Q(x - 1, y - 1, z, 2) -
                                  Q(x - 1, y - 1, z - 1, 2));
                dv[z][y][x] = ux + vy + wz;
            }
    const double f4 = 0.25;  // half**(ndim-1)
    const bool myref = a.refined[(size_t)o * 8 + ind] != 0;
    double un[5];
    for (int n = 0; n < 5; n++) un[n] = a.unew[(size_t)o * 40 + n * 8 + ind];
    for (int d = 0; d < 3; d++) {
        double D[2][5];
        for (int side = 0; side < 2; side++) {
            // face at the cell's low (side 0) or high (side 1) boundary of dimension d
            double div1;
            if (d == 0) {
                div1 = f4 * dv[0][0][side];
                div1 = div1 + f4 * dv[0][1][side];
                div1 = div1 + f4 * (dv[1][0][side] + dv[1][1][side]);
            } else if (d == 1) {
                div1 = 0.0;
                div1 = div1 + f4 * (dv[0][side][0] + dv[0][side][1]);
                div1 = div1 + f4 * (dv[1][side][0] + dv[1][side][1]);
            } else {
                div1 = f4 * (dv[side][0][0] + dv[side][0][1] + dv[side][1][0] + dv[side][1][1]);
            }
            div1 = a.difmag * (div1 < 0.0 ? div1 : 0.0);
            const int k = 1 + 2 * d + side;
            const bool z = myref || rf[k];
            for (int n = 0; n < 5; n++) {
                const double du = side == 0 ? uf[0][n] - uf[k][n] : uf[k][n] - uf[0][n];
                D[side][n] = z ? 0.0 : dt * div1 * du;
            }
            // coarse-fine face: the neighbour is an interpolated ghost oct
            if (a.reflux && !z && vf[k] > a.ngridmax) {
                const int sub = d == 0 ? ((ind >> 1) & 1) + 2 * ((ind >> 2) & 1)
                              : d == 1 ? (ind & 1) + 2 * ((ind >> 2) & 1)
                                       : (ind & 1) + 2 * ((ind >> 1) & 1);
                // ghost on the low side sends -D/8 (its high face slot), on the high side +D/8
                reflux_add(a.unew, a.ghost_parent, a.ngridmax, vf[k], side == 0 ? -1.0 : 1.0, D[side], 8, a.racc,
                           a.rdet, side == 0 ? 4 * (2 * d + 1) + sub : 4 * (2 * d) + sub, a.rdet != 0);
            }
        }
        for (int n = 0; n < 5; n++) un[n] = un[n] + (D[0][n] - D[1][n]);
    }
    for (int n = 0; n < 5; n++) a.unew[(size_t)o * 40 + n * 8 + ind] = un[n];
}

hipError_t launch_diffusion3d(const DiffArgs &a, hipStream_t s)
{
    const long n = (long)a.noct * 8;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_diffusion3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_godunov1d(const God1Args &a, int prec, hipStream_t s)
{
    const int n = 2 * a.noct;
    if (n <= 0) return hipSuccess;
    const int nb = (n + 255) / 256;
    if (prec == 0)
        hipLaunchKernelGGL((k_godunov1d<double>), dim3(nb), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_godunov1d<float>), dim3(nb), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_courant_reduce(const double *partials, int nparts, double dtmax, double *out, double *dt_inout,
                                 double *t_acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(256), 0, s, partials, nparts, dtmax, out, dt_inout, t_acc, 1);
    return hipGetLastError();
}

hipError_t launch_courant_reduce_batch(const CourantJob *jobs, int njobs, hipStream_t s)
{
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_courant_final_batch, dim3(njobs), dim3(256), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_courant_ml(const double *uold, const uint8_t *refined, const int32_t *lev, int head, int noct,
                             double boxlen, double dx_finest, double gamma, double smallr, double smallc,
                             double courant_factor, const double g[3], double *partials, int max_blocks, double *out,
                             double *dt_out, hipStream_t s, double *t_acc, bool mixed)
{
    CourantK c;
    c.dx = dx_finest;
    c.vol = 0.0;
    c.gamma = gamma;
    c.smallr = smallr;
    c.smallc = smallc;
    c.courant_factor = courant_factor;
    c.smallp = (smallc * smallc) / gamma;
    double gs = 0.0;
    for (int d = 0; d < 3; d++) gs = gs + __builtin_fabs(g[d]);
    c.gsum_dx = gs;  // times dx per oct in the kernel
    c.ndim = 3;
    const long ncell = (long)noct * 8;
    int nb = (int)((ncell + kCourantThreads - 1) / kCourantThreads);
    nb = nb < max_blocks ? nb : max_blocks;
    nb = nb < 1 ? 1 : nb;
    if (mixed)
        hipLaunchKernelGGL(k_courant_ml<true>, dim3(nb), dim3(kCourantThreads), 0, s, uold, refined, lev, head, noct, c,
                           boxlen, partials);
    else
        hipLaunchKernelGGL(k_courant_ml<false>, dim3(nb), dim3(kCourantThreads), 0, s, uold, refined, lev, head, noct,
                           c, boxlen, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(256), 0, s, partials, nb, dx_finest / smallc, out, dt_out,
                       t_acc, 1);
    return hipGetLastError();
}

hipError_t launch_zero(double *p, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long nb = (n + 255) / 256;
    nb = nb < 4096 ? nb : 4096;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_reflux_fold(double *unew, double *racc, const int32_t *ghost_parent, const int32_t *order,
                              int nghost, int ttd, hipStream_t s, int nslot)
{
    if (nghost <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reflux_fold, dim3((nghost + 255) / 256), dim3(256), 0, s, unew, racc, ghost_parent, order,
                       nghost, ttd, nslot);
    return hipGetLastError();
}

hipError_t launch_courant(const double *uold, const uint8_t *refined, int head, int noct, int ndim, double dx,
                          double gamma, double smallr, double smallc, double courant_factor, const double g[3],
                          double *partials, int max_blocks, double *out, double *dt_out, hipStream_t s,
                          double *t_acc)
{
    CourantK c;
    c.dx = dx;
    c.vol = ndim == 1 ? dx : dx * dx * dx;
    c.gamma = gamma;
    c.smallr = smallr;
    c.smallc = smallc;
    c.courant_factor = courant_factor;
    c.smallp = (smallc * smallc) / gamma;
    double gs = 0.0;
    for (int d = 0; d < ndim; d++) gs = gs + __builtin_fabs(g[d]);
    c.gsum_dx = gs * dx;
    c.ndim = ndim;
    const int ttd = 1 << ndim;
    const long ncell = (long)noct * ttd;
    int nb = (int)((ncell + kCourantThreads - 1) / kCourantThreads);
    if (nb > max_blocks) nb = max_blocks;
    if (nb < 1) nb = 1;
    if (ttd == 8)
        hipLaunchKernelGGL(k_courant<8>, dim3(nb), dim3(kCourantThreads), 0, s, uold, refined, head, noct, c, partials);
    else
        hipLaunchKernelGGL(k_courant<2>, dim3(nb), dim3(kCourantThreads), 0, s, uold, refined, head, noct, c, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(256), 0, s, partials, nb, dx / smallc, out, dt_out, t_acc, 0);
    return hipGetLastError();
}

}  // namespace rh
