// Device physics of the MUSCL-Hancock Godunov step, templated on the
// reconstruction/flux precision T (double for the FP64 kernel, float for the
// mixed kernel).  Conversion to primitive variables (the pressure recovery
// p=(gamma-1)(E-rho v^2/2)) is always done in double; see rh_prim_from_cons.
//
// Each routine restates the reference CPU routine it names, in the same
// expression order, so that the FP64 instantiation agrees with the reference
// to rounding (FMA contraction aside).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rh {

// Scalar helpers -----------------------------------------------------------------
// max/min as single v_max/v_min instructions (the reference's MAX/MIN; the two
// differ only in the sign of a zero result and for NaN inputs)
__device__ __forceinline__ double tmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float tmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float tmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double tcopysign(double a, double b) { return __builtin_copysign(a, b); }
__device__ __forceinline__ float tcopysign(float a, float b) { return __builtin_copysignf(a, b); }
__device__ __forceinline__ double fabs_(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float fabs_(float a) { return __builtin_fabsf(a); }

// Division and square root: IEEE in double; the hardware reciprocal and square
// root (about 1 ulp) in float, where the mixed kernel tolerates 1e-7 relative.
#ifndef RH_FAST_F64_DIV
#define RH_FAST_F64_DIV 1
#endif
// 1/b from the hardware reciprocal refined by two Newton steps (5 instructions
// instead of the ~10 of the IEEE division sequence); within 1 ulp for the
// normal-range operands of the hydro kernels (no range scaling).  Used for the
// divisions of the Godunov kernels (RH_FAST_F64_DIV, measured 2.5% faster on
// the FP64 kernel); the Courant dt keeps IEEE division.
__device__ __forceinline__ double rcp_nr(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double tdiv(double a, double b) { return RH_FAST_F64_DIV ? a * rcp_nr(b) : a / b; }
__device__ __forceinline__ float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double tsqrt(double a) { return __builtin_sqrt(a); }
// reciprocal used for a denominator shared by several quotients: a/b is then
// formed as a*(1/b), within an ulp of the reference's separate divisions
__device__ __forceinline__ double trcp(double b) { return RH_FAST_F64_DIV ? rcp_nr(b) : 1.0 / b; }
__device__ __forceinline__ float trcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float tsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }

// Square root of an argument known to lie in [2^-900, 2^1000] (the sound
// speeds of the Riemann solvers, whose argument is floored by smallc^2, see
// make_k): the hardware reciprocal square root refined by the Goldschmidt
// iteration (Markstein, "IA-64 and Elementary Functions", 2000, ch. 9): one
// coupled step for g ~ sqrt(x) and h ~ 1/(2 sqrt(x)), then two residual
// corrections d = x - g^2.  Ten instructions; it skips the range scaling and
// the zero/infinity/NaN fix-ups of the general double sqrt (about 17), which
// this range never needs, and returns the same correctly rounded results there.
__device__ __forceinline__ double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
__device__ __forceinline__ float sqrt_pos(float x) { return __builtin_amdgcn_sqrtf(x); }

// Constants of one Godunov launch (hydro_params_t + derived quantities).
template <typename T> struct HydroK {
    T gamma;
    T smallr;
    T smallc2;        // smallc**2, floored at 2^-900 (the range of sqrt_pos)
    T smallp_trace;   // smallr*smallc**2       (trace3d, umuscl.f90:934)
    T smallp_riem;    // smallc**2/gamma        (riemann_*, godunov_utils.f90:392)
    T entho;          // one/(gamma-one)
    T dtdx;           // dt/dx                  (trace3d, umuscl.f90:929)
    T smax_cap;       // courant_factor*dx/dt   (umuscl.f90:99)
    int slope_type;
    int riemann;
};

// ctoprim, hydro/umuscl.f90:1739-1901, for one cell, in double.
// u: conservative (rho, rho u, rho v, rho w, E); q: (rho, u, v, w, p).
// FAST (the mixed-precision kernel, whose primitives are rounded to FP32)
// forms 1/rho with the Newton-refined hardware reciprocal.
template <bool FAST = false>
__device__ __forceinline__ void prim_from_cons(const double u[5], double q[5], double smallr, double smalle,
                                               double gamma, const double gdt_half[3])
{
    double r = u[0] > smallr ? u[0] : smallr;
    double oneoverrho = FAST ? rcp_nr(r) : trcp(r);
    double vx = u[1] * oneoverrho;
    double vy = u[2] * oneoverrho;
    double vz = u[3] * oneoverrho;
    double ekin = 0.5 * r * (vx * vx);
    ekin = ekin + 0.5 * r * (vy * vy);
    ekin = ekin + 0.5 * r * (vz * vz);
    double e = u[4] - ekin;
    double se = smalle * r;
    double eint = e > se ? e : se;
    q[0] = r;
    q[1] = vx + gdt_half[0];
    q[2] = vy + gdt_half[1];
    q[3] = vz + gdt_half[2];
    q[4] = (gamma - 1.0) * eint;
}

// uslope, hydro/umuscl.f90:1906-2387: one-dimensional limited slope,
// returned HALVED (dq/2, the quantity the Hancock predictor uses; the halving
// is exact, so the trace sees the values of the reference's 0.5*dq).
// slope_type 1 = MinMod (:2132-2170), 2 = MonCen (:2171-2206).  With
// a = q0-qm1, b = qp1-q0:
//   MinMod  dq = a, b or 0: the one of smaller magnitude when a*b > 0;
//   MonCen  dq = dsgn*min(2 min(|a|,|b|), |a+b|/2) when a*b > 0, else 0.
// Both are formed branch-free as max(min(...), 0) + min(max(...), 0): when
// all operands share a sign one term is the limited value and the other 0,
// and when a and b differ in sign both terms are 0.  The values equal the
// reference's (min/max are exact), apart from the sign of a zero slope and
// slopes below 1e-154, where the reference's product a*b underflows to 0.
template <typename T> __device__ __forceinline__ T slope_half(int st, T qm1, T q0, T qp1)
{
    const T a = q0 - qm1;
    const T b = qp1 - q0;
    if (st == 1) {
        const T lo = tmax(tmin(a, b), T(0)), hi = tmin(tmax(a, b), T(0));
        return T(0.5) * (lo + hi);
    }
    const T c = (a + b) * T(0.25);  // dcen/2 = ((2a + 2b)/2)/2/2
    const T lo = tmax(tmin(tmin(a, b), c), T(0));
    const T hi = tmin(tmax(tmax(a, b), c), T(0));
    return lo + hi;
}

// Normal-velocity slope bound, hydro/umuscl.f90:2361-2385, on a halved slope:
// |dq| <= 2 min(rho_L |du_L|, rho_R |du_R|)/rho becomes
// |dq/2| <= min(rho_L |du_L|, rho_R |du_R|)/rho.
template <typename T>
__device__ __forceinline__ T vel_bound_half(T dqh, T rl, T ul, T inv_r0, T u0, T rr, T ur)
{
    T dlft = rl * fabs_(u0 - ul);
    T drgt = rr * fabs_(ur - u0);
    T dlimh = tmin(dlft, drgt) * inv_r0;
    return tcopysign(tmin(fabs_(dqh), dlimh), dqh);
}

// trace3d, hydro/umuscl.f90:860-1466, hydro part for one cell.
// q: cell primitives; dqh[d][n]: halved slopes (the reference's 0.5*dq).
// Outputs qp[d][n] (right state at the left interface) and qm[d][n] (left
// state at the right interface).
template <typename T>
__device__ __forceinline__ void trace3(const T q[5], const T dqh[3][5], T qp[3][5], T qm[3][5], const HydroK<T> &k,
                                       T inv_r)
{
    T r = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
    T drx = dqh[0][0], dux = dqh[0][1], dvx = dqh[0][2], dwx = dqh[0][3], dpx = dqh[0][4];
    T dry = dqh[1][0], duy = dqh[1][1], dvy = dqh[1][2], dwy = dqh[1][3], dpy = dqh[1][4];
    T drz = dqh[2][0], duz = dqh[2][1], dvz = dqh[2][2], dwz = dqh[2][3], dpz = dqh[2][4];
    T divv = dux + dvy + dwz;
    T sr0 = -u * drx - v * dry - w * drz - divv * r;
    T su0 = -u * dux - v * duy - w * duz - dpx * inv_r;
    T sv0 = -u * dvx - v * dvy - w * dvz - dpy * inv_r;
    T sw0 = -u * dwx - v * dwy - w * dwz - dpz * inv_r;
    T sp0 = -u * dpx - v * dpy - w * dpz - divv * k.gamma * p;
    T ws[5];
    ws[0] = r + sr0 * k.dtdx;
    ws[1] = u + su0 * k.dtdx;
    ws[2] = v + sv0 * k.dtdx;
    ws[3] = w + sw0 * k.dtdx;
    ws[4] = p + sp0 * k.dtdx;
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
            T h = dqh[d][n];
            qp[d][n] = ws[n] - h;
            qm[d][n] = ws[n] + h;
        }
        if (qp[d][0] < k.smallr) qp[d][0] = r;
        if (qp[d][4] < k.smallp_trace) qp[d][4] = p;
        if (qm[d][0] < k.smallr) qm[d][0] = r;
        if (qm[d][4] < k.smallp_trace) qm[d][4] = p;
    }
}

// trace1d, hydro/umuscl.f90:297-481, on halved slopes dqh.
template <typename T>
__device__ __forceinline__ void trace1(const T q[5], const T dqh[5], T qp[5], T qm[5], const HydroK<T> &k, T inv_r)
{
    T r = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
    T drx = dqh[0], dux = dqh[1], dvx = dqh[2], dwx = dqh[3], dpx = dqh[4];
    T sr0 = -u * drx - dux * r;
    T sp0 = -u * dpx - dux * k.gamma * p;
    T su0 = -u * dux - dpx * inv_r;
    T sv0 = -u * dvx;
    T sw0 = -u * dwx;
    T ws[5];
    ws[0] = r + sr0 * k.dtdx;
    ws[1] = u + su0 * k.dtdx;
    ws[2] = v + sv0 * k.dtdx;
    ws[3] = w + sw0 * k.dtdx;
    ws[4] = p + sp0 * k.dtdx;
    T h[5] = {drx, dux, dvx, dwx, dpx};
#pragma unroll
    for (int n = 0; n < 5; n++) {
        qp[n] = ws[n] - h[n];
        qm[n] = ws[n] + h[n];
    }
    if (qp[0] < k.smallr) qp[0] = r;
    if (qp[4] < k.smallp_trace) qp[4] = p;
    if (qm[0] < k.smallr) qm[0] = r;
    if (qm[4] < k.smallp_trace) qm[4] = p;
}

// riemann_hllc, hydro/godunov_utils.f90:638-857 (mass, momenta, energy fluxes).
// Only the star state of the side that is sampled is formed; the sampled
// values are those of the reference, which forms both.
template <typename T>
__device__ __forceinline__ void riemann_hllc(const T ql[5], const T qr[5], T f[5], const HydroK<T> &k)
{
    const T half = T(0.5);
    T rl = tmax(ql[0], k.smallr);
    T ul = ql[1], vl = ql[2], wl = ql[3];
    T Pl = tmax(ql[4], rl * k.smallp_riem);
    T el = Pl * k.entho;
    T etotl = el + half * rl * (ul * ul + vl * vl + wl * wl);
    T rr = tmax(qr[0], k.smallr);
    T ur = qr[1], vr = qr[2], wr = qr[3];
    T Pr = tmax(qr[4], rr * k.smallp_riem);
    T er = Pr * k.entho;
    T etotr = er + half * rr * (ur * ur + vr * vr + wr * wr);
    // max(sqrt(a), sqrt(b)) = sqrt(max(a, b)) for a correctly rounded
    // (monotonic) square root: one root instead of the two sound speeds
    T cmax = sqrt_pos(tmax(tmax(tdiv(k.gamma * Pl, rl), tdiv(k.gamma * Pr, rr)), k.smallc2));
    T SL = tmin(ul, ur) - cmax;
    T SR = tmax(ul, ur) + cmax;
    T rcl = rl * (ul - SL);
    T rcr = rr * (SR - ur);
    T inv_den = trcp(rcr + rcl);
    T ustar = (rcr * ur + rcl * ul + (Pl - Pr)) * inv_den;
    T Ptotstar = (rcr * Pl + rcl * Pr + rcl * rcr * (ul - ur)) * inv_den;
    // The reference samples, in this order: SL > 0 -> left state; ustar > 0
    // -> left star state; SR > 0 -> right star state; else right state.  The
    // side whose values are used is left exactly when SL > 0 or ustar > 0,
    // and the star state is taken when SL <= 0 and (ustar > 0 or SR > 0); in
    // both star cases the side is the reference's ustar > 0.  So one select
    // of the side's seven values and one of four values (state or star) give
    // the reference's result (its chain costs three selects per value).
    const bool sl_pos = SL > T(0), us_pos = ustar > T(0);
    const bool left = sl_pos || us_pos;
    const bool star = !sl_pos && (us_pos || SR > T(0));
    T S = left ? SL : SR;
    T rs = left ? rl : rr, us = left ? ul : ur, Ps = left ? Pl : Pr, es = left ? etotl : etotr;
    T vo = left ? vl : vr, wo = left ? wl : wr;
    T inv_sd = trcp(S - ustar);
    T rstar = rs * (S - us) * inv_sd;
    T etotstar = ((S - us) * es - Ps * us + Ptotstar * ustar) * inv_sd;
    T ro = star ? rstar : rs, uo = star ? ustar : us, Po = star ? Ptotstar : Ps, eo = star ? etotstar : es;
    f[0] = ro * uo;
    f[1] = ro * uo * uo + Po;
    f[2] = ro * uo * vo;
    f[3] = ro * uo * wo;
    f[4] = (eo + Po) * uo;
}

// Conservative state and physical flux of one side (riemann_llf/hll,
// godunov_utils.f90:420-470 and :565-615).
template <typename T>
__device__ __forceinline__ void side_state(const T q[5], T entho, T u[5])
{
    const T half = T(0.5);
    u[0] = q[0];
    u[1] = q[0] * q[1];
    u[2] = q[0] * q[2];
    u[3] = q[0] * q[3];
    T e = q[4] * entho + half * q[0] * (q[1] * q[1]);
    e = e + half * q[0] * (q[2] * q[2]);
    u[4] = e + half * q[0] * (q[3] * q[3]);
}

// riemann_llf, hydro/godunov_utils.f90:338-485.
template <typename T>
__device__ __forceinline__ void riemann_llf(const T ql[5], const T qr[5], T f[5], const HydroK<T> &k)
{
    T rl = tmax(ql[0], k.smallr), pl = tmax(ql[4], rl * k.smallp_riem);
    T cl = tsqrt(tdiv(k.gamma * pl, rl));
    T rr = tmax(qr[0], k.smallr), pr = tmax(qr[4], rr * k.smallp_riem);
    T cr = tsqrt(tdiv(k.gamma * pr, rr));
    T cmax = tmin(tmax(fabs_(ql[1]) + cl, fabs_(qr[1]) + cr), k.smax_cap);
    T uL[5], uR[5];
    side_state(ql, k.entho, uL);
    side_state(qr, k.entho, uR);
    T fL[5], fR[5];
    fL[0] = ql[1] * uL[0];
    fR[0] = qr[1] * uR[0];
    fL[1] = ql[1] * uL[1] + ql[4];
    fR[1] = qr[1] * uR[1] + qr[4];
    fL[2] = ql[1] * uL[2];
    fR[2] = qr[1] * uR[2];
    fL[3] = ql[1] * uL[3];
    fR[3] = qr[1] * uR[3];
    fL[4] = ql[1] * (uL[4] + ql[4]);
    fR[4] = qr[1] * (uR[4] + qr[4]);
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = T(0.5) * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
}

// riemann_hll, hydro/godunov_utils.f90:486-637.
template <typename T>
__device__ __forceinline__ void riemann_hll(const T ql[5], const T qr[5], T f[5], const HydroK<T> &k)
{
    T rl = tmax(ql[0], k.smallr), ul = ql[1], pl = tmax(ql[4], rl * k.smallp_riem);
    T cl = tsqrt(tdiv(k.gamma * pl, rl));
    T rr = tmax(qr[0], k.smallr), ur = qr[1], pr = tmax(qr[4], rr * k.smallp_riem);
    T cr = tsqrt(tdiv(k.gamma * pr, rr));
    T SL = tmax(tmin(tmin(ul, ur) - tmax(cl, cr), T(0)), -k.smax_cap);
    T SR = tmin(tmax(tmax(ul, ur) + tmax(cl, cr), T(0)), k.smax_cap);
    T uL[5], uR[5];
    side_state(ql, k.entho, uL);
    side_state(qr, k.entho, uR);
    T fL[5], fR[5];
    fL[0] = uL[1];
    fR[0] = uR[1];
    fL[1] = ql[4] + uL[1] * ql[1];
    fR[1] = qr[4] + uR[1] * qr[1];
    fL[2] = ql[1] * uL[2];
    fR[2] = qr[1] * uR[2];
    fL[3] = ql[1] * uL[3];
    fR[3] = qr[1] * uR[3];
    fL[4] = ql[1] * (uL[4] + ql[4]);
    fR[4] = qr[1] * (uR[4] + qr[4]);
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = tdiv(SR * fL[n] - SL * fR[n] + SR * SL * (uR[n] - uL[n]), SR - SL);
}

// cmpflxm, hydro/umuscl.f90:1472-1607: rotate to the face normal idim, solve,
// rotate back.  ql/qr are (rho,u,v,w,p) in grid axes.
// RS: 3 = HLLC at compile time, 0 = the solver named by k.riemann.
template <int RS = 0, typename T>
__device__ __forceinline__ void face_flux(int idim, const T ql[5], const T qr[5], T f[5], const HydroK<T> &k)
{
    const int ln = idim + 1;
    const int lt1 = idim == 0 ? 2 : 1;
    const int lt2 = idim == 2 ? 2 : 3;
    T a[5] = {ql[0], ql[ln], ql[lt1], ql[lt2], ql[4]};
    T b[5] = {qr[0], qr[ln], qr[lt1], qr[lt2], qr[4]};
    T g[5];
    if (RS == 3 || k.riemann == 3)
        riemann_hllc(a, b, g, k);
    else if (k.riemann == 1)
        riemann_llf(a, b, g, k);
    else
        riemann_hll(a, b, g, k);
    f[0] = g[0];
    f[ln] = g[1];
    f[lt1] = g[2];
    f[lt2] = g[3];
    f[4] = g[4];
}

}  // namespace rh
