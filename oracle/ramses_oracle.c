/*
 * CPU oracle: plain-C restatement of the RAMSES II CPU hydro path.
 *
 * TEST INFRASTRUCTURE ONLY (see ramses_oracle.h).  Every function cites the
 * reference routine it restates; expressions keep the Fortran evaluation order
 * (left to right, parenthesised as written) so that, compiled with
 * -ffp-contract=off, results agree bit for bit with the flang-built reference.
 */
#include "ramses_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALF 0.5
#define ONE 1.0
#define TWO 2.0
#define ZERO 0.0

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
/* Fortran SIGN(a,b) */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------------------ */
/* Riemann solvers: hydro/godunov_utils.f90                                  */
/* ------------------------------------------------------------------------ */

/* riemann_llf, hydro/godunov_utils.f90:338-485 */
void ora_riemann_llf(const double *ql, const double *qr, double *f, const ora_params_t *p)
{
    double gamma = p->gamma, smallr = p->smallr, smallc = p->smallc;
    double smallp = (smallc * smallc) / gamma;
    double entho = ONE / (gamma - ONE);
    double rl = dmax(ql[0], smallr), ul = ql[1], pl = dmax(ql[4], rl * smallp);
    double cl = gamma * pl;
    cl = sqrt(cl / rl);
    double rr = dmax(qr[0], smallr), ur = qr[1], pr = dmax(qr[4], rr * smallp);
    double cr = gamma * pr;
    cr = sqrt(cr / rr);
    double cmax = dmin(dmax(fabs(ul) + cl, fabs(ur) + cr), p->smax_cap);
    double uL[6], uR[6], fL[6], fR[6];
    uL[0] = ql[0];
    uR[0] = qr[0];
    uL[1] = ql[0] * ql[1];
    uR[1] = qr[0] * qr[1];
    for (int n = 2; n < 4; n++) {
        uL[n] = ql[0] * ql[n];
        uR[n] = qr[0] * qr[n];
    }
    uL[4] = ql[4] * entho + HALF * ql[0] * (ql[1] * ql[1]);
    uR[4] = qr[4] * entho + HALF * qr[0] * (qr[1] * qr[1]);
    uL[4] = uL[4] + HALF * ql[0] * (ql[2] * ql[2]);
    uR[4] = uR[4] + HALF * qr[0] * (qr[2] * qr[2]);
    uL[4] = uL[4] + HALF * ql[0] * (ql[3] * ql[3]);
    uR[4] = uR[4] + HALF * qr[0] * (qr[3] * qr[3]);
    uL[5] = ql[4] * entho;
    uR[5] = qr[4] * entho;
    fL[0] = ql[1] * uL[0];
    fR[0] = qr[1] * uR[0];
    fL[1] = ql[1] * uL[1] + ql[4];
    fR[1] = qr[1] * uR[1] + qr[4];
    for (int n = 2; n < 4; n++) {
        fL[n] = ql[1] * uL[n];
        fR[n] = qr[1] * uR[n];
    }
    fL[4] = ql[1] * (uL[4] + ql[4]);
    fR[4] = qr[1] * (uR[4] + qr[4]);
    fL[5] = ql[1] * uL[5];
    fR[5] = qr[1] * uR[5];
    for (int n = 0; n < 6; n++) f[n] = HALF * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
}

/* riemann_hll, hydro/godunov_utils.f90:486-637 */
void ora_riemann_hll(const double *ql, const double *qr, double *f, const ora_params_t *p)
{
    double gamma = p->gamma, smallr = p->smallr, smallc = p->smallc;
    double smallp = (smallc * smallc) / gamma;
    double entho = ONE / (gamma - ONE);
    double rl = dmax(ql[0], smallr), ul = ql[1], pl = dmax(ql[4], rl * smallp);
    double cl = gamma * pl;
    cl = sqrt(cl / rl);
    double rr = dmax(qr[0], smallr), ur = qr[1], pr = dmax(qr[4], rr * smallp);
    double cr = gamma * pr;
    cr = sqrt(cr / rr);
    double SL = dmax(dmin(dmin(ul, ur) - dmax(cl, cr), ZERO), -p->smax_cap);
    double SR = dmin(dmax(dmax(ul, ur) + dmax(cl, cr), ZERO), p->smax_cap);
    double uL[6], uR[6], fL[6], fR[6];
    uL[0] = ql[0];
    uR[0] = qr[0];
    uL[1] = ql[0] * ql[1];
    uR[1] = qr[0] * qr[1];
    for (int n = 2; n < 4; n++) {
        uL[n] = ql[0] * ql[n];
        uR[n] = qr[0] * qr[n];
    }
    uL[4] = ql[4] * entho + HALF * ql[0] * (ql[1] * ql[1]);
    uR[4] = qr[4] * entho + HALF * qr[0] * (qr[1] * qr[1]);
    uL[4] = uL[4] + HALF * ql[0] * (ql[2] * ql[2]);
    uR[4] = uR[4] + HALF * qr[0] * (qr[2] * qr[2]);
    uL[4] = uL[4] + HALF * ql[0] * (ql[3] * ql[3]);
    uR[4] = uR[4] + HALF * qr[0] * (qr[3] * qr[3]);
    uL[5] = ql[4] * entho;
    uR[5] = qr[4] * entho;
    fL[0] = uL[1];
    fR[0] = uR[1];
    fL[1] = ql[4] + uL[1] * ql[1];
    fR[1] = qr[4] + uR[1] * qr[1];
    for (int n = 2; n < 4; n++) {
        fL[n] = ql[1] * uL[n];
        fR[n] = qr[1] * uR[n];
    }
    fL[4] = ql[1] * (uL[4] + ql[4]);
    fR[4] = qr[1] * (uR[4] + qr[4]);
    fL[5] = ql[1] * uL[5];
    fR[5] = qr[1] * uR[5];
    for (int n = 0; n < 6; n++)
        f[n] = (SR * fL[n] - SL * fR[n] + SR * SL * (uR[n] - uL[n])) / (SR - SL);
}

/* riemann_hllc, hydro/godunov_utils.f90:638-857 */
void ora_riemann_hllc(const double *ql, const double *qr, double *f, const ora_params_t *p)
{
    double gamma = p->gamma, smallr = p->smallr, smallc = p->smallc;
    double smallp = (smallc * smallc) / gamma;
    double entho = ONE / (gamma - ONE);

    double rl = dmax(ql[0], smallr);
    double ul = ql[1], vl = ql[2], wl = ql[3];
    double Pl = dmax(ql[4], rl * smallp);
    double el = Pl * entho;
    double ekinl = HALF * rl * (ul * ul + vl * vl + wl * wl);
    double etotl = el + ekinl;
    double Ptotl = Pl;

    double rr = dmax(qr[0], smallr);
    double ur = qr[1], vr = qr[2], wr = qr[3];
    double Pr = dmax(qr[4], rr * smallp);
    double er = Pr * entho;
    double ekinr = HALF * rr * (ur * ur + vr * vr + wr * wr);
    double etotr = er + ekinr;
    double Ptotr = Pr;

    double cfastl = gamma * Pl;
    cfastl = sqrt(dmax(cfastl / rl, smallc * smallc));
    double cfastr = gamma * Pr;
    cfastr = sqrt(dmax(cfastr / rr, smallc * smallc));

    double SL = dmin(ul, ur) - dmax(cfastl, cfastr);
    double SR = dmax(ul, ur) + dmax(cfastl, cfastr);

    double rcl = rl * (ul - SL);
    double rcr = rr * (SR - ur);

    double ustar = (rcr * ur + rcl * ul + (Ptotl - Ptotr)) / (rcr + rcl);
    double Ptotstar = (rcr * Ptotl + rcl * Ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);

    double rstarl = rl * (SL - ul) / (SL - ustar);
    double etotstarl = ((SL - ul) * etotl - Ptotl * ul + Ptotstar * ustar) / (SL - ustar);
    double estarl = el * (SL - ul) / (SL - ustar);

    double rstarr = rr * (SR - ur) / (SR - ustar);
    double etotstarr = ((SR - ur) * etotr - Ptotr * ur + Ptotstar * ustar) / (SR - ustar);
    double estarr = er * (SR - ur) / (SR - ustar);

    double ro, uo, vo, wo, Ptoto, etoto, eo;
    if (SL > 0.0) {
        ro = rl; uo = ul; vo = vl; wo = wl; Ptoto = Ptotl; etoto = etotl; eo = el;
    } else if (ustar > 0.0) {
        ro = rstarl; uo = ustar; vo = vl; wo = wl; Ptoto = Ptotstar; etoto = etotstarl; eo = estarl;
    } else if (SR > 0.0) {
        ro = rstarr; uo = ustar; vo = vr; wo = wr; Ptoto = Ptotstar; etoto = etotstarr; eo = estarr;
    } else {
        ro = rr; uo = ur; vo = vr; wo = wr; Ptoto = Ptotr; etoto = etotr; eo = er;
    }
    f[0] = ro * uo;
    f[1] = ro * uo * uo + Ptoto;
    f[2] = ro * uo * vo;
    f[3] = ro * uo * wo;
    f[4] = (etoto + Ptoto) * uo;
    f[5] = uo * eo;
}

/* ------------------------------------------------------------------------ */
/* unsplit and its helpers: hydro/umuscl.f90                                 */
/* ------------------------------------------------------------------------ */

/* Bounds of hydro_kernel_t, hydro/hydro_commons.f90 init_hydro_kernel */
typedef struct {
    int ndim, nn;
    int iu1, iu2, ju1, ju2, ku1, ku2;
    int if1, if2, jf1, jf2, kf1, kf2;
    int ni, nj, nk;    /* extents of the u-arrays */
    int fi, fj, fk;    /* extents of the flux array */
} kbounds_t;

static void kbounds_init(kbounds_t *b, int ndim, int nn)
{
    b->ndim = ndim;
    b->nn = nn;
    b->iu1 = -1;
    b->iu2 = nn + 2;
    b->ju1 = (1 - ndim / 2) - 1 * (ndim / 2);
    b->ju2 = (1 - ndim / 2) + (nn + 2) * (ndim / 2);
    b->ku1 = (1 - ndim / 3) - 1 * (ndim / 3);
    b->ku2 = (1 - ndim / 3) + (nn + 2) * (ndim / 3);
    b->if1 = 1;
    b->if2 = nn + 1;
    b->jf1 = 1;
    b->jf2 = (1 - ndim / 2) + (nn + 1) * (ndim / 2);
    b->kf1 = 1;
    b->kf2 = (1 - ndim / 3) + (nn + 1) * (ndim / 3);
    b->ni = b->iu2 - b->iu1 + 1;
    b->nj = b->ju2 - b->ju1 + 1;
    b->nk = b->ku2 - b->ku1 + 1;
    b->fi = b->if2 - b->if1 + 1;
    b->fj = b->jf2 - b->jf1 + 1;
    b->fk = b->kf2 - b->kf1 + 1;
}

/* u-array element (i,j,k,n[,d]) with n,d 1-based */
#define UI(b, i, j, k, n) \
    (((i) - (b)->iu1) + (b)->ni * (((j) - (b)->ju1) + (b)->nj * (((k) - (b)->ku1) + (b)->nk * ((n) - 1))))
#define UI5(b, i, j, k, n, d) (UI(b, i, j, k, n) + (b)->ni * (b)->nj * (b)->nk * ORA_NVAR * ((d) - 1))
#define UC(b, i, j, k) (((i) - (b)->iu1) + (b)->ni * (((j) - (b)->ju1) + (b)->nj * ((k) - (b)->ku1)))
/* flux element (i,j,k,n,d) */
#define FI(b, i, j, k, n, d)                                                                    \
    (((i) - (b)->if1) + (b)->fi * (((j) - (b)->jf1) + (b)->fj * (((k) - (b)->kf1) +             \
     (b)->fk * (((n) - 1) + ORA_NVAR * ((d) - 1)))))

/* ctoprim, hydro/umuscl.f90:1739-1901 (hydro, no MHD/non-thermal energies) */
static void ctoprim(const kbounds_t *b, const double *uin, double *q, double *c, const double *gin,
                    double dt, const ora_params_t *p)
{
    double gamma = p->gamma, smallr = p->smallr, smallc = p->smallc;
    double smalle = (smallc * smallc) / gamma / (gamma - ONE);
    double dtxhalf = dt * HALF;
    for (int k = b->ku1; k <= b->ku2; k++)
        for (int j = b->ju1; j <= b->ju2; j++)
            for (int i = b->iu1; i <= b->iu2; i++) {
                double r = dmax(uin[UI(b, i, j, k, 1)], smallr);
                q[UI(b, i, j, k, 1)] = r;
                double oneoverrho = ONE / r;
                double vx = uin[UI(b, i, j, k, 2)] * oneoverrho;
                double vy = uin[UI(b, i, j, k, 3)] * oneoverrho;
                double vz = uin[UI(b, i, j, k, 4)] * oneoverrho;
                double ekin = HALF * r * (vx * vx);
                ekin = ekin + HALF * r * (vy * vy);
                ekin = ekin + HALF * r * (vz * vz);
                double eint = dmax(uin[UI(b, i, j, k, 5)] - ekin, smalle * r);
                double pr = (gamma - ONE) * eint;
                q[UI(b, i, j, k, 5)] = pr;
                double cc = gamma * pr;
                c[UC(b, i, j, k)] = sqrt(cc * oneoverrho);
                /* gravity predictor: gin(..,1:ndim); components beyond ndim untouched */
                vx = vx + gin[UI(b, i, j, k, 1)] * dtxhalf;
                if (b->ndim > 1) vy = vy + gin[UI(b, i, j, k, 2)] * dtxhalf;
                if (b->ndim > 2) vz = vz + gin[UI(b, i, j, k, 3)] * dtxhalf;
                q[UI(b, i, j, k, 2)] = vx;
                q[UI(b, i, j, k, 3)] = vy;
                q[UI(b, i, j, k, 4)] = vz;
            }
}

/* uslope, hydro/umuscl.f90:1906-2387.  1D: MinMod/MonCen via the MIN(st,2)
 * form (:1958-1966); 3D: MinMod (:2132-2170), MonCen (:2171-2206),
 * positivity-preserving (:2207-2270); then the normal-velocity bound (:2361-2385). */
static void uslope(const kbounds_t *b, const double *q, double *dq, const ora_params_t *p)
{
    int ilo = b->iu1 + 1 < 1 ? b->iu1 + 1 : 1, ihi = b->iu2 - 1 > 1 ? b->iu2 - 1 : 1;
    int jlo = b->ju1 + 1 < 1 ? b->ju1 + 1 : 1, jhi = b->ju2 - 1 > 1 ? b->ju2 - 1 : 1;
    int klo = b->ku1 + 1 < 1 ? b->ku1 + 1 : 1, khi = b->ku2 - 1 > 1 ? b->ku2 - 1 : 1;
    int st = p->slope_type;
    int ndim = b->ndim;
    if (st == 0) {
        memset(dq, 0, sizeof(double) * (size_t)b->ni * b->nj * b->nk * ORA_NVAR * ndim);
        return;
    }
    for (int n = 1; n <= ORA_NVAR; n++)
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    if (ndim == 1) {
                        double fac = (double)(st < 2 ? st : 2);
                        double dlft = fac * (q[UI(b, i, j, k, n)] - q[UI(b, i - 1, j, k, n)]);
                        double drgt = fac * (q[UI(b, i + 1, j, k, n)] - q[UI(b, i, j, k, n)]);
                        double dcen = HALF * (dlft + drgt) / fac;
                        double dsgn = copysign(ONE, dcen);
                        double slop = dmin(fabs(dlft), fabs(drgt));
                        double dlim = slop;
                        if ((dlft * drgt) <= ZERO) dlim = ZERO;
                        dq[UI5(b, i, j, k, n, 1)] = dsgn * dmin(dlim, fabs(dcen));
                        continue;
                    }
                    if (st == 1) {
                        for (int d = 1; d <= 3; d++) {
                            int di = d == 1, dj = d == 2, dk = d == 3;
                            double dlft = q[UI(b, i, j, k, n)] - q[UI(b, i - di, j - dj, k - dk, n)];
                            double drgt = q[UI(b, i + di, j + dj, k + dk, n)] - q[UI(b, i, j, k, n)];
                            double v;
                            if ((dlft * drgt) <= ZERO)
                                v = ZERO;
                            else if (dlft > 0)
                                v = dmin(dlft, drgt);
                            else
                                v = dmax(dlft, drgt);
                            dq[UI5(b, i, j, k, n, d)] = v;
                        }
                    } else if (st == 2) {
                        for (int d = 1; d <= 3; d++) {
                            int di = d == 1, dj = d == 2, dk = d == 3;
                            double dlft = 2.0 * (q[UI(b, i, j, k, n)] - q[UI(b, i - di, j - dj, k - dk, n)]);
                            double drgt = 2.0 * (q[UI(b, i + di, j + dj, k + dk, n)] - q[UI(b, i, j, k, n)]);
                            double dcen = HALF * (dlft + drgt) / 2.0;
                            double dsgn = copysign(ONE, dcen);
                            double slop = dmin(fabs(dlft), fabs(drgt));
                            double dlim = slop;
                            if ((dlft * drgt) <= ZERO) dlim = ZERO;
                            dq[UI5(b, i, j, k, n, d)] = dsgn * dmin(dlim, fabs(dcen));
                        }
                    } else if (st == 3) {
                        double qc = q[UI(b, i, j, k, n)];
                        double vmin = 0, vmax = 0;
                        int first = 1;
                        /* same 27 differences, min/max are order independent */
                        for (int kk = -1; kk <= 1; kk++)
                            for (int jj = -1; jj <= 1; jj++)
                                for (int ii = -1; ii <= 1; ii++) {
                                    double df = q[UI(b, i + ii, j + jj, k + kk, n)] - qc;
                                    if (first) {
                                        vmin = vmax = df;
                                        first = 0;
                                    } else {
                                        vmin = dmin(vmin, df);
                                        vmax = dmax(vmax, df);
                                    }
                                }
                        double dfx = HALF * (q[UI(b, i + 1, j, k, n)] - q[UI(b, i - 1, j, k, n)]);
                        double dfy = HALF * (q[UI(b, i, j + 1, k, n)] - q[UI(b, i, j - 1, k, n)]);
                        double dfz = HALF * (q[UI(b, i, j, k + 1, n)] - q[UI(b, i, j, k - 1, n)]);
                        double dff = HALF * (fabs(dfx) + fabs(dfy) + fabs(dfz));
                        double slop;
                        if (dff > ZERO)
                            slop = dmin(ONE, dmin(fabs(vmin), fabs(vmax)) / dff);
                        else
                            slop = ONE;
                        double dlim = slop;
                        dq[UI5(b, i, j, k, n, 1)] = dlim * dfx;
                        dq[UI5(b, i, j, k, n, 2)] = dlim * dfy;
                        dq[UI5(b, i, j, k, n, 3)] = dlim * dfz;
                    }
                }
    /* normal-velocity slope bound */
    for (int idim = 1; idim <= ndim; idim++) {
        int n = idim + 1;
        int di = idim == 1, dj = idim == 2, dk = idim == 3;
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    double dlft = q[UI(b, i - di, j - dj, k - dk, 1)] *
                                  fabs(q[UI(b, i, j, k, n)] - q[UI(b, i - di, j - dj, k - dk, n)]);
                    double drgt = q[UI(b, i + di, j + dj, k + dk, 1)] *
                                  fabs(q[UI(b, i + di, j + dj, k + dk, n)] - q[UI(b, i, j, k, n)]);
                    double dlim = TWO * dmin(dlft, drgt) / q[UI(b, i, j, k, 1)];
                    double d = dq[UI5(b, i, j, k, n, idim)];
                    dq[UI5(b, i, j, k, n, idim)] = fsign(dmin(fabs(d), dlim), d);
                }
    }
}

/* trace1d, hydro/umuscl.f90:297-481 and trace3d, hydro/umuscl.f90:860-1466 */
static void trace(const kbounds_t *b, const double *q, const double *dq, double *qm, double *qp,
                  double dx, double dt, const ora_params_t *p)
{
    double dtdx = dt / dx;
    double gamma = p->gamma, smallr = p->smallr, smallc = p->smallc;
    double smallp = smallr * (smallc * smallc);
    int ilo = b->iu1 + 1 < 1 ? b->iu1 + 1 : 1, ihi = b->iu2 - 1 > 1 ? b->iu2 - 1 : 1;
    int jlo = b->ju1 + 1 < 1 ? b->ju1 + 1 : 1, jhi = b->ju2 - 1 > 1 ? b->ju2 - 1 : 1;
    int klo = b->ku1 + 1 < 1 ? b->ku1 + 1 : 1, khi = b->ku2 - 1 > 1 ? b->ku2 - 1 : 1;
    int ndim = b->ndim;
    for (int k = klo; k <= khi; k++)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++) {
                double r = q[UI(b, i, j, k, 1)], u = q[UI(b, i, j, k, 2)], v = q[UI(b, i, j, k, 3)];
                double w = q[UI(b, i, j, k, 4)], pp = q[UI(b, i, j, k, 5)];
                double sr0, su0, sv0, sw0, sp0;
                double d[4][6] = {{0}}; /* d[dim][var], half slopes */
                for (int dd = 1; dd <= ndim; dd++)
                    for (int n = 1; n <= 5; n++) d[dd][n] = HALF * dq[UI5(b, i, j, k, n, dd)];
                if (ndim == 1) {
                    double drx = d[1][1], dux = d[1][2], dvx = d[1][3], dwx = d[1][4], dpx = d[1][5];
                    sr0 = -u * drx - (dux)*r;
                    sp0 = -u * dpx - (dux)*gamma * pp;
                    su0 = -u * dux - (dpx) / r;
                    sv0 = -u * dvx;
                    sw0 = -u * dwx;
                } else {
                    double drx = d[1][1], dux = d[1][2], dvx = d[1][3], dwx = d[1][4], dpx = d[1][5];
                    double dry = d[2][1], duy = d[2][2], dvy = d[2][3], dwy = d[2][4], dpy = d[2][5];
                    double drz = d[3][1], duz = d[3][2], dvz = d[3][3], dwz = d[3][4], dpz = d[3][5];
                    sr0 = -u * drx - v * dry - w * drz - (dux + dvy + dwz) * r;
                    su0 = -u * dux - v * duy - w * duz - (dpx) / r;
                    sv0 = -u * dvx - v * dvy - w * dvz - (dpy) / r;
                    sw0 = -u * dwx - v * dwy - w * dwz - (dpz) / r;
                    sp0 = -u * dpx - v * dpy - w * dpz - (dux + dvy + dwz) * gamma * pp;
                }
                r = r + sr0 * dtdx;
                u = u + su0 * dtdx;
                v = v + sv0 * dtdx;
                w = w + sw0 * dtdx;
                pp = pp + sp0 * dtdx;
                double ws[6] = {0, r, u, v, w, pp};
                for (int dd = 1; dd <= ndim; dd++) {
                    /* right state at left interface (qp), left state at right interface (qm) */
                    for (int n = 1; n <= 5; n++) qp[UI5(b, i, j, k, n, dd)] = ws[n] - d[dd][n];
                    if (qp[UI5(b, i, j, k, 1, dd)] < smallr) qp[UI5(b, i, j, k, 1, dd)] = q[UI(b, i, j, k, 1)];
                    if (qp[UI5(b, i, j, k, 5, dd)] < smallp) qp[UI5(b, i, j, k, 5, dd)] = q[UI(b, i, j, k, 5)];
                    for (int n = 1; n <= 5; n++) qm[UI5(b, i, j, k, n, dd)] = ws[n] + d[dd][n];
                    if (qm[UI5(b, i, j, k, 1, dd)] < smallr) qm[UI5(b, i, j, k, 1, dd)] = q[UI(b, i, j, k, 1)];
                    if (qm[UI5(b, i, j, k, 5, dd)] < smallp) qm[UI5(b, i, j, k, 5, dd)] = q[UI(b, i, j, k, 5)];
                }
            }
}

/* cmpflxm, hydro/umuscl.f90:1472-1607, plus the flux*dt/dx store of unsplit.
 * Face (i,j,k) in direction idim sits between cell (i,j,k)-e_idim (left, qm)
 * and cell (i,j,k) (right, qp).                                               */
static void cmpflxm(const kbounds_t *b, const double *qm, const double *qp, int idim,
                    int ilo, int ihi, int jlo, int jhi, int klo, int khi, double *flux,
                    double dx, double dt, const ora_params_t *p)
{
    int ln = idim + 1, lt1, lt2;
    if (idim == 1) { lt1 = 3; lt2 = 4; }
    else if (idim == 2) { lt1 = 2; lt2 = 4; }
    else { lt1 = 2; lt2 = 3; }
    int di = idim == 1, dj = idim == 2, dk = idim == 3;
    for (int k = klo; k <= khi; k++)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++) {
                double ql[5], qr[5], fg[6];
                ql[0] = qm[UI5(b, i - di, j - dj, k - dk, 1, idim)];
                ql[1] = qm[UI5(b, i - di, j - dj, k - dk, ln, idim)];
                ql[2] = qm[UI5(b, i - di, j - dj, k - dk, lt1, idim)];
                ql[3] = qm[UI5(b, i - di, j - dj, k - dk, lt2, idim)];
                ql[4] = qm[UI5(b, i - di, j - dj, k - dk, 5, idim)];
                qr[0] = qp[UI5(b, i, j, k, 1, idim)];
                qr[1] = qp[UI5(b, i, j, k, ln, idim)];
                qr[2] = qp[UI5(b, i, j, k, lt1, idim)];
                qr[3] = qp[UI5(b, i, j, k, lt2, idim)];
                qr[4] = qp[UI5(b, i, j, k, 5, idim)];
                if (p->riemann == ORA_LLF)
                    ora_riemann_llf(ql, qr, fg, p);
                else if (p->riemann == ORA_HLLC)
                    ora_riemann_hllc(ql, qr, fg, p);
                else
                    ora_riemann_hll(ql, qr, fg, p);
                double fx[6];
                fx[1] = fg[0];
                fx[ln] = fg[1];
                fx[lt1] = fg[2];
                fx[lt2] = fg[3];
                fx[5] = fg[4];
                for (int n = 1; n <= 5; n++) flux[FI(b, i, j, k, n, idim)] = fx[n] * dt / dx;
            }
}

/* cmpdivu + cmpdiff, hydro/umuscl.f90:2392-2540 (artificial diffusion, difmag>0) */
static void cmpdiff(const kbounds_t *b, const double *uin, const double *q, double *flux, double dt,
                    double dx, double difmag)
{
    int ndim = b->ndim;
    double *div = (double *)calloc((size_t)b->fi * b->fj * b->fk, sizeof(double));
#define DV(i, j, k) div[((i) - b->if1) + b->fi * (((j) - b->jf1) + b->fj * ((k) - b->kf1))]
#define Q(i, j, k, n) q[UI(b, i, j, k, n)]
#define U(i, j, k, n) uin[UI(b, i, j, k, n)]
    /* cmpdivu: fact = half**(ndim-1)/dx */
    double fact = (ndim == 1 ? 1.0 : (ndim == 2 ? HALF : HALF * HALF)) / dx;
    for (int k = b->kf1; k <= b->kf2; k++)
        for (int j = b->jf1; j <= b->jf2; j++)
            for (int i = b->if1; i <= b->if2; i++) {
                double ux = ZERO, vy = ZERO, wz = ZERO;
                ux = ux + fact * (Q(i, j, k, 2) - Q(i - 1, j, k, 2));
                if (ndim == 3) {
                    ux = ux + fact * (Q(i, j - 1, k, 2) - Q(i - 1, j - 1, k, 2));
                    vy = vy + fact * (Q(i, j, k, 3) - Q(i, j - 1, k, 3) + Q(i - 1, j, k, 3) - Q(i - 1, j - 1, k, 3));
                    ux = ux + fact * (Q(i, j, k - 1, 2) - Q(i - 1, j, k - 1, 2) + Q(i, j - 1, k - 1, 2) -
                                      Q(i - 1, j - 1, k - 1, 2));
                    vy = vy + fact * (Q(i, j, k - 1, 3) - Q(i, j - 1, k - 1, 3) + Q(i - 1, j, k - 1, 3) -
                                      Q(i - 1, j - 1, k - 1, 3));
                    wz = wz + fact * (Q(i, j, k, 4) - Q(i, j, k - 1, 4) + Q(i, j - 1, k, 4) - Q(i, j - 1, k - 1, 4) +
                                      Q(i - 1, j, k, 4) - Q(i - 1, j, k - 1, 4) + Q(i - 1, j - 1, k, 4) -
                                      Q(i - 1, j - 1, k - 1, 4));
                }
                DV(i, j, k) = ux + vy + wz;
            }
    /* cmpdiff: fact = half**(ndim-1) */
    double fd = ndim == 1 ? 1.0 : (ndim == 2 ? HALF : HALF * HALF);
    int khi = b->ku2 - 2 > b->kf1 ? b->ku2 - 2 : b->kf1;
    int jhi = b->ju2 - 2 > b->jf1 ? b->ju2 - 2 : b->jf1;
    for (int n = 1; n <= ORA_NVAR; n++) {
        for (int k = b->kf1; k <= khi; k++)
            for (int j = b->jf1; j <= jhi; j++)
                for (int i = b->if1; i <= b->if2; i++) {
                    double div1 = fd * DV(i, j, k);
                    if (ndim == 3) {
                        div1 = div1 + fd * DV(i, j + 1, k);
                        div1 = div1 + fd * (DV(i, j, k + 1) + DV(i, j + 1, k + 1));
                    }
                    div1 = difmag * dmin(ZERO, div1);
                    flux[FI(b, i, j, k, n, 1)] = flux[FI(b, i, j, k, n, 1)] + dt * div1 * (U(i, j, k, n) - U(i - 1, j, k, n));
                }
        if (ndim < 3) continue;
        for (int k = b->kf1; k <= khi; k++)
            for (int j = b->jf1; j <= b->jf2; j++)
                for (int i = b->iu1 + 2; i <= b->iu2 - 2; i++) {
                    double div1 = ZERO;
                    div1 = div1 + fd * (DV(i, j, k) + DV(i + 1, j, k));
                    div1 = div1 + fd * (DV(i, j, k + 1) + DV(i + 1, j, k + 1));
                    div1 = difmag * dmin(ZERO, div1);
                    flux[FI(b, i, j, k, n, 2)] = flux[FI(b, i, j, k, n, 2)] + dt * div1 * (U(i, j, k, n) - U(i, j - 1, k, n));
                }
        for (int k = b->kf1; k <= b->kf2; k++)
            for (int j = b->ju1 + 2; j <= b->ju2 - 2; j++)
                for (int i = b->iu1 + 2; i <= b->iu2 - 2; i++) {
                    double div1 = fd * (DV(i, j, k) + DV(i + 1, j, k) + DV(i, j + 1, k) + DV(i + 1, j + 1, k));
                    div1 = difmag * dmin(ZERO, div1);
                    flux[FI(b, i, j, k, n, 3)] = flux[FI(b, i, j, k, n, 3)] + dt * div1 * (U(i, j, k, n) - U(i, j, k - 1, n));
                }
    }
#undef DV
#undef Q
#undef U
    free(div);
}

/* unsplit, hydro/umuscl.f90:20-292 */
void ora_unsplit(int ndim, int nn, const double *uin, const double *gin, double *flux, double dx,
                 double dt, ora_params_t *p)
{
    kbounds_t b;
    kbounds_init(&b, ndim, nn);
    size_t nc = (size_t)b.ni * b.nj * b.nk;
    double *q = (double *)malloc(sizeof(double) * nc * ORA_NVAR);
    double *c = (double *)malloc(sizeof(double) * nc);
    double *dq = (double *)calloc(nc * ORA_NVAR * ndim, sizeof(double));
    double *qm = (double *)calloc(nc * ORA_NVAR * ndim, sizeof(double));
    double *qp = (double *)calloc(nc * ORA_NVAR * ndim, sizeof(double));

    int ilo = b.iu1 + 2 < 1 ? b.iu1 + 2 : 1, ihi = b.iu2 - 2 > 1 ? b.iu2 - 2 : 1;
    int jlo = b.ju1 + 2 < 1 ? b.ju1 + 2 : 1, jhi = b.ju2 - 2 > 1 ? b.ju2 - 2 : 1;
    int klo = b.ku1 + 2 < 1 ? b.ku1 + 2 : 1, khi = b.ku2 - 2 > 1 ? b.ku2 - 2 : 1;

    p->smax_cap = p->courant_factor * dx / dt;

    ctoprim(&b, uin, q, c, gin, dt, p);
    uslope(&b, q, dq, p);
    trace(&b, q, dq, qm, qp, dx, dt, p);

    cmpflxm(&b, qm, qp, 1, b.if1, b.if2, jlo, jhi, klo, khi, flux, dx, dt, p);
    if (ndim > 1) cmpflxm(&b, qm, qp, 2, ilo, ihi, b.jf1, b.jf2, klo, khi, flux, dx, dt, p);
    if (ndim > 2) cmpflxm(&b, qm, qp, 3, ilo, ihi, jlo, jhi, b.kf1, b.kf2, flux, dx, dt, p);

    if (p->difmag > 0.0) cmpdiff(&b, uin, q, flux, dt, dx, p->difmag);

    free(q);
    free(c);
    free(dq);
    free(qm);
    free(qp);
}

/* cmpdt, hydro/godunov_utils.f90:228-337 (hydro, no MHD: b2=0, bb=0) */
void ora_cmpdt(int ndim, double *uu, const double *gg, double dx, const ora_params_t *p, double *dt)
{
    double smallp = (p->smallc * p->smallc) / p->gamma;
    uu[0] = dmax(uu[0], p->smallr);
    for (int idim = 1; idim <= 3; idim++) uu[idim] = uu[idim] / uu[0];
    double b2 = ZERO;
    for (int idim = 1; idim <= 3; idim++) uu[4] = uu[4] - HALF * uu[0] * (uu[idim] * uu[idim]);
    uu[4] = dmax((p->gamma - ONE) * uu[4], uu[0] * smallp);
    double a2 = p->gamma * uu[4];
    a2 = a2 / uu[0];
    double ctot = ZERO;
    for (int idim = 1; idim <= ndim; idim++) {
        double bb = 0.0;
        double c2 = HALF * (b2 / uu[0] + a2);
        double cfast2 = c2 + sqrt(c2 * c2 - a2 * (bb * bb) / uu[0]);
        ctot = ctot + fabs(uu[idim]) + sqrt(cfast2);
    }
    double grav = ZERO;
    for (int idim = 0; idim < ndim; idim++) grav = grav + fabs(gg[idim]);
    grav = grav * dx / (ctot * ctot);
    grav = dmax(grav, 0.0001);
    double d = p->courant_factor * dx / p->smallc;
    double dtcell = TWO * p->courant_factor * dx / (ctot * (ONE + sqrt(ONE + TWO * p->courant_factor * grav)));
    *dt = dmin(d, dtcell);
}

/* ------------------------------------------------------------------------ */
/* interpol_hydro, hydro/interpol_hydro.f90:5-113 (+ limiters 610-752)       */
/* ------------------------------------------------------------------------ */
static void cell_xc(int ndim, int ind /*1-based*/, double *xc)
{
    int iz = (ind - 1) / 4;
    int iy = (ind - 1 - 4 * iz) / 2;
    int ix = (ind - 1 - 2 * iy - 4 * iz);
    xc[0] = ((double)ix - 0.5);
    xc[1] = ndim > 1 ? ((double)iy - 0.5) : 0.0;
    xc[2] = ndim > 2 ? ((double)iz - 0.5) : 0.0;
}

void ora_interpol_hydro(int ndim, double *u1, double *u2, int interpol_var, int interpol_type, double smallr)
{
    int twondim = 2 * ndim, twotondim = 1 << ndim, n1 = twondim + 1;
#define U1(j, ivar) u1[(j) + n1 * ((ivar) - 1)]
#define U2(ind, ivar) u2[((ind) - 1) + twotondim * ((ivar) - 1)]
    if (interpol_var == 1) {
        for (int j = 0; j <= twondim; j++) {
            double ekin = 0.0;
            for (int idim = 1; idim <= 3; idim++)
                ekin = ekin + 0.5 * (U1(j, idim + 1) * U1(j, idim + 1)) / dmax(U1(j, 1), smallr);
            double erad = 0.0;
            U1(j, 5) = U1(j, 5) - ekin - erad;
        }
    }
    for (int ivar = 1; ivar <= ORA_NVAR; ivar++) {
        double a[7], w[3] = {0, 0, 0};
        for (int j = 0; j <= twondim; j++) a[j] = U1(j, ivar);
        if (interpol_type == 1) {
            for (int idim = 1; idim <= ndim; idim++) {
                double diff_left = 0.5 * (a[2 * idim] - a[0]);
                double diff_right = 0.5 * (a[0] - a[2 * idim - 1]);
                double mm;
                if (diff_left * diff_right <= 0.0)
                    mm = 0.0;
                else
                    mm = dmin(fabs(diff_left), fabs(diff_right)) * diff_left / fabs(diff_left);
                w[idim - 1] = mm;
            }
        } else if (interpol_type == 2) {
            for (int idim = 1; idim <= ndim; idim++) w[idim - 1] = 0.25 * (a[2 * idim] - a[2 * idim - 1]);
            double ac[8] = {0};
            for (int ind = 1; ind <= twotondim; ind++) ac[ind - 1] = a[0];
            for (int idim = 1; idim <= ndim; idim++)
                for (int ind = 1; ind <= twotondim; ind++) {
                    double xc[3];
                    cell_xc(ndim, ind, xc);
                    double corner = ac[ind - 1] + 2.0 * w[idim - 1] * xc[idim - 1];
                    ac[ind - 1] = corner;
                }
            double corner = ac[0];
            for (int j = 2; j <= twotondim; j++) corner = dmax(corner, ac[j - 1]);
            double kernel = a[1];
            for (int j = 2; j <= twondim; j++) kernel = dmax(kernel, a[j]);
            double diff_kernel = a[0] - kernel, diff_corner = a[0] - corner;
            double max_limiter = 0.0;
            if (diff_kernel * diff_corner > 0.0) max_limiter = dmin(1.0, diff_kernel / diff_corner);
            corner = ac[0];
            for (int j = 2; j <= twotondim; j++) corner = dmin(corner, ac[j - 1]);
            kernel = a[1];
            for (int j = 2; j <= twondim; j++) kernel = dmin(kernel, a[j]);
            diff_kernel = a[0] - kernel;
            diff_corner = a[0] - corner;
            double min_limiter = 0.0;
            if (diff_kernel * diff_corner > 0.0) min_limiter = dmin(1.0, diff_kernel / diff_corner);
            double limiter = dmin(min_limiter, max_limiter);
            for (int idim = 1; idim <= ndim; idim++) w[idim - 1] = w[idim - 1] * limiter;
        } else if (interpol_type == 3) {
            for (int idim = 1; idim <= ndim; idim++) w[idim - 1] = 0.25 * (a[2 * idim] - a[2 * idim - 1]);
        }
        for (int ind = 1; ind <= twotondim; ind++) {
            double xc[3];
            cell_xc(ndim, ind, xc);
            U2(ind, ivar) = a[0];
            for (int idim = 1; idim <= ndim; idim++) U2(ind, ivar) = U2(ind, ivar) + w[idim - 1] * xc[idim - 1];
        }
    }
    if (interpol_var == 1) {
        for (int ind = 1; ind <= twotondim; ind++) {
            double ekin = 0.0;
            for (int idim = 1; idim <= 3; idim++)
                ekin = ekin + 0.5 * (U2(ind, idim + 1) * U2(ind, idim + 1)) / dmax(U2(ind, 1), smallr);
            double erad = 0.0;
            U2(ind, 5) = U2(ind, 5) + ekin + erad;
        }
    }
#undef U1
#undef U2
}

/* ------------------------------------------------------------------------ */
/* Mesh lookup: (lev, ckey) -> oct, the role of hash_getp (amr/hash.f90)     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t *keys; /* packed key, -1 = empty */
    int32_t *vals;
    int64_t cap;
} omap_t;

static int64_t pack_key(int lev, int64_t x, int64_t y, int64_t z)
{
    /* 21 bits per coordinate (offset so that -1 fits), 6 bits of level */
    return ((int64_t)lev << 58) | (((x + 1) & 0x1FFFFF) << 37) | (((y + 1) & 0x1FFFFF) << 16) |
           ((z + 1) & 0xFFFF);
}

static uint64_t mix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

static void omap_init(omap_t *m, int64_t n)
{
    int64_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    m->cap = cap;
    m->keys = (int64_t *)malloc(sizeof(int64_t) * cap);
    m->vals = (int32_t *)malloc(sizeof(int32_t) * cap);
    for (int64_t i = 0; i < cap; i++) m->keys[i] = -1;
}

static void omap_put(omap_t *m, int64_t key, int32_t val)
{
    uint64_t h = mix64((uint64_t)key) & (uint64_t)(m->cap - 1);
    while (m->keys[h] != -1 && m->keys[h] != key) h = (h + 1) & (uint64_t)(m->cap - 1);
    m->keys[h] = key;
    m->vals[h] = val;
}

static int32_t omap_get(const omap_t *m, int64_t key)
{
    uint64_t h = mix64((uint64_t)key) & (uint64_t)(m->cap - 1);
    while (m->keys[h] != -1) {
        if (m->keys[h] == key) return m->vals[h];
        h = (h + 1) & (uint64_t)(m->cap - 1);
    }
    return 0;
}

static void omap_free(omap_t *m)
{
    free(m->keys);
    free(m->vals);
}

static void omap_build(omap_t *h, const ora_mesh_t *m)
{
    omap_init(h, m->ngrid);
    for (int i = 0; i < m->ngrid; i++) {
        const int32_t *ck = m->ckey + 3 * i;
        omap_put(h, pack_key(m->lev[i], ck[0], m->ndim > 1 ? ck[1] : 0, m->ndim > 2 ? ck[2] : 0), i + 1);
    }
}

/* Ghost oct storage created on the fly for domain boundaries, as get_grid
 * does for keys outside the computational domain (amr/nbors_utils.f90:356-420,
 * amr/boundaries.f90:39-165).                                                */
typedef struct {
    double u[8 * ORA_NVAR];
    uint8_t refined[8];
} ghost_t;

/* Index of cell ind (1-based) of an oct: u(ind,ivar,igrid) */
#define UOCT(u, nvar, igrid, ivar, ind, ttd) (u)[((size_t)((igrid) - 1) * (nvar) + ((ivar) - 1)) * (ttd) + ((ind) - 1)]

typedef struct {
    const ora_mesh_t *m;
    const omap_t *h;
    int ttd;
} lookup_t;

static int32_t box_max(int lev) { return lev >= 1 ? (1 << (lev - 1)) : 1; }

/* box_ckey_min/max(d, lev) (amr/init_amr.f90:250-280) */
static int64_t box_lo(const ora_mesh_t *m, int lev, int d)
{
    if (m->box_lev <= 0) return 0;
    return lev >= m->box_lev ? (int64_t)m->box_ckey_min[d] << (lev - m->box_lev)
                             : (int64_t)m->box_ckey_min[d] >> (m->box_lev - lev);
}
static int64_t box_hi(const ora_mesh_t *m, int lev, int d)
{
    if (m->box_lev <= 0) return box_max(lev);
    return lev >= m->box_lev ? (int64_t)m->box_ckey_max[d] << (lev - m->box_lev)
                             : (int64_t)m->box_ckey_max[d] >> (m->box_lev - lev);
}

/* get_grid for a (possibly out-of-domain) key: returns an oct index (>0),
 * -1 when a boundary ghost was built into *g, or 0 when missing.             */
static int get_grid(const lookup_t *L, int lev, int64_t *key, ghost_t *g)
{
    const ora_mesh_t *m = L->m;
    int ndim = m->ndim;
    int32_t c = omap_get(L->h, pack_key(lev, key[0], ndim > 1 ? key[1] : 0, ndim > 2 ? key[2] : 0));
    if (c > 0) return c;
    int in_domain = 1, bdim = -1, side = 0;
    for (int d = 0; d < ndim; d++) {
        if (key[d] < box_lo(m, lev, d)) { in_domain = 0; if (bdim < 0) { bdim = d; side = 0; } }
        else if (key[d] >= box_hi(m, lev, d)) { in_domain = 0; if (bdim < 0) { bdim = d; side = 1; } }
    }
    if (in_domain) return 0;
    /* boundary ghost: reference oct one step back into the domain */
    int64_t kr[3] = {key[0], ndim > 1 ? key[1] : 0, ndim > 2 ? key[2] : 0};
    int shift = side == 0 ? +1 : -1;
    kr[bdim] += shift;
    int32_t cref = omap_get(L->h, pack_key(lev, kr[0], kr[1], kr[2]));
    if (cref <= 0) return 0;
    int type = m->bound_type[bdim][side];
    int dir = bdim + 1;
    /* tables of amr/boundaries.f90:55-80 (1-based cell indices) */
    static const int ref_right[3][8] = {{0, 1, 0, 3, 0, 5, 0, 7}, {0, 0, 1, 2, 0, 0, 5, 6}, {0, 0, 0, 0, 1, 2, 3, 4}};
    static const int ref_left[3][8] = {{2, 0, 4, 0, 6, 0, 8, 0}, {3, 4, 0, 0, 7, 8, 0, 0}, {5, 6, 7, 8, 0, 0, 0, 0}};
    static const int ind1[3][8] = {{2, 1, 4, 3, 6, 5, 8, 7}, {3, 4, 1, 2, 7, 8, 5, 6}, {5, 6, 7, 8, 1, 2, 3, 4}};
    static const int ind2_right[3][8] = {{1, 1, 3, 3, 5, 5, 7, 7}, {1, 2, 1, 2, 5, 6, 5, 6}, {1, 2, 3, 4, 1, 2, 3, 4}};
    static const int ind2_left[3][8] = {{2, 2, 4, 4, 6, 6, 8, 8}, {3, 4, 3, 4, 7, 8, 7, 8}, {5, 6, 7, 8, 5, 6, 7, 8}};
    int ttd = L->ttd, nvar = m->nvar;
    for (int ind = 1; ind <= ttd; ind++) {
        int r = shift == +1 ? ref_right[dir - 1][ind - 1] : ref_left[dir - 1][ind - 1];
        g->refined[ind - 1] = r > 0 ? m->refined[(size_t)(cref - 1) * 8 + (r - 1)] : 0;
    }
    for (int ivar = 1; ivar <= nvar; ivar++) {
        double reverse = (ivar == 1 + dir) ? -1.0 : 1.0;
        for (int ind = 1; ind <= ttd; ind++) {
            int src;
            if (type == 1)
                src = ind1[dir - 1][ind - 1];
            else
                src = shift == +1 ? ind2_right[dir - 1][ind - 1] : ind2_left[dir - 1][ind - 1];
            double v = UOCT(m->uold, nvar, cref, ivar, src, ttd);
            g->u[(ivar - 1) * ttd + (ind - 1)] = type == 1 ? v * reverse : v;
        }
    }
    return -1;
}

/* Periodic wrap as in godfine1 (hydro/godunov_fine.f90:463-468) */
static void wrap_key(const ora_mesh_t *m, int lev, int64_t *key)
{
    for (int d = 0; d < m->ndim; d++)
        if (m->periodic[d]) {
            if (key[d] < box_lo(m, lev, d)) key[d] = box_hi(m, lev, d) - 1;
            if (key[d] >= box_hi(m, lev, d)) key[d] = box_lo(m, lev, d);
        }
}

/* Parent key of a cell key: key/2 as the reference divides (Fortran integer
 * division, amr/nbors_utils.f90:180,200,241) for the non-negative keys of its
 * meshes, whose non-periodic boxes start at box_ckey_min >= 1 with boundary
 * octs below (amr/init_amr.f90:255-283).  Floor division extends this to boxes
 * that start at key 0, where a neighbour key -1 belongs to the boundary oct
 * -1 (cell bit 1), exactly as in the reference's mesh shifted by an even
 * number of keys; truncation would give parent 0 and cell bit -1. */
static int64_t fdiv2(int64_t a) { return a >= 0 ? a / 2 : -((1 - a) / 2); }

/* ------------------------------------------------------------------------ */
/* godfine1, hydro/godunov_fine.f90:244-1342 (nsubgrid = 1)                  */
/* ------------------------------------------------------------------------ */
static int godfine1(const lookup_t *L, int ind_grid, int ilevel, double dt, ora_params_t *p, double *uloc,
                    double *gloc, uint8_t *okloc, double *flux, const kbounds_t *b)
{
    const ora_mesh_t *m = L->m;
    int ndim = m->ndim, nvar = m->nvar, ttd = L->ttd;
    double dx = m->boxlen / (double)(1LL << ilevel);
    double oneontwotondim = 1.0 / (double)ttd;
    /* io1=0, io2=2 for nn=2 */
    int childloc[27], gridloc[27], cellloc[27];
    int i1max = 2;
    int j1min = ndim > 1 ? 0 : 1, j1max = ndim > 1 ? 2 : 1;
    int k1min = ndim > 2 ? 0 : 1, k1max = ndim > 2 ? 2 : 1;
#define OI(i1, j1, k1) ((i1) + 3 * (((j1)-j1min) + 3 * ((k1)-k1min)))
    int64_t ckey_corner[3] = {0, 0, 0};
    for (int d = 0; d < ndim; d++) ckey_corner[d] = m->ckey[(size_t)(ind_grid - 1) * 3 + d];
    for (int k1 = k1min; k1 <= k1max; k1++)
        for (int j1 = j1min; j1 <= j1max; j1++)
            for (int i1 = 0; i1 <= i1max; i1++) {
                int okx = i1 > 0 && i1 < i1max;
                int oky = ndim < 2 || (j1 > j1min && j1 < j1max);
                int okz = ndim < 3 || (k1 > k1min && k1 < k1max);
                int ichild = 0, igrid = 0, icell = 0;
                ghost_t gh;
                double u2[8 * ORA_NVAR];
                int use_ghost = 0, use_interp = 0;
                if (okx && oky && okz) {
                    ichild = ind_grid;
                } else {
                    int64_t key[3] = {ckey_corner[0] + i1 - 1, ndim > 1 ? ckey_corner[1] + j1 - 1 : 0,
                                      ndim > 2 ? ckey_corner[2] + k1 - 1 : 0};
                    wrap_key(m, ilevel, key);
                    int c = get_grid(L, ilevel, key, &gh);
                    if (c > 0)
                        ichild = c;
                    else if (c == -1)
                        use_ghost = 1;
                    if (c == 0) {
                        /* get_parent_cell (amr/nbors_utils.f90:222-249) */
                        int64_t kf[3] = {0, 0, 0};
                        int ii[3] = {0, 0, 0};
                        for (int d = 0; d < ndim; d++) {
                            kf[d] = fdiv2(key[d]);
                            ii[d] = (int)(key[d] - 2 * kf[d]);
                        }
                        icell = 1;
                        for (int d = 0; d < ndim; d++) icell += (1 << d) * ii[d];
                        /* a stencil oct outside the box whose reference oct is
                         * missing has its parent among the coarse level's
                         * boundary octs (get_grid builds one in the cache,
                         * amr/nbors_utils.f90:580-640): igrid = -1, values gp */
                        ghost_t gp;
                        igrid = get_grid(L, ilevel - 1, kf, &gp);
                        if (igrid == 0) return 1; /* GODUNOV: parent_cell should exist */
                        /* get_twondim_nbor_parent_cell (amr/nbors_utils.f90:149-217) */
                        double u1[7 * ORA_NVAR];
                        int n1 = 2 * ndim + 1;
                        for (int ivar = 1; ivar <= nvar; ivar++)
                            u1[0 + n1 * (ivar - 1)] = igrid > 0 ? UOCT(m->uold, nvar, igrid, ivar, icell, ttd)
                                                                : gp.u[(ivar - 1) * ttd + (icell - 1)];
                        for (int inbor = 1; inbor <= 2 * ndim; inbor++) {
                            int d = (inbor - 1) / 2;
                            int64_t kn[3] = {key[0], key[1], key[2]};
                            kn[d] += (inbor % 2 == 1) ? -1 : +1;
                            wrap_key(m, ilevel, kn);
                            int64_t knf[3] = {0, 0, 0};
                            int jj[3] = {0, 0, 0};
                            for (int e = 0; e < ndim; e++) {
                                knf[e] = fdiv2(kn[e]);
                                jj[e] = (int)(kn[e] - 2 * knf[e]);
                            }
                            int ind = 1;
                            for (int e = 0; e < ndim; e++) ind += (1 << e) * jj[e];
                            ghost_t gq;
                            int gn = get_grid(L, ilevel - 1, knf, &gq);
                            if (gn == 0) return 2; /* GODUNOV: parent neighbors should exist */
                            for (int ivar = 1; ivar <= nvar; ivar++)
                                u1[inbor + n1 * (ivar - 1)] = gn > 0 ? UOCT(m->uold, nvar, gn, ivar, ind, ttd)
                                                                     : gq.u[(ivar - 1) * ttd + (ind - 1)];
                        }
                        ora_interpol_hydro(ndim, u1, u2, m->interpol_var, m->interpol_type, p->smallr);
                        use_interp = 1;
                    }
                }
                /* a boundary ghost oct is an oct of the cache in the reference
                 * (childloc > 0): no reflux through its faces */
                childloc[OI(i1, j1, k1)] = use_ghost ? -1 : ichild;
                gridloc[OI(i1, j1, k1)] = igrid;
                cellloc[OI(i1, j1, k1)] = icell;
                /* 2^ndim cells of this oct into the stencil */
                for (int k2 = 0; k2 <= (ndim > 2 ? 1 : 0); k2++)
                    for (int j2 = 0; j2 <= (ndim > 1 ? 1 : 0); j2++)
                        for (int i2 = 0; i2 <= 1; i2++) {
                            int ind_son = 1 + i2 + 2 * j2 + 4 * k2;
                            int i3 = 1 + 2 * (i1 - 1) + i2;
                            int j3 = ndim > 1 ? 1 + 2 * (j1 - 1) + j2 : 1;
                            int k3 = ndim > 2 ? 1 + 2 * (k1 - 1) + k2 : 1;
                            for (int ivar = 1; ivar <= nvar; ivar++) {
                                double v;
                                if (ichild > 0)
                                    v = UOCT(m->uold, nvar, ichild, ivar, ind_son, ttd);
                                else if (use_ghost)
                                    v = gh.u[(ivar - 1) * ttd + (ind_son - 1)];
                                else
                                    v = u2[(ivar - 1) * ttd + (ind_son - 1)];
                                uloc[UI(b, i3, j3, k3, ivar)] = v;
                            }
                            for (int idim = 1; idim <= ndim; idim++)
                                gloc[UI(b, i3, j3, k3, idim)] = m->constant_gravity[idim - 1];
                            uint8_t ok;
                            if (ichild > 0)
                                ok = m->refined[(size_t)(ichild - 1) * 8 + (ind_son - 1)];
                            else if (use_ghost)
                                ok = gh.refined[ind_son - 1];
                            else
                                ok = 0;
                            (void)use_interp;
                            okloc[UC(b, i3, j3, k3)] = ok;
                        }
                /* a boundary ghost oct plays the role of an existing child */
                if (use_ghost) childloc[OI(i1, j1, k1)] = -1;
            }

    for (size_t i = 0; i < (size_t)b->fi * b->fj * b->fk * ORA_NVAR * 3; i++) flux[i] = 0.0;
    ora_unsplit(ndim, 2, uloc, gloc, flux, dx, dt, p);

    /* reset flux at refined interfaces (:669-685) */
    int i3min = b->iu1 + 2, i3max = b->iu2 - 2;
    int j3min = b->ju1 + 2, j3max = b->ju2 - 2;
    int k3min = b->ku1 + 2, k3max = b->ku2 - 2;
    if (ndim < 2) { j3min = 1; j3max = 1; }
    if (ndim < 3) { k3min = 1; k3max = 1; }
    for (int idim = 1; idim <= ndim; idim++) {
        int i0 = idim == 1, j0 = idim == 2, k0 = idim == 3;
        for (int k3 = k3min; k3 <= k3max + k0; k3++)
            for (int j3 = j3min; j3 <= j3max + j0; j3++)
                for (int i3 = i3min; i3 <= i3max + i0; i3++)
                    if (okloc[UC(b, i3 - i0, j3 - j0, k3 - k0)] || okloc[UC(b, i3, j3, k3)])
                        for (int ivar = 1; ivar <= ORA_NVAR; ivar++) flux[FI(b, i3, j3, k3, ivar, idim)] = 0.0;
    }

    /* conservative update at level ilevel (:737-816) */
    for (int idim = 1; idim <= ndim; idim++) {
        int i0 = idim == 1, j0 = idim == 2, k0 = idim == 3;
        int ichild = ind_grid;
        for (int k2 = 0; k2 <= (ndim > 2 ? 1 : 0); k2++)
            for (int j2 = 0; j2 <= (ndim > 1 ? 1 : 0); j2++)
                for (int i2 = 0; i2 <= 1; i2++) {
                    int ind_son = 1 + i2 + 2 * j2 + 4 * k2;
                    int i3 = 1 + i2, j3 = ndim > 1 ? 1 + j2 : 1, k3 = ndim > 2 ? 1 + k2 : 1;
                    for (int ivar = 1; ivar <= 5; ivar++) {
                        double *un = &UOCT(m->unew, nvar, ichild, ivar, ind_son, ttd);
                        *un = *un + (flux[FI(b, i3, j3, k3, ivar, idim)] - flux[FI(b, i3 + i0, j3 + j0, k3 + k0, ivar, idim)]);
                    }
                }
    }

    /* conservative update at level ilevel-1 (:894-1052) */
    if (ilevel > m->levelmin) {
        for (int idim = 1; idim <= ndim; idim++) {
            int i0 = idim == 1, j0 = idim == 2, k0 = idim == 3;
            for (int side = 0; side < 2; side++) {
                int i1 = 1, j1 = 1, k1 = 1;
                if (idim == 1) i1 = side == 0 ? 0 : i1max;
                if (idim == 2) j1 = side == 0 ? j1min : j1max;
                if (idim == 3) k1 = side == 0 ? k1min : k1max;
                if (ndim < 2) j1 = 1;
                if (ndim < 3) k1 = 1;
                int ichild = childloc[OI(i1, j1, k1)];
                if (ichild != 0) continue;
                int igrid = gridloc[OI(i1, j1, k1)], icell = cellloc[OI(i1, j1, k1)];
                /* parent cell in a boundary oct of the coarse level: a cache oct
                 * that is never flushed back, so the reflux is dropped */
                if (igrid < 0) continue;
                int k2lo = 0, k2hi = ndim > 2 ? 1 : 0, j2lo = 0, j2hi = ndim > 1 ? 1 : 0, i2lo = 0, i2hi = 1;
                if (side == 0) { i2hi -= i0; j2hi -= j0; k2hi -= k0; }
                else { i2lo += i0; j2lo += j0; k2lo += k0; }
                for (int k2 = k2lo; k2 <= k2hi; k2++)
                    for (int j2 = j2lo; j2 <= j2hi; j2++)
                        for (int i2 = i2lo; i2 <= i2hi; i2++) {
                            int sgn = side == 0 ? +1 : -1;
                            int i3 = 1 + 2 * (i1 + sgn * i0 - 1) + i2;
                            int j3 = ndim > 1 ? 1 + 2 * (j1 + sgn * j0 - 1) + j2 : 1;
                            int k3 = ndim > 2 ? 1 + 2 * (k1 + sgn * k0 - 1) + k2 : 1;
                            for (int ivar = 1; ivar <= 5; ivar++) {
                                double *un = &UOCT(m->unew, nvar, igrid, ivar, icell, ttd);
                                if (side == 0)
                                    *un = *un - flux[FI(b, i3, j3, k3, ivar, idim)] * oneontwotondim;
                                else
                                    *un = *un + flux[FI(b, i3 + i0, j3 + j0, k3 + k0, ivar, idim)] * oneontwotondim;
                            }
                        }
            }
        }
    }
#undef OI
    return 0;
}

static void level_range(const ora_mesh_t *m, int ilevel, int *head, int *noct)
{
    *head = m->head[ilevel - 1];
    *noct = m->noct[ilevel - 1];
}

/* godunov_fine, hydro/godunov_fine.f90:43-79 */
int ora_godunov_fine(const ora_mesh_t *m, int ilevel, double dt, const ora_params_t *pin)
{
    ora_params_t p = *pin;
    omap_t h;
    omap_build(&h, m);
    lookup_t L = {m, &h, 1 << m->ndim};
    kbounds_t b;
    kbounds_init(&b, m->ndim, 2);
    size_t nc = (size_t)b.ni * b.nj * b.nk;
    double *uloc = (double *)calloc(nc * ORA_NVAR, sizeof(double));
    double *gloc = (double *)calloc(nc * 3, sizeof(double));
    uint8_t *okloc = (uint8_t *)calloc(nc, 1);
    double *flux = (double *)calloc((size_t)b.fi * b.fj * b.fk * ORA_NVAR * 3, sizeof(double));
    int head, noct, rc = 0;
    level_range(m, ilevel, &head, &noct);
    /* The reflux into the parent level follows the oct order of the level, as
     * the serial loop of godunov_fine does.                                   */
    for (int igrid = head; igrid < head + noct && rc == 0; igrid++)
        rc = godfine1(&L, igrid, ilevel, dt, &p, uloc, gloc, okloc, flux, &b);
    free(uloc);
    free(gloc);
    free(okloc);
    free(flux);
    omap_free(&h);
    return rc;
}

/* set_unew / set_uold, hydro/godunov_fine.f90:114-148,183-214 */
void ora_set_unew(const ora_mesh_t *m, int ilevel)
{
    int head, noct;
    level_range(m, ilevel, &head, &noct);
    size_t per = (size_t)m->nvar << m->ndim;
    memcpy(m->unew + (size_t)(head - 1) * per, m->uold + (size_t)(head - 1) * per, sizeof(double) * per * noct);
}

void ora_set_uold(const ora_mesh_t *m, int ilevel)
{
    int head, noct;
    level_range(m, ilevel, &head, &noct);
    size_t per = (size_t)m->nvar << m->ndim;
    memcpy(m->uold + (size_t)(head - 1) * per, m->unew + (size_t)(head - 1) * per, sizeof(double) * per * noct);
}

/* hydro_refine, hydro/godunov_utils.f90:6-227, density and pressure criteria
 * (no MHD, no non-thermal energies, no passive scalars), in its expression
 * order; ug/um/ud are modified in place as there. */
static int hydro_refine_dp(double gamma, double smallr, double err_grad_d, double err_grad_p, double floor_d,
                           double floor_p, double *ug, double *um, double *ud, int ok)
{
    ug[0] = ug[0] > smallr ? ug[0] : smallr;
    um[0] = um[0] > smallr ? um[0] : smallr;
    ud[0] = ud[0] > smallr ? ud[0] : smallr;
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
    /* (gamma-1)*(E - ekin - emag - erad) with emag = erad = 0 */
    double pg = (gamma - 1.0) * (ug[4] - eking - 0.0 - 0.0);
    double pm = (gamma - 1.0) * (um[4] - ekinm - 0.0 - 0.0);
    double pd = (gamma - 1.0) * (ud[4] - ekind - 0.0 - 0.0);
    if (err_grad_d >= 0.0) {
        double dg = ug[0], dm = um[0], dd = ud[0];
        double a = fabs((dd - dm) / (dd + dm + floor_d)), b = fabs((dm - dg) / (dm + dg + floor_d));
        double error = 2.0 * (a > b ? a : b);
        ok = ok || error > err_grad_d;
    }
    if (err_grad_p >= 0.0) {
        double a = fabs((pd - pm) / (pd + pm + floor_p)), b = fabs((pm - pg) / (pm + pg + floor_p));
        double error = 2.0 * (a > b ? a : b);
        ok = ok || error > err_grad_p;
    }
    return ok;
}

/* hydro_flag, hydro/hydro_flag.f90:7-159 (gradient criteria) */
int ora_hydro_flag(const ora_mesh_t *m, int ilevel, double gamma, double smallr, double err_grad_d,
                   double err_grad_p, double floor_d, double floor_p, int32_t *flag1)
{
    if (err_grad_d == -1.0 && err_grad_p == -1.0) return 0;
    omap_t h;
    omap_build(&h, m);
    const int ndim = m->ndim, ttd = 1 << ndim, nvar = m->nvar;
    lookup_t L = {m, &h, ttd};
    int head, noct;
    level_range(m, ilevel, &head, &noct);
    int nflag = 0, rc = 0;
    for (int igrid = head; igrid < head + noct && !rc; igrid++) {
        for (int ind = 1; ind <= ttd; ind++) {
            /* cell key at level ilevel+1 (iii of hydro_flag: bits of ind-1) */
            int64_t hk[3] = {0, 0, 0};
            for (int d = 0; d < ndim; d++)
                hk[d] = 2 * (int64_t)m->ckey[(size_t)(igrid - 1) * 3 + d] + (((ind - 1) >> d) & 1);
            int ok = 0;
            /* the 2 ndim neighbour cells (shift: -x +x -y +y -z +z): the cell
             * of level ilevel that holds it, else the one of level ilevel-1 */
            int gn[6];
            double un[6][5];
            for (int inb = 0; inb < 2 * ndim && !rc; inb++) {
                int64_t kn[3] = {hk[0], hk[1], hk[2]};
                kn[inb / 2] += (inb % 2 == 0) ? -1 : 1;
                wrap_key(m, ilevel + 1, kn);
                int64_t kp[3] = {0, 0, 0};
                int icell = 1;
                for (int d = 0; d < ndim; d++) {
                    kp[d] = fdiv2(kn[d]);
                    icell += (1 << d) * (int)(kn[d] - 2 * kp[d]);
                }
                ghost_t g;
                int c = get_grid(&L, ilevel, kp, &g);
                if (c == 0) {
                    int64_t kq[3] = {0, 0, 0};
                    icell = 1;
                    for (int d = 0; d < ndim; d++) {
                        int64_t k2 = fdiv2(kn[d]);
                        kq[d] = fdiv2(k2);
                        icell += (1 << d) * (int)(k2 - 2 * kq[d]);
                    }
                    c = get_grid(&L, ilevel - 1, kq, &g);
                    if (c == 0) {
                        rc = 1;
                        break;
                    }
                }
                gn[inb] = c;
                for (int v = 1; v <= nvar; v++)
                    un[inb][v - 1] = c > 0 ? UOCT(m->uold, nvar, c, v, icell, ttd) : g.u[(v - 1) * ttd + (icell - 1)];
            }
            if (rc) break;
            (void)gn;
            for (int d = 0; d < ndim; d++) {
                double ug[5], um[5], ud[5];
                for (int v = 1; v <= nvar; v++) {
                    ug[v - 1] = un[2 * d][v - 1];
                    um[v - 1] = UOCT(m->uold, nvar, igrid, v, ind, ttd);
                    ud[v - 1] = un[2 * d + 1][v - 1];
                }
                ok = hydro_refine_dp(gamma, smallr, err_grad_d, err_grad_p, floor_d, floor_p, ug, um, ud, ok);
            }
            int32_t *f = &flag1[(size_t)(igrid - 1) * ttd + (ind - 1)];
            if (*f == 0 && ok) nflag++;
            if (ok) *f = 1;
        }
    }
    omap_free(&h);
    return rc ? -1 : nflag;
}

/* upload_fine, hydro/upload.f90:74-283 */
/* get_grid of one key for tests: > 0 an oct of the mesh, -1 a boundary ghost
 * (its state and refined map returned), 0 missing.  u[nvar][ttd], refined[ttd]. */
int ora_probe_get_grid(const ora_mesh_t *m, int lev, const int64_t *key, double *u, uint8_t *refined)
{
    omap_t h;
    omap_build(&h, m);
    lookup_t L = {m, &h, 1 << m->ndim};
    int64_t k[3] = {key[0], key[1], key[2]};
    ghost_t g;
    const int ttd = L.ttd, nvar = m->nvar;
    int c = get_grid(&L, lev, k, &g);
    for (int ivar = 1; ivar <= nvar; ivar++)
        for (int ind = 1; ind <= ttd; ind++)
            u[(ivar - 1) * ttd + (ind - 1)] = c > 0    ? UOCT(m->uold, nvar, c, ivar, ind, ttd)
                                              : c < 0 ? g.u[(ivar - 1) * ttd + (ind - 1)]
                                                      : 0.0;
    for (int ind = 0; ind < ttd; ind++)
        refined[ind] = c > 0 ? m->refined[(size_t)(c - 1) * 8 + ind] : c < 0 ? g.refined[ind] : 0;
    omap_free(&h);
    return c;
}

int ora_upload_fine(const ora_mesh_t *m, int ilevel, double smallr)
{
    int head, noct, fhead, fnoct;
    level_range(m, ilevel, &head, &noct);
    if (ilevel >= m->nlevelmax) return 0;
    level_range(m, ilevel + 1, &fhead, &fnoct);
    int ndim = m->ndim, nvar = m->nvar, ttd = 1 << ndim;
    for (int ioct = head; ioct < head + noct; ioct++)
        for (int ivar = 1; ivar <= nvar; ivar++)
            for (int ind = 1; ind <= ttd; ind++)
                if (m->refined[(size_t)(ioct - 1) * 8 + (ind - 1)]) UOCT(m->uold, nvar, ioct, ivar, ind, ttd) = 0.0;
    omap_t h;
    omap_build(&h, m);
    lookup_t L = {m, &h, ttd};
    int rc = 0;
    for (int ioct = fhead; ioct < fhead + fnoct; ioct++) {
        int64_t kf[3] = {0, 0, 0};
        int icell = 1;
        for (int d = 0; d < ndim; d++) {
            int64_t k = m->ckey[(size_t)(ioct - 1) * 3 + d];
            kf[d] = fdiv2(k);
            icell += (1 << d) * (int)(k - 2 * kf[d]);
        }
        ghost_t gp;
        int igrid = get_grid(&L, ilevel, kf, &gp);
        if (igrid <= 0) {
            rc = 1;
            break;
        }
        for (int ivar = 1; ivar <= nvar; ivar++) {
            double average = 0.0;
            for (int ind = 1; ind <= ttd; ind++) average = average + UOCT(m->uold, nvar, ioct, ivar, ind, ttd);
            UOCT(m->uold, nvar, igrid, ivar, icell, ttd) = average / (double)ttd;
        }
        if (m->interpol_var == 1) {
            double average = 0.0;
            for (int ind = 1; ind <= ttd; ind++) {
                double ekin = 0.0;
                for (int idim = 1; idim <= 3; idim++) {
                    double mm = UOCT(m->uold, nvar, ioct, idim + 1, ind, ttd);
                    ekin = ekin + 0.5 * (mm * mm) / dmax(UOCT(m->uold, nvar, ioct, 1, ind, ttd), smallr);
                }
                average = average + UOCT(m->uold, nvar, ioct, 5, ind, ttd) - ekin - 0.0 - 0.0;
            }
            double ekin = 0.0;
            for (int idim = 1; idim <= 3; idim++) {
                double mm = UOCT(m->uold, nvar, igrid, idim + 1, icell, ttd);
                ekin = ekin + 0.5 * (mm * mm) / dmax(UOCT(m->uold, nvar, igrid, 1, icell, ttd), smallr);
            }
            UOCT(m->uold, nvar, igrid, 5, icell, ttd) = average / (double)ttd + ekin + 0.0 + 0.0;
        }
    }
    omap_free(&h);
    return rc;
}

/* courant_fine, hydro/courant_fine.f90:57-168 */
void ora_courant_fine(const ora_mesh_t *m, int ilevel, const ora_params_t *p, int nthreads, double *out)
{
    int head, noct;
    level_range(m, ilevel, &head, &noct);
    int ndim = m->ndim, nvar = m->nvar, ttd = 1 << ndim;
    double dx = m->boxlen / (double)(1LL << ilevel);
    double vol = ndim == 1 ? dx : (ndim == 2 ? dx * dx : dx * dx * dx);
    if (nthreads < 1) nthreads = 1;
    double *part = (double *)calloc((size_t)nthreads * 5, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
#endif
    for (int t = 0; t < nthreads; t++) {
        int lo = head + (int)((int64_t)noct * t / nthreads);
        int hi = head + (int)((int64_t)noct * (t + 1) / nthreads);
        double mass = 0, ekin = 0, eint = 0, emag = 0, dt = dx / p->smallc;
        for (int igrid = lo; igrid < hi; igrid++)
            for (int ind = 1; ind <= ttd; ind++) {
                if (m->refined[(size_t)(igrid - 1) * 8 + (ind - 1)]) continue;
                double uu[5];
                for (int ivar = 1; ivar <= 5; ivar++) uu[ivar - 1] = UOCT(m->uold, nvar, igrid, ivar, ind, ttd);
                mass = mass + uu[0] * vol;
                ekin = ekin + uu[4] * vol;
                eint = eint + uu[4] * vol;
                for (int idim = 1; idim <= 3; idim++)
                    eint = eint - 0.5 * (uu[idim] * uu[idim]) / dmax(uu[0], p->smallr) * vol;
                double dt_lev;
                ora_cmpdt(ndim, uu, m->constant_gravity, dx, p, &dt_lev);
                dt = dmin(dt, dt_lev);
            }
        part[5 * t + 0] = mass;
        part[5 * t + 1] = ekin;
        part[5 * t + 2] = eint;
        part[5 * t + 3] = emag;
        part[5 * t + 4] = dt;
    }
    out[0] = out[1] = out[2] = out[3] = 0.0;
    out[4] = dx / p->smallc;
    for (int t = 0; t < nthreads; t++) {
        for (int k = 0; k < 4; k++) out[k] = out[k] + part[5 * t + k];
        out[4] = dmin(out[4], part[5 * t + 4]);
    }
    free(part);
}

/* Uniform periodic level, OpenMP over octs.  Same per-oct arithmetic as
 * godfine1 (gather from the 3x3x3 neighbour octs, unsplit, update); used as
 * the CPU throughput baseline.                                               */
int ora_godunov_uniform_omp(const ora_mesh_t *m, int ilevel, double dt, const ora_params_t *pin,
                            const int32_t *octmap, int nthreads)
{
    int head, noct;
    level_range(m, ilevel, &head, &noct);
    int ndim = m->ndim, nvar = m->nvar, ttd = 1 << ndim;
    if (ndim != 3) return 3;
    if (m->box_lev > 0) return 4; /* full periodic boxes only */
    int nx = box_max(ilevel);
    double dx = m->boxlen / (double)(1LL << ilevel);
    if (nthreads < 1) nthreads = 1;
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        ora_params_t p = *pin;
        kbounds_t b;
        kbounds_init(&b, 3, 2);
        size_t nc = (size_t)b.ni * b.nj * b.nk;
        double *uloc = (double *)calloc(nc * ORA_NVAR, sizeof(double));
        double *gloc = (double *)calloc(nc * 3, sizeof(double));
        double *flux = (double *)calloc((size_t)b.fi * b.fj * b.fk * ORA_NVAR * 3, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int igrid = head; igrid < head + noct; igrid++) {
            const int32_t *ck = m->ckey + (size_t)(igrid - 1) * 3;
            for (int k1 = 0; k1 <= 2; k1++)
                for (int j1 = 0; j1 <= 2; j1++)
                    for (int i1 = 0; i1 <= 2; i1++) {
                        int ix = (ck[0] + i1 - 1 + nx) % nx, iy = (ck[1] + j1 - 1 + nx) % nx, iz = (ck[2] + k1 - 1 + nx) % nx;
                        int oc = octmap[((size_t)iz * nx + iy) * nx + ix];
                        for (int k2 = 0; k2 <= 1; k2++)
                            for (int j2 = 0; j2 <= 1; j2++)
                                for (int i2 = 0; i2 <= 1; i2++) {
                                    int ind_son = 1 + i2 + 2 * j2 + 4 * k2;
                                    int i3 = 1 + 2 * (i1 - 1) + i2, j3 = 1 + 2 * (j1 - 1) + j2, k3 = 1 + 2 * (k1 - 1) + k2;
                                    for (int ivar = 1; ivar <= nvar; ivar++)
                                        uloc[UI(&b, i3, j3, k3, ivar)] = UOCT(m->uold, nvar, oc, ivar, ind_son, ttd);
                                    for (int idim = 1; idim <= 3; idim++)
                                        gloc[UI(&b, i3, j3, k3, idim)] = m->constant_gravity[idim - 1];
                                }
                    }
            ora_unsplit(3, 2, uloc, gloc, flux, dx, dt, &p);
            for (int idim = 1; idim <= 3; idim++) {
                int i0 = idim == 1, j0 = idim == 2, k0 = idim == 3;
                for (int k2 = 0; k2 <= 1; k2++)
                    for (int j2 = 0; j2 <= 1; j2++)
                        for (int i2 = 0; i2 <= 1; i2++) {
                            int ind_son = 1 + i2 + 2 * j2 + 4 * k2;
                            int i3 = 1 + i2, j3 = 1 + j2, k3 = 1 + k2;
                            for (int ivar = 1; ivar <= 5; ivar++) {
                                double *un = &UOCT(m->unew, nvar, igrid, ivar, ind_son, ttd);
                                *un = *un + (flux[FI(&b, i3, j3, k3, ivar, idim)] -
                                             flux[FI(&b, i3 + i0, j3 + j0, k3 + k0, ivar, idim)]);
                            }
                        }
            }
        }
        free(uloc);
        free(gloc);
        free(flux);
    }
    return rc;
}
