/*
 * tbl_oracle.c -- CPU FP64 restatement of the reference's TBL time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Used by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never linked into the shipped library.
 *
 * Each function restates a block of dnssetbl/dnssetbl.f90 (cited per
 * function), keeping the reference's arithmetic order (compiled with
 * -ffp-contract=off so that no multiply-add is fused), its index tables and
 * its REAL(4) literals (1.835, 3.4, 1.7, 1e-12).  The k-slab ("MPI rank")
 * semantics that change results -- per-slab Gauss-Seidel with halos exchanged
 * after each sweep, and per-slab block SSOR -- are emulated for P slabs.
 */
#include "tbl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* threads the OpenMP loops use (1 without OpenMP) */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define NXYZ(g) ((size_t)(g)->nX * (size_t)(g)->nY * (size_t)(g)->nZ)
/* 1-based (i,j,k) -> linear offset, Fortran order */
#define I3(i, j, k) ((size_t)((i) - 1) + (size_t)nX * ((size_t)((j) - 1) + (size_t)nY * (size_t)((k) - 1)))
/* 1-based (i,k) -> (nX,nZ) offset */
#define I2(i, k) ((size_t)((i) - 1) + (size_t)nX * (size_t)((k) - 1))

/* ------------------------------------------------------------------------ */
/* a1: mesh, metrics, coefficients, index tables (dnssetbl.f90:76-92,183-247) */
/* ------------------------------------------------------------------------ */
orc_grid *orc_grid_create(int nX, int nY, int nZ, int P,
                          double Lx, double Ly, double Lz, double delta,
                          double dt, double Re, double Ue, double omega,
                          int i_rec, int freq_tim_avg, const int *planes)
{
    if (nX < 6 || nY < 6 || nZ < 6 || P < 1 || P > nZ / 3 || P > ORC_MAX_SLABS) return NULL;
    orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
    g->nX = nX; g->nY = nY; g->nZ = nZ; g->P = P;
    g->Lx = Lx; g->Ly = Ly; g->Lz = Lz; g->delta = delta;
    g->dt = dt; g->Re = Re; g->Ue = Ue; g->omega = omega;
    g->i_rec = i_rec; g->freq_tim_avg = freq_tim_avg;
    for (int q = 0; q < 5; q++) g->planes[q] = planes[q];

    g->x1 = (double *)calloc(nX + 1, sizeof(double));
    g->x2 = (double *)calloc(nY + 1, sizeof(double));
    g->x3 = (double *)calloc(nZ + 1, sizeof(double));
    g->Xi = (double *)calloc(nZ + 1, sizeof(double));
    g->Jac = (double *)calloc(nZ + 1, sizeof(double));
    g->Jac2 = (double *)calloc(nZ + 1, sizeof(double));
    g->JacInv = (double *)calloc(nZ + 1, sizeof(double));
    g->Ab = (double *)calloc(nZ + 1, sizeof(double));
    g->At = (double *)calloc(nZ + 1, sizeof(double));
    g->Ap = (double *)calloc(nZ + 1, sizeof(double));
    g->rAp = (double *)calloc(nZ + 1, sizeof(double));
    g->jm1 = (int *)calloc(nY + 1, sizeof(int));
    g->jm2 = (int *)calloc(nY + 1, sizeof(int));
    g->jp1 = (int *)calloc(nY + 1, sizeof(int));
    g->jp2 = (int *)calloc(nY + 1, sizeof(int));
    g->b = (int *)calloc(P + 1, sizeof(int));
    g->e = (int *)calloc(P + 1, sizeof(int));
    g->bn = (int *)calloc(P + 1, sizeof(int));
    g->en = (int *)calloc(P + 1, sizeof(int));

    /* slab bounds, dnssetbl.f90:76-92 (remainder planes go to the high ranks) */
    int it1 = nZ / P, it2 = P - nZ % P;
    g->b[0] = 1;
    for (int r = 0; r < P; r++) {
        if (r == it2) it1 = it1 + 1;
        g->e[r] = g->b[r] + it1 - 1;
        g->b[r + 1] = g->e[r] + 1;
    }
    for (int r = 0; r < P; r++) { g->bn[r] = g->b[r]; g->en[r] = g->e[r]; }
    g->bn[0] = g->b[0] + 1;
    g->en[P - 1] = g->e[P - 1] - 1;

    /* uniform coordinates by repeated addition, dnssetbl.f90:183-209 */
    double *x1 = g->x1, *x2 = g->x2, *x3 = g->x3, *Xi = g->Xi;
    x1[1] = 0.0; x2[1] = 0.0; Xi[1] = 0.0; x3[1] = 0.0;
    x1[nX] = Lx; x2[nY] = Ly; x3[nZ] = Lz;
    g->dx = x1[nX] / (double)(nX - 1);
    g->dy = x2[nY] / (double)(nY - 1);
    g->dXi = x3[nZ] / (double)(nZ - 1);
    g->invdx = 1.0 / g->dx;
    g->invdy = 1.0 / g->dy;
    g->invdXi = 1.0 / g->dXi;
    for (int i = 1; i <= nX - 1; i++) x1[i + 1] = x1[i] + g->dx;
    for (int i = 1; i <= nY - 1; i++) x2[i + 1] = x2[i] + g->dy;
    for (int k = 1; k <= nZ - 1; k++) Xi[k + 1] = Xi[k] + g->dXi;

    /* tanh map, dnssetbl.f90:212-216.  The reference writes 3.4d0 / 1.7d0 where
     * the literal is double and 3.4 / 1.7 (REAL(4)) elsewhere; the same
     * rounding is kept for a general Lz (half = Lz/2). */
    const double Lzd = Lz, Lhd = Lz * 0.5;
    const double Lzf = (double)(float)Lz, Lhf = (double)(float)(Lz * 0.5);
    for (int k = 1; k <= nZ; k++) {
        double a = delta * 0.5 * (Xi[k] - Lzf);
        double ch = cosh(0.5 * delta * (Xi[k] - Lzf));
        x3[k] = Lzd + Lzd * tanh(a) / (tanh(Lhd * delta));
        g->Jac[k] = Lhd * delta / (tanh(Lhf * delta) * (ch * ch));
        g->Jac2[k] = -(Lhd * (delta * delta) * tanh(0.5 * delta * (Xi[k] - Lzf)) /
                       (tanh(Lhf * delta) * (ch * ch)));
    }

    /* coefficients, dnssetbl.f90:219-232 */
    g->df1 = x3[2] - x3[1];
    g->df2 = x3[3] - x3[1];
    double dx = g->dx, dy = g->dy, dXi = g->dXi;
    g->Aw = 1.0 / (dx * dx);
    g->An = 1.0 / (dy * dy);
    g->Ae = 1.0 / (dx * dx);
    g->As = 1.0 / (dy * dy);
    for (int k = 1; k <= nZ; k++) {
        double J = g->Jac[k];
        g->Ab[k] = 1.0 / (dXi * dXi * J * J) + g->Jac2[k] / (2.0 * dXi * J * J * J);
        g->At[k] = 1.0 / (dXi * dXi * J * J) - g->Jac2[k] / (2.0 * dXi * J * J * J);
        g->Ap[k] = -2.0 * ((1.0 / (dx * dx)) + (1.0 / (dy * dy)) + (1.0 / (dXi * dXi * J * J)));
        g->rAp[k] = 1.0 / g->Ap[k];
        g->JacInv[k] = 1.0 / J;
    }

    /* periodic span tables, dnssetbl.f90:234-243 */
    for (int j = 2; j <= nY - 1; j++) {
        g->jm1[j] = j - 1; if (j == 2) g->jm1[j] = nY - 1;
        g->jm2[j] = g->jm1[j] - 1; if (j == 3) g->jm2[j] = nY - 1;
        g->jp1[j] = j + 1; if (j == nY - 1) g->jp1[j] = 2;
        g->jp2[j] = g->jp1[j] + 1; if (j == nY - 2) g->jp2[j] = 2;
    }
    g->onetwelfth = 1.0 / 12.0;
    return g;
}

void orc_grid_destroy(orc_grid *g)
{
    if (!g) return;
    free(g->x1); free(g->x2); free(g->x3); free(g->Xi);
    free(g->Jac); free(g->Jac2); free(g->JacInv);
    free(g->Ab); free(g->At); free(g->Ap); free(g->rAp);
    free(g->jm1); free(g->jm2); free(g->jp1); free(g->jp2);
    free(g->b); free(g->e); free(g->bn); free(g->en);
    free(g->ws_v); free(g->ws_r); free(g->ws_rh);
    free(g);
}

void orc_grid_mesh(const orc_grid *g, double *x1, double *x2, double *x3, double *Xi,
                   double *Jac, double *Jac2, double *JacInv,
                   double *Ab, double *At, double *Ap, double *rAp)
{
    memcpy(x1, g->x1 + 1, g->nX * sizeof(double));
    memcpy(x2, g->x2 + 1, g->nY * sizeof(double));
    memcpy(x3, g->x3 + 1, g->nZ * sizeof(double));
    memcpy(Xi, g->Xi + 1, g->nZ * sizeof(double));
    memcpy(Jac, g->Jac + 1, g->nZ * sizeof(double));
    memcpy(Jac2, g->Jac2 + 1, g->nZ * sizeof(double));
    memcpy(JacInv, g->JacInv + 1, g->nZ * sizeof(double));
    memcpy(Ab, g->Ab + 1, g->nZ * sizeof(double));
    memcpy(At, g->At + 1, g->nZ * sizeof(double));
    memcpy(Ap, g->Ap + 1, g->nZ * sizeof(double));
    memcpy(rAp, g->rAp + 1, g->nZ * sizeof(double));
}

void orc_grid_scalars(const orc_grid *g, double *o)
{
    o[0] = g->dx; o[1] = g->dy; o[2] = g->dXi;
    o[3] = g->Aw; o[4] = g->Ae; o[5] = g->As; o[6] = g->An;
    o[7] = g->df1; o[8] = g->df2;
}

/* ------------------------------------------------------------------------ */
/* a2: recycling inflow (dnssetbl.f90:259-303)                              */
/* ------------------------------------------------------------------------ */
void orc_recycle_inflow(const orc_grid *g, int nstep, double lamda,
                        const double *usum2, const double *vsum2, const double *wsum2,
                        const double *Wt, const int *newk, const double *Ume,
                        const double *Ufluc_rec, double *Ut)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ, ir = g->i_rec;
    const size_t N = NXYZ(g);
    double eps = nstep * (5.0e-5);
    if (eps > 1.0) eps = 1.0;
    if (!(nstep > g->freq_tim_avg)) return;
#define UF(j, k, p) Ufluc_rec[(size_t)((j) - 1) + (size_t)nY * ((size_t)((k) - 1) + (size_t)nZ * (size_t)((p) - 1))]
    for (int k = 2; k <= nZ - 1; k++) {
        if (usum2[I2(ir, k)] < (0.99 * g->Ue)) {
            int nk = newk[k - 1];
            double Um = eps * (usum2[I2(ir, nk)]) + (1 - eps) * Ume[k - 1];
            double Vm = lamda * vsum2[I2(ir, nk)];
            double Wm = wsum2[I2(ir, nk)];
            double W1 = Wt[I2(1, k)];
            for (int j = 2; j <= nY - 1; j++) {
                double f1 = lamda * UF(j, nk, 1);
                double f2 = lamda * UF(j, nk, 2);
                double f3 = lamda * UF(j, nk, 3);
                Ut[I3(1, j, k)]         = (Um + f1) * (1.0 - W1) + (Um + f1) * W1;
                Ut[I3(1, j, k) + N]     = (Vm + f2) * (1.0 - W1) + (Vm + f2) * W1;
                Ut[I3(1, j, k) + 2 * N] = (Wm + f3) * (1.0 - W1) + (Wm + f3) * W1;
            }
        } else {
            for (int j = 2; j <= nY - 1; j++) {
                Ut[I3(1, j, k)] = usum2[I2(ir, k)];
                Ut[I3(1, j, k) + N] = vsum2[I2(ir, k)];
                Ut[I3(1, j, k) + 2 * N] = wsum2[I2(ir, k)];
            }
        }
    }
#undef UF
}

void orc_grid_set_scheme(orc_grid *g, int scheme, const double *forc)
{
    g->scheme = scheme;
    for (int q = 0; q < 3; q++) g->forc[q] = forc ? forc[q] : 0.0;
}

/* sign(1, ceiling(a)) of dnstbl.f90:650,663,689 (INTEGER SIGN: +1 for a zero
 * second argument) */
static double sgn_ceil(double a)
{
    return ceil(a) < 0.0 ? -1.0 : 1.0;
}

/* dnstbl generation of a3+a4 (dnstbl/dnstbl.f90:620-766): Kawamura-Kuwahara
 * only inside the cell-Peclet window 2 < Pe < 50 (z: 2 < Pe < 5) with
 * Pe = sgn*Re*u*dx; u (F(i+1) - F(i-1)) / (2 dx) at i <= 2, i >= nX-1;
 * first-order upwind at j = 2, nY-1 (:670-677); Q = (U + dt (forc - conv))
 * - dt gradp (:755-761; no scalar field, so the Ri term is absent).  Its
 * Jac(k) is 1/Jac (dnstbl.f90:535), the JacInv(k) here. */
static void momentum_rhs_dnstbl(const orc_grid *g, const double *Ut, const double *Press, double *conv, double *Q)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double tf = g->Re, dx = g->dx, dy = g->dy, invdx = g->invdx, invdy = g->invdy,
                 invdXi = g->invdXi, c12 = g->onetwelfth, dt = g->dt;
    const double *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    for (int p = 1; p <= 3; p++) {
        const double *F = Ut + (size_t)(p - 1) * N;
        #pragma omp parallel for schedule(static)
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++) {
                const int jm1 = g->jm1[j], jm2 = g->jm2[j], jp1 = g->jp1[j], jp2 = g->jp2[j];
                for (int i = 2; i <= nX - 1; i++) {
                    const size_t c = I3(i, j, k);
                    const size_t cm1 = I3(i, jm1, k), cm2 = I3(i, jm2, k), cp1 = I3(i, jp1, k), cp2 = I3(i, jp2, k);
                    const double u = U[c], v = V[c], w = W[c];
                    double d1, d2, d3, Pe, t1, t2;
                    Pe = sgn_ceil(u) * tf * u * dx;
                    if (i <= 2 || i >= nX - 1) {
                        d1 = u * (F[c + 1] - F[c - 1]) / (2.0 * dx);
                    } else if (Pe > 2.0 && Pe < 50.0) {
                        t1 = (fabs(u)) * (F[c + 2] - (4 * (F[c + 1])) + (6 * (F[c])) - (4 * (F[c - 1])) + F[c - 2]) * (0.25 * invdx);
                        t2 = (u) * (-F[c + 2] + (8 * (F[c + 1])) - (8 * (F[c - 1])) + F[c - 2]) * (c12 * invdx);
                        d1 = t1 + t2;
                    } else {
                        d1 = u * ((F[c - 2] - 8 * F[c - 1] + 8 * F[c + 1] - F[c + 2]) * (c12 * invdx));
                    }
                    Pe = sgn_ceil(v) * tf * v * dy;
                    if (j == 2 || j == nY - 1) {
                        const double a1 = v > 0.0 ? v : 0.0, b1 = v < 0.0 ? v : 0.0;
                        t1 = a1 * ((F[c] - F[cm1]) * (invdy));
                        t2 = b1 * ((F[cp1] - F[c]) * (invdy));
                        d2 = t1 + t2;
                    } else if (Pe > 2.0 && Pe < 50.0) {
                        t1 = (fabs(v)) * (F[cp2] - (4 * (F[cp1])) + (6 * (F[c])) - (4 * (F[cm1])) + F[cm2]) * (0.25 * invdy);
                        t2 = (v) * (-F[cp2] + (8 * (F[cp1])) - (8 * (F[cm1])) + F[cm2]) * (c12 * invdy);
                        d2 = t1 + t2;
                    } else {
                        d2 = v * ((F[cm2] - 8 * F[cm1] + 8 * F[cp1] - F[cp2]) * (c12 * invdy));
                    }
                    Pe = (sgn_ceil(w) * tf * w * (g->x3[k + 1] - g->x3[k]));
                    if (k <= 2 || k >= nZ - 1) {
                        d3 = (w * (F[c + sk] - F[c - sk]) * (0.5 * invdXi));
                    } else if (Pe > 2.0 && Pe < 5.0) {
                        t1 = (fabs(w)) * (F[c + 2 * sk] - (4 * (F[c + sk])) + (6 * (F[c])) - (4 * (F[c - sk])) + F[c - 2 * sk]) * (0.25 * invdXi);
                        t2 = (w) * (-F[c + 2 * sk] + (8 * (F[c + sk])) - (8 * (F[c - sk])) + F[c - 2 * sk]) * (c12 * invdXi);
                        d3 = (t1 + t2);
                    } else {
                        d3 = w * ((F[c - 2 * sk] - 8 * F[c - sk] + 8 * F[c + sk] - F[c + 2 * sk]) * (c12 * invdXi));
                    }
                    const double cv = d1 + d2 + (d3 * g->JacInv[k]);
                    if (conv) conv[c + (size_t)(p - 1) * N] = cv;
                    double gp;
                    if (p == 1) gp = 0.5 * (Press[c + 1] - Press[c - 1]) * (invdx);
                    else if (p == 2) gp = 0.5 * (Press[cp1] - Press[cm1]) * (invdy);
                    else gp = 0.5 * (Press[c + sk] - Press[c - sk]) * (invdXi * g->JacInv[k]);
                    Q[c + (size_t)(p - 1) * N] = (F[c] + (dt * (g->forc[p - 1] - cv))) - (dt * gp);
                }
            }
    }
}

/* ------------------------------------------------------------------------ */
/* a3+a4: convection + old-pressure gradient + Q (dnssetbl.f90:305-369)     */
/* ------------------------------------------------------------------------ */
void orc_momentum_rhs(const orc_grid *g, const double *Ut, const double *Press,
                      double *conv, double *Q)
{
    if (g->scheme == 1) {
        momentum_rhs_dnstbl(g, Ut, Press, conv, Q);
        return;
    }
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g);
    const double Re = g->Re, dx = g->dx, dy = g->dy, invdx = g->invdx, invdy = g->invdy,
                 invdXi = g->invdXi, c12 = g->onetwelfth, dt = g->dt;
    const double *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    for (int p = 1; p <= 3; p++) {
        const double *F = Ut + (size_t)(p - 1) * N;
        #pragma omp parallel for schedule(static)
        for (int k = 2; k <= nZ - 1; k++) {
            for (int j = 2; j <= nY - 1; j++) {
                const int jm1 = g->jm1[j], jm2 = g->jm2[j], jp1 = g->jp1[j], jp2 = g->jp2[j];
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k);
                    double rt1, rt2, rt3, rt11, rt12, Pe;
                    double u = U[c], v = V[c], w = W[c];
                    /* x */
                    Pe = fabs(Re * u * dx);
                    if (i <= 2 || i >= nX - 1) {
                        rt1 = u * (F[c + 1] - F[c - 1]) * 0.5 * invdx;
                    } else if (Pe > 2.0) {
                        rt11 = fabs(u) * (F[c + 2] - (4 * (F[c + 1])) + (6 * (F[c])) - (4 * (F[c - 1])) + F[c - 2]) * (0.25 * invdx);
                        rt12 = (u) * (-F[c + 2] + (8 * (F[c + 1])) - (8 * (F[c - 1])) + F[c - 2]) * (c12 * invdx);
                        rt1 = rt11 + rt12;
                    } else {
                        rt1 = u * ((F[c - 2] - 8 * F[c - 1] + 8 * F[c + 1] - F[c + 2]) * (c12 * invdx));
                    }
                    /* y (periodic tables) */
                    size_t cm1 = I3(i, jm1, k), cm2 = I3(i, jm2, k), cp1 = I3(i, jp1, k), cp2 = I3(i, jp2, k);
                    Pe = fabs(Re * v * dy);
                    if ((j == 2) || (j == (nY - 1))) {
                        rt2 = v * (F[cp1] - F[cm1]) * 0.5 * invdy;
                    } else if (Pe > 2.0) {
                        rt11 = (fabs(v)) * (F[cp2] - (4 * (F[cp1])) + (6 * (F[c])) - (4 * (F[cm1])) + F[cm2]) * (0.25 * invdy);
                        rt12 = (v) * (-F[cp2] + (8 * (F[cp1])) - (8 * (F[cm1])) + F[cm2]) * (c12 * invdy);
                        rt2 = rt11 + rt12;
                    } else {
                        rt2 = v * ((F[cm2] - 8 * F[cm1] + 8 * F[cp1] - F[cp2]) * (c12 * invdy));
                    }
                    /* z (stretched; Peclet uses the physical spacing) */
                    size_t sk = (size_t)nX * nY;
                    Pe = fabs(Re * w * (g->x3[k + 1] - g->x3[k]));
                    if ((k <= 2) || (k >= (nZ - 1))) {
                        rt3 = (w * (F[c + sk] - F[c - sk]) * (0.5 * invdXi));
                    } else if (Pe > 2.0) {
                        rt11 = (fabs(w)) * (F[c + 2 * sk] - (4 * (F[c + sk])) + (6 * (F[c])) - (4 * (F[c - sk])) + F[c - 2 * sk]) * (0.25 * invdXi);
                        rt12 = (w) * (-F[c + 2 * sk] + (8 * (F[c + sk])) - (8 * (F[c - sk])) + F[c - 2 * sk]) * (c12 * invdXi);
                        rt3 = (rt11 + rt12);
                    } else {
                        rt3 = w * ((F[c - 2 * sk] - 8 * F[c - sk] + 8 * F[c + sk] - F[c + 2 * sk]) * (c12 * invdXi));
                    }
                    if (conv) conv[c + (size_t)(p - 1) * N] = rt1 + rt2 + (rt3 * g->JacInv[k]);
                    double cv = rt1 + rt2 + (rt3 * g->JacInv[k]);
                    /* old-pressure gradient + Q, dnssetbl.f90:355-369 */
                    double gp;
                    if (p == 1) gp = 0.5 * (Press[c + 1] - Press[c - 1]) * (invdx);
                    else if (p == 2) gp = 0.5 * (Press[cp1] - Press[cm1]) * (invdy);
                    else gp = 0.5 * (Press[c + sk] - Press[c - sk]) * (invdXi * g->JacInv[k]);
                    Q[c + (size_t)(p - 1) * N] = F[c] - dt * (cv + gp);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a5: implicit diffusion by Gauss-Seidel (dnssetbl.f90:371-433)            */
/* ------------------------------------------------------------------------ */
double orc_helm_residual(const orc_grid *g, const double *Q, const double *Ut, int p)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double *u = Ut + (size_t)(p - 1) * N, *q = Q + (size_t)(p - 1) * N;
    const double dt = g->dt, Re = g->Re;
    double part[ORC_MAX_SLABS];
#pragma omp parallel for schedule(static, 1)
    for (int r = 0; r < g->P; r++) {            /* per-slab partial, then "allreduce" */
        double tmp = 0.0;
        for (int k = g->bn[r]; k <= g->en[r]; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k);
                    double rt1 = g->Ae * u[c + 1] + g->Aw * u[c - 1] + g->An * u[I3(i, g->jp1[j], k)] + g->As * u[I3(i, g->jm1[j], k)];
                    double rt2 = g->At[k] * u[c + sk] + g->Ab[k] * u[c - sk];
                    double resLoc = q[c] - (1.0 - dt / Re * g->Ap[k]) * u[c] + dt * (rt1 + rt2) / Re;
                    tmp = tmp + fabs(resLoc);
                }
        part[r] = tmp;
    }
    double total = 0.0;
    for (int r = 0; r < g->P; r++) total += part[r];
    return total;
}

/* Wall-normal line solve of the z-line Gauss-Seidel ordering.  The column
 * system  a(k) u(k-1) + b(k) u(k) + c(k) u(k+1) = rhs(k),  k = 2..nZ-1, with
 * a = -h Ab, b = 1 - h Ap, c = -h At (h = dt/Re; the wall couplings a(2) and
 * c(nZ-1) moved into rhs) is the same for every column.  The interior planes
 * are split into C balanced chunks of at most LC planes (LC = 16 for
 * nZ-2 <= 256, else 32).  Inside a chunk of M rows a forward and a backward
 * elimination write every row through the chunk's first and last unknowns
 * (partitioned Thomas); the 2C first/last unknowns form a tridiagonal system
 * solved by Thomas; the rest follow.  All coefficient parts depend on k only
 * and are computed here once; only the rhs parts are per column. */
int orc_zline_chunks(int nz)
{
    const int LC = nz <= 256 ? 16 : 32;
    return (nz + LC - 1) / LC;
}
int orc_zline_begin(int ch, int C, int nz) { return 2 + (int)(((long)ch * nz) / C); }

int orc_zline_factors(const orc_grid *g, orc_zline *z)
{
    const int nZ = g->nZ, nz = nZ - 2;
    if (nz < 2 || nz > 512) return 1;
    const int C = orc_zline_chunks(nz);
    const double h = g->dt / g->Re;
    z->C = C;
    z->h = h;
    double *buf = (double *)calloc((size_t)5 * (nZ + 1) + (size_t)7 * C, sizeof(double));
    z->a = buf; z->R = buf + (nZ + 1); z->cf = buf + 2 * (nZ + 1); z->ap2 = buf + 3 * (nZ + 1);
    z->cp2 = buf + 4 * (nZ + 1);
    z->c0 = buf + 5 * (nZ + 1); z->rA = z->c0 + C; z->rg = z->rA + 2 * C; z->re = z->rg + 2 * C;
    double *ap = (double *)calloc(nZ + 1, sizeof(double)), *cp = (double *)calloc(nZ + 1, sizeof(double));
    double *B = (double *)calloc(2 * C, sizeof(double)), *Cc = (double *)calloc(2 * C, sizeof(double));
    for (int ch = 0; ch < C; ch++) {
        const int kb = orc_zline_begin(ch, C, nz), ke = orc_zline_begin(ch + 1, C, nz) - 1;
        const int M = ke - kb + 1;
        double a0 = (kb == 2) ? 0.0 : -(h * g->Ab[kb]);
        double b0 = 1.0 - h * g->Ap[kb];
        double c0 = (kb == nZ - 1) ? 0.0 : -(h * g->At[kb]);
        for (int k = kb + 1; k <= ke; k++) {
            const double a = -(h * g->Ab[k]), b = 1.0 - h * g->Ap[k];
            const double c = (k == nZ - 1) ? 0.0 : -(h * g->At[k]);
            if (k == kb + 1) {
                const double R = 1.0 / b;
                ap[k] = a * R; cp[k] = c * R;
                z->a[k] = 0.0; z->R[k] = R;
            } else {
                const double R = 1.0 / (b - a * cp[k - 1]);
                ap[k] = -(a * ap[k - 1]) * R; cp[k] = c * R;
                z->a[k] = a; z->R[k] = R;
            }
            z->cf[k] = cp[k];
        }
        /* backward: rows kb+1..ke-1 in terms of the first and last unknowns */
        if (M >= 3) {
            z->ap2[ke - 1] = ap[ke - 1];
            z->cp2[ke - 1] = cp[ke - 1];
            for (int k = ke - 2; k >= kb + 1; k--) {
                z->ap2[k] = ap[k] - cp[k] * z->ap2[k + 1];
                z->cp2[k] = -(cp[k] * z->cp2[k + 1]);
            }
        }
        /* reduced rows: F (first) and L (last) */
        z->rA[2 * ch] = a0;
        if (M >= 3) {
            B[2 * ch] = b0 - c0 * z->ap2[kb + 1];
            Cc[2 * ch] = -(c0 * z->cp2[kb + 1]);
            z->c0[ch] = c0;
        } else {
            B[2 * ch] = b0;
            Cc[2 * ch] = c0;
            z->c0[ch] = 0.0;
        }
        z->rA[2 * ch + 1] = ap[ke];
        B[2 * ch + 1] = 1.0;
        Cc[2 * ch + 1] = cp[ke];
    }
    /* Thomas factors of the reduced system */
    z->rg[0] = 1.0 / B[0];
    z->re[0] = Cc[0] * z->rg[0];
    for (int s2 = 1; s2 < 2 * C; s2++) {
        z->rg[s2] = 1.0 / (B[s2] - z->rA[s2] * z->re[s2 - 1]);
        z->re[s2] = Cc[s2] * z->rg[s2];
    }
    free(ap); free(cp); free(B); free(Cc);
    return 0;
}
void orc_zline_free(orc_zline *z) { free(z->a); }

static inline double helm_update(const orc_grid *g, double q, double ue, double uw, double un,
                                 double us, double ut, double ub, int k)
{
    const double dt = g->dt, Re = g->Re;
    double rt1 = g->Ae * ue + g->Aw * uw + g->An * un + g->As * us;
    double rt2 = g->At[k] * ut + g->Ab[k] * ub;
    return (q + dt / Re * (rt1 + rt2)) / (1.0 - dt / Re * g->Ap[k]);
}

void orc_helm_sweep(const orc_grid *g, const double *Q, double *Ut, int p, int order)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    double *u = Ut + (size_t)(p - 1) * N;
    const double *q = Q + (size_t)(p - 1) * N;
    if (order == ORC_GS_ZLINE_RB || order == ORC_GS_ZEBRA) {
        orc_zline zl;
        if (orc_zline_factors(g, &zl)) return;
        const int C = zl.C, nz = nZ - 2;
        double *rhs = (double *)malloc(sizeof(double) * (nZ + 1));
        double *D = (double *)malloc(sizeof(double) * 2 * C);
        /* zebra: the x neighbours of a pass are read from its start-of-pass copy */
        double *xsrc = u, *snap = NULL;
        if (order == ORC_GS_ZEBRA) snap = (double *)malloc(N * sizeof(double));
        for (int color = 0; color < 2; color++) {
            if (snap) {
                memcpy(snap, u, N * sizeof(double));
                xsrc = snap;
            }
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    const int par = (order == ORC_GS_ZEBRA) ? (j & 1) : ((i + j) & 1);
                    if (par != color) continue;
                    const size_t dn = I3(i, g->jp1[j], 1) - I3(i, j, 1), ds = I3(i, g->jm1[j], 1) - I3(i, j, 1);
                    const double h = zl.h;
                    for (int k = 2; k <= nZ - 1; k++) {
                        size_t c = I3(i, j, k);
                        double rt1 = g->Ae * xsrc[c + 1] + g->Aw * xsrc[c - 1] + g->An * u[c + dn] + g->As * u[c + ds];
                        double r = q[c] + h * rt1;
                        if (k == 2) r = r + h * (g->Ab[k] * u[c - sk]);
                        if (k == nZ - 1) r = r + h * (g->At[k] * u[c + sk]);
                        rhs[k] = r;
                    }
                    /* chunk-local elimination: rhs[] becomes d'' (rows 1..M-2), d' (row M-1) */
                    for (int ch = 0; ch < C; ch++) {
                        const int kb = orc_zline_begin(ch, C, nz), ke = orc_zline_begin(ch + 1, C, nz) - 1;
                        for (int k = kb + 1; k <= ke; k++) rhs[k] = (rhs[k] - zl.a[k] * rhs[k - 1]) * zl.R[k];
                        for (int k = ke - 2; k >= kb + 1; k--) rhs[k] = rhs[k] - zl.cf[k] * rhs[k + 1];
                        D[2 * ch] = rhs[kb] - zl.c0[ch] * rhs[kb + 1];
                        D[2 * ch + 1] = rhs[ke];
                    }
                    /* the 2C first/last unknowns: Thomas with the precomputed factors */
                    double y = D[0] * zl.rg[0];
                    D[0] = y;
                    for (int s2 = 1; s2 < 2 * C; s2++) {
                        y = (D[s2] - zl.rA[s2] * y) * zl.rg[s2];
                        D[s2] = y;
                    }
                    double x = y;
                    for (int s2 = 2 * C - 2; s2 >= 0; s2--) {
                        x = D[s2] - zl.re[s2] * x;
                        D[s2] = x;
                    }
                    for (int ch = 0; ch < C; ch++) {
                        const int kb = orc_zline_begin(ch, C, nz), ke = orc_zline_begin(ch + 1, C, nz) - 1;
                        const double xf = D[2 * ch], xl = D[2 * ch + 1];
                        u[I3(i, j, kb)] = xf;
                        for (int k = kb + 1; k <= ke - 1; k++) u[I3(i, j, k)] = rhs[k] - zl.ap2[k] * xf - zl.cp2[k] * xl;
                        u[I3(i, j, ke)] = xl;
                    }
                }
        }
        free(snap);
        free(rhs);
        free(D);
        orc_zline_free(&zl);
        return;
    }
    if (order == ORC_GS_RB) {
        for (int color = 0; color < 2; color++)
            for (int k = 2; k <= nZ - 1; k++)
                for (int j = 2; j <= nY - 1; j++)
                    for (int i = 2; i <= nX - 1; i++) {
                        if (((i + j + k) & 1) != color) continue;
                        size_t c = I3(i, j, k);
                        u[c] = helm_update(g, q[c], u[c + 1], u[c - 1], u[I3(i, g->jp1[j], k)],
                                           u[I3(i, g->jm1[j], k)], u[c + sk], u[c - sk], k);
                    }
        return;
    }
    /* per-slab lexicographic sweep; the ghost planes bn-1 / en+1 hold the
     * values from the last halo exchange (i.e. from before this sweep). */
    double *gh = (double *)malloc(sizeof(double) * 2 * sk * (size_t)g->P);
    for (int r = 0; r < g->P; r++) {
        memcpy(gh + (size_t)(2 * r) * sk, u + (size_t)(g->bn[r] - 2) * sk, sk * sizeof(double));
        memcpy(gh + (size_t)(2 * r + 1) * sk, u + (size_t)(g->en[r]) * sk, sk * sizeof(double));
    }
    #pragma omp parallel for schedule(static, 1)
    for (int r = 0; r < g->P; r++) {
        for (int k = g->bn[r]; k <= g->en[r]; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k);
                    size_t pl = c - (size_t)(k - 1) * sk;   /* offset inside the plane */
                    double ub = (k == g->bn[r]) ? gh[(size_t)(2 * r) * sk + pl] : u[c - sk];
                    double ut = (k == g->en[r]) ? gh[(size_t)(2 * r + 1) * sk + pl] : u[c + sk];
                    u[c] = helm_update(g, q[c], u[c + 1], u[c - 1], u[I3(i, g->jp1[j], k)],
                                       u[I3(i, g->jm1[j], k)], ut, ub, k);
                }
    }
    free(gh);
}

void orc_diffusion_solve(const orc_grid *g, const double *Q, double *Ut,
                         int order, double errorstop, int mmax, int *sweeps, double *res_out)
{
    const double tiny = (double)1e-12f;   /* REAL(4) literal, dnssetbl.f90:392 */
    for (int p = 1; p <= 3; p++) {
        double sumnor = 0.0, res = 0.0;
        int ns = 0;
        for (int it = 1; it <= mmax; it++) {
            if (it % 2 == 1) {
                res = orc_helm_residual(g, Q, Ut, p);
                if (it == 1) sumnor = res;
                double sumav = res / sumnor;
                if (sumav < errorstop || res < tiny) break;
            }
            orc_helm_sweep(g, Q, Ut, p, order);
            ns++;
        }
        if (sweeps) sweeps[p - 1] = ns;
        if (res_out) res_out[p - 1] = res;
    }
}

/* ------------------------------------------------------------------------ */
/* a6: PPE right-hand side (dnssetbl.f90:437-447)                           */
/* ------------------------------------------------------------------------ */
void orc_ppe_rhs(const orc_grid *g, const double *Ut, double *pRhs)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 2; i <= nX - 1; i++) {
                size_t c = I3(i, j, k), cp = I3(i, g->jp1[j], k), cm = I3(i, g->jm1[j], k);
                double rt1 = (((U[c + 1] + U[c]) * 0.5) - (((U[c] + U[c - 1]) * 0.5))) * g->invdx;
                double rt2 = (((V[cp] + V[c]) * 0.5) - (((V[c] + V[cm]) * 0.5))) * g->invdy;
                double rt3 = 2.0 * (((W[c + sk] + W[c]) * 0.5) - (((W[c] + W[c - sk]) * 0.5))) / (g->x3[k + 1] - g->x3[k - 1]);
                pRhs[c] = (rt1 + rt2 + rt3) / g->dt;
            }
}

/* ------------------------------------------------------------------------ */
/* a7: Poisson operator (dnssetbl.f90:498-507, 574-583)                     */
/* ------------------------------------------------------------------------ */
void orc_poisson_apply(const orc_grid *g, const double *x, double *out)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t sk = (size_t)nX * nY;
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 2; i <= nX - 1; i++) {
                size_t c = I3(i, j, k);
                out[c] = g->Aw * x[c - 1] + g->Ae * x[c + 1] +
                         g->As * x[I3(i, g->jm1[j], k)] + g->An * x[I3(i, g->jp1[j], k)] +
                         g->Ab[k] * x[c - sk] + g->At[k] * x[c + sk] +
                         g->Ap[k] * x[c];
            }
}

/* ------------------------------------------------------------------------ */
/* a8: preconditioners (reference: dnssetbl.f90:509-532, 586-610)           */
/* ------------------------------------------------------------------------ */
/* the grid's zero-initialised work vector of N values (allocated on first use) */
static double *grid_work(const orc_grid *gc, double **slot)
{
    if (!*slot) *slot = (double *)calloc(NXYZ(gc), sizeof(double));
    return *slot;
}

void orc_precond_apply(const orc_grid *g, int pc, const double *r, double *out)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double w = g->omega;
    orc_grid *gm = (orc_grid *)g;          /* work arrays only; the grid's data are unchanged */
    double *rh = grid_work(g, &gm->ws_rh);
    if (pc == ORC_PC_NONE || pc == ORC_PC_JACOBI) {
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k);
                    rh[c] = (pc == ORC_PC_NONE) ? r[c] : r[c] * g->rAp[k];
                }
    } else if (pc == ORC_PC_SSOR_BLOCK) {
        /* one SSOR per slab; rhat outside the slab interior stays zero
         * (never written on that rank), so slabs are decoupled. */
        #pragma omp parallel for schedule(static, 1)
        for (int s = 0; s < g->P; s++) {
            const int k0 = g->bn[s], k1 = g->en[s];
            for (int k = k0; k <= k1; k++)
                for (int j = 2; j <= nY - 1; j++)
                    for (int i = 2; i <= nX - 1; i++) {
                        size_t c = I3(i, j, k);
                        double rb = (k == k0) ? 0.0 : rh[c - sk];
                        rh[c] = -w * (r[c] + g->Aw * rh[c - 1] + g->As * rh[c - nX] + g->Ab[k] * rb) * g->rAp[k];
                    }
            for (int k = k0; k <= k1; k++)
                for (int j = 2; j <= nY - 1; j++)
                    for (int i = 2; i <= nX - 1; i++) {
                        size_t c = I3(i, j, k);
                        rh[c] = ((2 - w) / w) * g->Ap[k] * rh[c];
                    }
            for (int k = k1; k >= k0; k--)
                for (int j = nY - 1; j >= 2; j--)
                    for (int i = nX - 1; i >= 2; i--) {
                        size_t c = I3(i, j, k);
                        double rt = (k == k1) ? 0.0 : rh[c + sk];
                        rh[c] = -w * (rh[c] + g->Ae * rh[c + 1] + g->An * rh[c + nX] + g->At[k] * rt) * g->rAp[k];
                    }
        }
    } else { /* ORC_PC_SSOR_RB: red (i+j+k even) first; full periodic couplings */
        for (int color = 0; color < 2; color++)
            for (int k = 2; k <= nZ - 1; k++)
                for (int j = 2; j <= nY - 1; j++)
                    for (int i = 2; i <= nX - 1; i++) {
                        if (((i + j + k) & 1) != color) continue;
                        size_t c = I3(i, j, k);
                        double low = 0.0;
                        if (color == 1)
                            low = g->Aw * rh[c - 1] + g->Ae * rh[c + 1] + g->As * rh[I3(i, g->jm1[j], k)] +
                                  g->An * rh[I3(i, g->jp1[j], k)] + g->Ab[k] * rh[c - sk] + g->At[k] * rh[c + sk];
                        rh[c] = -w * (r[c] + low) * g->rAp[k];
                    }
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k);
                    rh[c] = ((2 - w) / w) * g->Ap[k] * rh[c];
                }
        for (int color = 1; color >= 0; color--)
            for (int k = 2; k <= nZ - 1; k++)
                for (int j = 2; j <= nY - 1; j++)
                    for (int i = 2; i <= nX - 1; i++) {
                        if (((i + j + k) & 1) != color) continue;
                        size_t c = I3(i, j, k);
                        double up = 0.0;
                        if (color == 0)
                            up = g->Aw * rh[c - 1] + g->Ae * rh[c + 1] + g->As * rh[I3(i, g->jm1[j], k)] +
                                 g->An * rh[I3(i, g->jp1[j], k)] + g->Ab[k] * rh[c - sk] + g->At[k] * rh[c + sk];
                        rh[c] = -w * (rh[c] + up) * g->rAp[k];
                    }
    }
    memcpy(out, rh, N * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* a9: restarted left-preconditioned GMRES (dnssetbl.f90:449-671)           */
/* ------------------------------------------------------------------------ */
/* interior dot product as the reference's ranks form it: a partial sum per
 * slab (k = bn..en, in i-fastest order), then the partials summed in slab
 * order (one allreduce per dot, :615,620); the slabs run in parallel, so the
 * result depends on the slab count P but not on the thread count */
static double dot_int(const orc_grid *g, const double *a, const double *b)
{
    const int nX = g->nX, nY = g->nY;
    double part[ORC_MAX_SLABS];
#pragma omp parallel for schedule(static, 1)
    for (int r = 0; r < g->P; r++) {
        double s = 0.0;
        for (int k = g->bn[r]; k <= g->en[r]; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) { size_t c = I3(i, j, k); s += a[c] * b[c]; }
        part[r] = s;
    }
    double s = 0.0;
    for (int r = 0; r < g->P; r++) s += part[r];
    return s;
}

static void zero_faces(const orc_grid *g, double *x)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    for (int k = 1; k <= nZ; k++)
        for (int j = 1; j <= nY; j++)
            for (int i = 1; i <= nX; i++)
                if (i == 1 || i == nX || j == 1 || j == nY || k == 1 || k == nZ) x[I3(i, j, k)] = 0.0;
}

void orc_ppe_gmres(const orc_grid *g, const double *pRhs, double *PCorr,
                   int pc, double errtol, int kmax, int mmax, double *stats)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g);
    const int ld = kmax + 1;
    orc_grid *gm = (orc_grid *)g;          /* work arrays only; the grid's data are unchanged */
    if (gm->ws_nv < kmax + 1) {
        free(gm->ws_v);
        gm->ws_v = (double *)calloc(N * (size_t)(kmax + 1), sizeof(double));
        gm->ws_nv = kmax + 1;
    }
    double *v = gm->ws_v;
    double *r = grid_work(g, &gm->ws_r);
    double *h = (double *)calloc((size_t)ld * ld, sizeof(double));
    double *cs = (double *)calloc(ld + 1, sizeof(double)), *sn = (double *)calloc(ld + 1, sizeof(double));
    double *gg = (double *)calloc(ld + 1, sizeof(double)), *y = (double *)calloc(ld + 1, sizeof(double));
#define H(a, b) h[(size_t)((a) - 1) + (size_t)ld * (size_t)((b) - 1)]
#define VV(q) (v + (size_t)((q) - 1) * N)

    zero_faces(g, PCorr);   /* dnssetbl.f90:451-460 (interior kept: warm start) */
    double sumav = 1.0, sumnor = 1.0;
    int m = 0, total = 0, kit = 0;
    while ((sumav > errtol) && (m < mmax)) {
        m = m + 1;
        memset(h, 0, sizeof(double) * (size_t)ld * ld);
        memset(cs, 0, sizeof(double) * (ld + 1)); memset(sn, 0, sizeof(double) * (ld + 1));
        memset(gg, 0, sizeof(double) * (ld + 1)); memset(y, 0, sizeof(double) * (ld + 1));
        /* r = pRhs - A PCorr (dnssetbl.f90:498-507) */
        memset(r, 0, N * sizeof(double));
        #pragma omp parallel for schedule(static)
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k), sk = (size_t)nX * nY;
                    r[c] = pRhs[c] - (g->Aw * PCorr[c - 1] + g->Ae * PCorr[c + 1] +
                                      g->As * PCorr[I3(i, g->jm1[j], k)] + g->An * PCorr[I3(i, g->jp1[j], k)] +
                                      g->Ab[k] * PCorr[c - sk] + g->At[k] * PCorr[c + sk] +
                                      g->Ap[k] * PCorr[c]);
                }
        orc_precond_apply(g, pc, r, r);
        double rho = sqrt(dot_int(g, r, r));
        if (m == 1) sumnor = rho;
        sumav = rho / sumnor;
        if (rho == 0.0) { sumav = 0.0; kit = 0; break; }
        gg[1] = rho;
        #pragma omp parallel for schedule(static)
        for (size_t c = 0; c < N; c++) VV(1)[c] = r[c] / rho;   /* boundary of r is 0 */
        kit = 0;
        while ((sumav > errtol) && (kit < kmax)) {
            kit = kit + 1;
            total++;
            double *w = VV(kit + 1);
            orc_poisson_apply(g, VV(kit), w);
            orc_precond_apply(g, pc, w, w);
            /* modified Gram-Schmidt, one reduction per dot (dnssetbl.f90:613-617) */
            for (int q = 1; q <= kit; q++) {
                double hq = dot_int(g, VV(q), w);
                H(q, kit) = hq;
                const double *vq = VV(q);
                #pragma omp parallel for schedule(static)
                for (size_t c = 0; c < N; c++) w[c] = w[c] - hq * vq[c];
            }
            H(kit + 1, kit) = sqrt(dot_int(g, w, w));
            if (H(kit + 1, kit) > 0.0 || H(kit + 1, kit) < 0.0) {
                double hn = H(kit + 1, kit);
                #pragma omp parallel for schedule(static)
                for (size_t c = 0; c < N; c++) w[c] = w[c] / hn;
            }
            if (kit > 1) {   /* old Givens rotations (dnssetbl.f90:627-635) */
                for (int i = 1; i <= kit - 1; i++) {
                    double hik = cs[i] * H(i, kit) - sn[i] * H(i + 1, kit);
                    double hipk = sn[i] * H(i, kit) + cs[i] * H(i + 1, kit);
                    H(i, kit) = hik;
                    H(i + 1, kit) = hipk;
                }
            }
            double nu = sqrt(H(kit, kit) * H(kit, kit) + H(kit + 1, kit) * H(kit + 1, kit));
            if (nu > 0.0) {   /* new rotation (dnssetbl.f90:637-649) */
                cs[kit] = H(kit, kit) / nu;
                sn[kit] = -H(kit + 1, kit) / nu;
                H(kit, kit) = cs[kit] * H(kit, kit) - sn[kit] * H(kit + 1, kit);
                H(kit + 1, kit) = 0;
                double gk = cs[kit] * gg[kit] - sn[kit] * gg[kit + 1];
                double gkp = sn[kit] * gg[kit] + cs[kit] * gg[kit + 1];
                gg[kit] = gk;
                gg[kit + 1] = gkp;
            }
            rho = fabs(gg[kit + 1]);
            sumav = rho / sumnor;
        }
        if (kit == 0) continue;   /* the reference would index y(0) here (reads out of bounds) */
        /* back substitution + update (dnssetbl.f90:657-669) */
        y[kit] = gg[kit] / H(kit, kit);
        for (int i = kit - 1; i >= 1; i--) {
            y[i] = gg[i];
            for (int k = i + 1; k <= kit; k++) y[i] = y[i] - H(i, k) * y[k];
            y[i] = y[i] / H(i, i);
        }
        for (int i = 1; i <= kit; i++) {
            const double *vi = VV(i);
            #pragma omp parallel for schedule(static)
            for (int k = 2; k <= nZ - 1; k++)
                for (int j = 2; j <= nY - 1; j++)
                    for (int ii = 2; ii <= nX - 1; ii++) {
                        size_t c = I3(ii, j, k);
                        PCorr[c] = PCorr[c] + vi[c] * y[i];
                    }
        }
    }
    if (stats) {
        /* true residual after the solve */
        double s2 = 0.0;
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++)
                for (int i = 2; i <= nX - 1; i++) {
                    size_t c = I3(i, j, k), sk = (size_t)nX * nY;
                    double rr = pRhs[c] - (g->Aw * PCorr[c - 1] + g->Ae * PCorr[c + 1] +
                                           g->As * PCorr[I3(i, g->jm1[j], k)] + g->An * PCorr[I3(i, g->jp1[j], k)] +
                                           g->Ab[k] * PCorr[c - sk] + g->At[k] * PCorr[c + sk] +
                                           g->Ap[k] * PCorr[c]);
                    s2 += rr * rr;
                }
        stats[0] = m; stats[1] = total; stats[2] = kit; stats[3] = sumav;
        stats[4] = sumnor; stats[5] = sqrt(s2);
    }
#undef H
#undef VV
    free(h); free(cs); free(sn); free(gg); free(y);
}

/* ------------------------------------------------------------------------ */
/* a10: PCorr extrapolation + pressure update + projection (:697-726)      */
/* ------------------------------------------------------------------------ */
void orc_project(const orc_grid *g, double *PCorr, double *Press, double *Ut)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    /* wall-normal faces (rank 0 bottom, last rank top), :697-699 */
    for (int j = 2; j <= nY - 1; j++)
        for (int i = 1; i <= nX; i++) {
            PCorr[I3(i, j, 1)] = ((4.0 * PCorr[I3(i, j, 2)]) - PCorr[I3(i, j, 3)]) / 3.0;
            PCorr[I3(i, j, nZ)] = (5.0 * PCorr[I3(i, j, nZ - 1)] + (-4.0 * PCorr[I3(i, j, nZ - 2)] + PCorr[I3(i, j, nZ - 3)])) / 2.0;
        }
    /* span copies, :701-702 */
    for (int k = 1; k <= nZ; k++)
        for (int i = 1; i <= nX; i++) {
            PCorr[I3(i, 1, k)] = PCorr[I3(i, nY - 1, k)];
            PCorr[I3(i, nY, k)] = PCorr[I3(i, 2, k)];
        }
    /* inflow / outflow faces, :705-708 */
    for (int k = 1; k <= nZ; k++)
        for (int j = 2; j <= nY - 1; j++) {
            PCorr[I3(1, j, k)] = (4.0 * PCorr[I3(2, j, k)] - PCorr[I3(3, j, k)]) / 3.0;
            PCorr[I3(nX, j, k)] = (5.0 * PCorr[I3(nX - 1, j, k)] + (-4.0 * PCorr[I3(nX - 2, j, k)] + PCorr[I3(nX - 3, j, k)])) / 2.0;
        }
    /* pressure update + projection, :711-726 */
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 2; i <= nX - 1; i++) {
                size_t c = I3(i, j, k);
                Press[c] = Press[c] + PCorr[c];
            }
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 2; i <= nX - 1; i++) {
                size_t c = I3(i, j, k);
                double g1 = 0.5 * (PCorr[c + 1] - PCorr[c - 1]) * (g->invdx);
                double g2 = 0.5 * (PCorr[I3(i, g->jp1[j], k)] - PCorr[I3(i, g->jm1[j], k)]) * (g->invdy);
                double g3 = 0.5 * (PCorr[c + sk] - PCorr[c - sk]) * (g->invdXi * g->JacInv[k]);
                Ut[c] = Ut[c] - (g->dt * g1);
                Ut[c + N] = Ut[c + N] - (g->dt * g2);
                Ut[c + 2 * N] = Ut[c + 2 * N] - (g->dt * g3);
            }
}

/* ------------------------------------------------------------------------ */
/* a11: velocity and pressure boundary conditions (:728-797)               */
/* ------------------------------------------------------------------------ */
void orc_apply_bcs(const orc_grid *g, double *Ut, const double *Utold, double *Press,
                   const double *wall_w)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double dx = g->dx, dy = g->dy, dXi = g->dXi, Re = g->Re, dt = g->dt;
    const double df1 = g->df1, df2 = g->df2;
    for (int p = 0; p < 3; p++) {
        double *u = Ut + (size_t)p * N;
        /* wall: no slip (w = wall_w is the blowing/suction control hook), :729 */
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 1; i <= nX; i++)
                u[I3(i, j, 1)] = (p == 2 && wall_w) ? wall_w[(size_t)(i - 1) + (size_t)nX * (j - 1)] : 0.0;
        /* top: 2nd-order zero gradient, :731 */
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 1; i <= nX; i++)
                u[I3(i, j, nZ)] = (4.0 * u[I3(i, j, nZ - 1)] - u[I3(i, j, nZ - 2)]) / 3.0;
        /* outflow: 2nd-order zero gradient, :733 */
        for (int k = 1; k <= nZ; k++)
            for (int j = 2; j <= nY - 1; j++)
                u[I3(nX, j, k)] = (4.0 * u[I3(nX - 1, j, k)] - u[I3(nX - 2, j, k)]) / 3.0;
        /* span periodic copies, :735-736 */
        for (int k = 1; k <= nZ; k++)
            for (int i = 1; i <= nX; i++) {
                u[I3(i, 1, k)] = u[I3(i, nY - 1, k)];
                u[I3(i, nY, k)] = u[I3(i, 2, k)];
            }
    }
    const double *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    /* bottom pressure from wall-normal momentum, :738-755 */
    for (int j = 2; j <= nY - 1; j++)
        for (int i = 2; i <= nX - 1; i++) {
            double w1 = W[I3(i, j, 1)], w2 = W[I3(i, j, 2)], w3 = W[I3(i, j, 3)];
            double rt11 = w1 * (w2 - w1) / df1;
            double rt1 = (W[I3(i - 1, j, 1)] - 2.0 * w1 + W[I3(i + 1, j, 1)]) / (dx * dx);
            double rt2 = (W[I3(i, j - 1, 1)] - 2.0 * w1 + W[I3(i, j + 1, 1)]) / (dy * dy);
            double rt3 = (2.0 * df1 * w3 - 2.0 * df2 * w2 + 2.0 * (df2 - df1) * w1) / (df1 * (df2 * df2) - (df2 * (df1 * df1)));
            double rt12 = -rt11 + (rt1 + rt2 + rt3) / Re;
            Press[I3(i, j, 1)] = Press[I3(i, j, 2)] - rt12 * df1;
        }
    /* top pressure, :757-765 */
    for (int j = 2; j <= nY - 1; j++)
        for (int i = 2; i <= nX - 1; i++) {
            double rt1 = (11.0 * W[I3(i, j, nZ)] - 18.0 * W[I3(i, j, nZ - 1)] + 9.0 * W[I3(i, j, nZ - 2)] - 2.0 * W[I3(i, j, nZ - 3)]) / (6.0 * dXi);
            double rt2 = (rt1 * g->JacInv[nZ]);
            Press[I3(i, j, nZ)] = rt2 / Re;
        }
    /* outflow pressure, :767-771 */
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            Press[I3(nX, j, k)] = ((3.0 * U[I3(nX, j, k)] - 4.0 * U[I3(nX - 1, j, k)] + U[I3(nX - 2, j, k)]) / (2.0 * dx * Re));
    /* inflow pressure from streamwise momentum, :773-794 */
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++) {
            size_t c = I3(1, j, k);
            double u1 = U[c];
            double rt1 = u1 * ((U[c + 1] - u1) / (dx));
            double rt2 = V[c] * ((U[I3(1, j + 1, k)] - U[I3(1, j - 1, k)]) / (2.0 * dy));
            double rt3 = W[c] * ((U[c + sk] - U[c - sk]) / (2.0 * dXi));
            double rt4 = (u1 + U[c + 2] - 2 * U[c + 1]) / (dx * dx);
            double rt5 = (U[I3(1, j + 1, k)] + U[I3(1, j - 1, k)] - 2 * u1) / (dy * dy);
            double rt11, rt6;
            if (k <= 2) {
                rt11 = (U[c + sk] - u1) / dXi;
                rt6 = (U[c + 2 * sk] - 2.0 * U[c + sk] + u1) / (dXi * dXi);
            } else if (k >= nZ - 1) {
                rt11 = (u1 - U[c - sk]) / dXi;
                rt6 = (U[c - 2 * sk] - 2.0 * U[c - sk] + u1) / (dXi * dXi);
            } else {
                rt11 = (U[c + sk] - U[c - sk]) / (2.0 * dXi);
                rt6 = (-2 * u1 + U[c + sk] + U[c - sk]) / (dXi * dXi);
            }
            double Ji = g->JacInv[k];
            rt6 = rt6 * (Ji * Ji) - g->Jac2[k] * rt11 * (Ji * Ji * Ji);
            double rt12 = (u1 - Utold[c]) / dt;
            Press[c] = ((2 * dx * ((rt12 + rt1 + rt2 + rt3 * Ji) - ((rt4 + rt5 + rt6) / Re))) - Press[c + 2] + (4 * Press[c + 1])) / 3.0;
        }
    /* span periodic pressure, :796-797 */
    for (int k = 1; k <= nZ; k++)
        for (int i = 1; i <= nX; i++) {
            Press[I3(i, 1, k)] = Press[I3(i, nY - 1, k)];
            Press[I3(i, nY, k)] = Press[I3(i, 2, k)];
        }
}

/* ------------------------------------------------------------------------ */
/* a12: divergence check (:851-882)                                         */
/* ------------------------------------------------------------------------ */
double orc_div_check(const orc_grid *g, const double *Ut, double *signed_sum)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    const double c12 = g->onetwelfth;
    double tmp = 0.0;
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 2; i <= nX - 1; i++) {
                size_t c = I3(i, j, k);
                double rt1, rt2, rt3;
                if ((i == 2) || (i == (nX - 1)))
                    rt1 = (U[c + 1] - U[c - 1]) * (0.5 * g->invdx);
                else
                    rt1 = ((U[c - 2] - 8 * U[c - 1] + 8 * U[c + 1] - U[c + 2]) * (c12 * g->invdx));
                size_t cm1 = I3(i, g->jm1[j], k), cm2 = I3(i, g->jm2[j], k), cp1 = I3(i, g->jp1[j], k), cp2 = I3(i, g->jp2[j], k);
                if ((j == 2) || (j == (nY - 1)))
                    rt2 = (V[cp1] - V[cm1]) * (0.5 * g->invdy);
                else
                    rt2 = ((V[cm2] - 8 * V[cm1] + 8 * V[cp1] - V[cp2]) * (c12 * g->invdy));
                if (k <= 2 || k >= (nZ - 1))
                    rt3 = ((W[c + sk] - W[c - sk]) * (0.5 * g->invdXi));
                else
                    rt3 = ((W[c - 2 * sk] - 8 * W[c - sk] + 8 * W[c + sk] - W[c + 2 * sk]) * (c12 * g->invdXi));
                tmp = tmp + (rt1 + rt2 + rt3 * g->JacInv[k]);
            }
    if (signed_sum) *signed_sum = tmp;
    return fabs(tmp) / ((nY - 2) * (nX - 2) * (nZ - 2));
}

/* ------------------------------------------------------------------------ */
/* a13: mass flux through 5 planes (:885-904)                               */
/* ------------------------------------------------------------------------ */
void orc_mass_flux(const orc_grid *g, const double *Ut, double *mass5)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t sk = (size_t)nX * nY;
    for (int q = 0; q < 5; q++) {
        int pl = g->planes[q];
        double m = 0.0;
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++) {
                double dAv = g->dy * (g->x3[k + 1] - g->x3[k]);
                size_t c = I3(pl, j, k), cp = I3(pl, g->jp1[j], k);
                m = m + (0.25 * (Ut[c] + Ut[cp] + Ut[c + sk] + Ut[cp + sk]) * (dAv));
            }
        mass5[q] = m / ((g->x3[nZ - 1] - g->x3[2]) * (g->x2[nY - 1] - g->x2[2]));
    }
}

/* ------------------------------------------------------------------------ */
/* a14: spanwise statistics accumulation (:936-984)                         */
/* order of the 9 sums: u v w uu vv ww uv uw vw                             */
/* ------------------------------------------------------------------------ */
void orc_stats_accumulate(const orc_grid *g, const double *Ut, double *sums)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ;
    const size_t N = NXYZ(g), M = (size_t)nX * nZ;
    const double *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    for (int i = 1; i <= nX; i++)
        for (int k = 1; k <= nZ; k++) {
            double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = 2; j <= nY - 1; j++) {
                size_t c = I3(i, j, k), cp = I3(i, g->jp1[j], k);
                double su = 0.50 * (U[c] + U[cp]);
                double sv = 0.50 * (V[c] + V[cp]);
                double sw = 0.50 * (W[c] + W[cp]);
                t[0] = t[0] + su; t[1] = t[1] + sv; t[2] = t[2] + sw;
                t[3] = t[3] + su * su; t[4] = t[4] + sv * sv; t[5] = t[5] + sw * sw;
                t[6] = t[6] + su * sv; t[7] = t[7] + su * sw; t[8] = t[8] + sv * sw;
            }
            double A = nY - 2;
            for (int q = 0; q < 9; q++) sums[q * M + I2(i, k)] = sums[q * M + I2(i, k)] + (t[q] / A);
        }
}

/* ------------------------------------------------------------------------ */
/* a15: time average + recycling rescale (:1015-1153, rank 0 part)          */
/* ------------------------------------------------------------------------ */
void orc_rescale(const orc_grid *g, const double *sums, int nsamples, const double *Ut,
                 double *avg, double *lamda, double *lamda03, double *Wt, int *newk,
                 double *bltkns, double *deltastar, double *theta, double *Ufluc_rec)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ, ir = g->i_rec;
    const size_t N = NXYZ(g), M = (size_t)nX * nZ;
    const double Ue = g->Ue, *x3 = g->x3;
    for (int q = 0; q < 9; q++)
        for (size_t c = 0; c < M; c++) avg[q * M + c] = (sums[q * M + c] / nsamples);
    const double *usum2 = avg, *vsum2 = avg + M, *wsum2 = avg + 2 * M;
    /* fluctuations at the recycling plane (:1034-1042) */
    if (Ufluc_rec && ir <= nX - 1)
        for (int k = 2; k <= nZ - 1; k++)
            for (int j = 2; j <= nY - 1; j++) {
                size_t c = I3(ir, j, k), o = (size_t)(j - 1) + (size_t)nY * (k - 1);
                Ufluc_rec[o] = Ut[c] - usum2[I2(ir, k)];
                Ufluc_rec[o + (size_t)nY * nZ] = Ut[c + N] - vsum2[I2(ir, k)];
                Ufluc_rec[o + 2 * (size_t)nY * nZ] = Ut[c + 2 * N] - wsum2[I2(ir, k)];
            }
    /* friction velocities and their ratio (:1055-1058) */
    double utau1 = sqrt(((usum2[I2(1, 2)] - usum2[I2(1, 1)]) / (x3[2] - x3[1])) / g->Re);
    double utau2 = sqrt(((usum2[I2(ir, 2)] - usum2[I2(ir, 1)]) / (x3[2] - x3[1])) / g->Re);
    *lamda = (utau1) / (utau2);
    /* delta99: first node with U >= 0.99 Ue, else unchanged (:1062-1070) */
    for (int i = 1; i <= nX; i++)
        for (int k = 1; k <= nZ; k++)
            if (usum2[I2(i, k)] >= (0.99 * Ue)) { bltkns[i - 1] = x3[k]; break; }
    *lamda03 = bltkns[ir - 1] / bltkns[0];
    /* Lund weight (:1080-1084) */
    for (int i = 1; i <= nX; i++)
        for (int k = 1; k <= nZ; k++) {
            double eta = x3[k] / bltkns[i - 1];
            Wt[I2(i, k)] = 0.5 * (1 + (tanh((4.0 * (eta - 0.2)) / ((1 - 2.0 * 0.2) * eta + 0.2)) / tanh(4.0)));
        }
    /* recycling index map new(k) (:1096-1129) */
    double *eta = (double *)calloc(nZ + 1, sizeof(double));
    double *zrecy = (double *)calloc(nZ + 1, sizeof(double));
    for (int k = 1; k <= nZ; k++) newk[k - 1] = nZ - 1;
    int ksep = 0;
    for (int k = 1; k <= nZ - 1; k++) {
        if (Wt[I2(ir, k)] <= 0.99) { ksep = k; eta[k] = *lamda; }
        else eta[k] = *lamda03;
    }
    for (int k = 1; k <= nZ - 1; k++) zrecy[k] = eta[k] * x3[k];
    int counter = 1;
    for (int k = 1; k <= nZ - 1; k++)
        if (counter <= ksep)
            if (x3[k] >= zrecy[counter]) { newk[counter - 1] = k; counter = counter + 1; }
    counter = ksep + 1;
    for (int k = ksep + 1; k <= nZ - 1; k++)
        if (counter <= (nZ - 1))
            if (x3[k] >= zrecy[counter]) { newk[counter - 1] = k; counter = counter + 1; }
    free(eta); free(zrecy);
    /* integral thicknesses (:1131-1145), fresh sums each call */
    for (int i = 1; i <= nX; i++) {
        double ds = 0.0;
        for (int k = 1; k <= nZ - 1; k++) {
            ds = ds + ((1.0 - (usum2[I2(i, k)] / Ue)) * (x3[k + 1] - x3[k]));
            if (x3[k] == bltkns[i - 1]) break;
        }
        deltastar[i - 1] = ds;
        double th = 0.0;
        for (int k = 1; k <= nZ - 1; k++) {
            th = th + ((usum2[I2(i, k)] / Ue) * (1.0 - (usum2[I2(i, k)] / Ue)) * (x3[k + 1] - x3[k]));
            if (x3[k] >= bltkns[i - 1]) break;
        }
        theta[i - 1] = th;
    }
}

/* ------------------------------------------------------------------------ */
/* a16: fluctuation kinetic energy + blow-up guard (:1223-1261)             */
/* ------------------------------------------------------------------------ */
double orc_ke_guard(const orc_grid *g, const double *Ut, const double *usum2,
                    const double *vsum2, const double *wsum2, int *guard, double *Ufluc_rec)
{
    const int nX = g->nX, nY = g->nY, nZ = g->nZ, ir = g->i_rec;
    const size_t N = NXYZ(g);
    double ke = 0.0;
    int gd = 0;
    for (int k = 2; k <= nZ - 1; k++)
        for (int j = 2; j <= nY - 1; j++)
            for (int i = 2; i <= nX - 1; i++) {
                size_t c = I3(i, j, k);
                double f1 = Ut[c] - usum2[I2(i, k)];
                double f2 = Ut[c + N] - vsum2[I2(i, k)];
                double f3 = Ut[c + 2 * N] - wsum2[I2(i, k)];
                ke = ke + f1 * f1 + f2 * f2 + f3 * f3;
                double mx = f1 > f2 ? f1 : f2; mx = mx > f3 ? mx : f3;
                if (mx > 25.0) gd = 1;
                if (i == ir && Ufluc_rec) {
                    size_t o = (size_t)(j - 1) + (size_t)nY * (k - 1);
                    Ufluc_rec[o] = f1;
                    Ufluc_rec[o + (size_t)nY * nZ] = f2;
                    Ufluc_rec[o + 2 * (size_t)nY * nZ] = f3;
                }
            }
    if (guard) *guard = gd;
    return ke / (double)(nX * (nY - 2) * (nZ - 2));
}

/* ------------------------------------------------------------------------ */
/* a17: wall shear per streamwise station (dnstbl/dnstbl.f90:1620-1668)      */
/* ------------------------------------------------------------------------ */
void orc_wall_shear(const orc_grid *g, const double *Ut, double *tau)
{
    const int nX = g->nX, nY = g->nY;
    const size_t N = NXYZ(g), sk = (size_t)nX * nY;
    const double *U = Ut, *W = Ut + 2 * N;
    const double dx = g->dx, df1 = g->df1, df2 = g->df2;
    for (int i = 1; i <= nX; i++) tau[i - 1] = 0.0;
    for (int i = 2; i <= nX - 1; i++) {
        double t = 0.0;
        for (int j = 2; j <= nY - 1; j++) {
            size_t c = I3(i, j, 1);
            double mfb1;
            if (i <= 3 || i >= nX - 2)
                mfb1 = (W[c + 1] - W[c - 1]) / (2.0 * dx);
            else
                mfb1 = (-W[c - 3] + (9.0 * W[c - 2]) - (45.0 * W[c - 1]) + (45.0 * W[c + 1]) - (9.0 * W[c + 2]) + W[c + 3]) / (60.0 * dx);
            double fder = (((U[c + 2 * sk] - U[c]) * (df1 * df1)) + ((U[c] - U[c + sk]) * (df2 * df2))) /
                          ((df2 * (df1 * df1)) - ((df2 * df2) * df1));
            t = t + ((mfb1 + fder) / g->Re);
        }
        tau[i - 1] = t;
    }
}
