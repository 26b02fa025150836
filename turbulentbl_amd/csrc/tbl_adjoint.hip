// tbl_adjoint.hip -- discrete adjoint of the TBL time step (FP64), and the
// direct Helmholtz solve that the forward steps of an adjoint window use.
//
// Inside a window that crosses no statistics rescale (dnssetbl.f90:986-1220)
// the recycling inflow is a constant overwrite of the i = 1 plane, and one
// step maps (Ut, Press) to (Ut, Press) through
//   Q = U - dt (conv(U) + Gp(P))                 (dnssetbl.f90:306-369)
//   U* = H^{-1} (Q + h B(U faces)), H = I - hA   (:371-433, solved exactly)
//   phi = A^{-1} D(U*)/dt                        (:437-671, solved exactly)
//   phi extended to the faces, P += phi, U = U* - dt G(phi)   (:673-726)
//   velocity and pressure boundary conditions    (:728-797)
// The kernels below apply the transposes of these blocks in reverse order.
// Every kernel gathers: one thread owns one output value and sums the
// contributions of the forward outputs that read it, so no two threads write
// the same address and the result is deterministic.  Indices are 0-based;
// the interior is i in [1,nX-2], j in [1,nY-2], k in [1,nZ-2].
#include "tbl_dev.h"
#include "tbl_launch.h"

namespace tbl {

__device__ __forceinline__ bool inner(const Geo &g, int i, int j, int k)
{
    return i >= 1 && i <= g.nX - 2 && j >= 1 && j <= g.nY - 2 && k >= 1 && k <= g.nZ - 2;
}
__device__ __forceinline__ long lin(const Geo &g, int i, int j, int k)
{
    return (long)i + (long)g.nX * ((long)j + (long)g.nY * k);
}
__device__ __forceinline__ double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

// all-points 3-D launch: 64 x 4 (i,j) tiles, one k plane per block row
static inline dim3 all3(const Geo &g) { return dim3((g.nX + 63) / 64, (g.nY + 3) / 4, g.nZ); }
#define ALL3_IDX                                                   \
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);            \
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);             \
    const int k = blockIdx.z;                                      \
    if (i >= g.nX || j >= g.nY) return;

// ============================================================================
// direct Helmholtz solve (forward): residual field and interior update
// ============================================================================
// r = Q - H u on the interior, in the reference's residual form
// (dnssetbl.f90:378-386) evaluated in double; faces of r are left untouched
template <class T>
__global__ __launch_bounds__(256) void k_helm_rfield(Geo g, const T *__restrict__ Q, const T *__restrict__ u,
                                                    T *__restrict__ r)
{
    ALL3_IDX
    if (!inner(g, i, j, k)) return;
    const long c = lin(g, i, j, k), sk = g.sZ;
    const long cp = lin(g, i, jp1_of(j, g.nY), k), cm = lin(g, i, jm1_of(j, g.nY), k);
    const double rt1 = g.Ae * (double)u[c + 1] + g.Aw * (double)u[c - 1] + g.An * (double)u[cp] + g.As * (double)u[cm];
    const double rt2 = g.At[k] * (double)u[c + sk] + g.Ab[k] * (double)u[c - sk];
    r[c] = (T)((double)Q[c] - (1.0 - g.dt / g.Re * g.Ap[k]) * (double)u[c] + g.dt * (rt1 + rt2) / g.Re);
}
template <class T, class Td>
__global__ __launch_bounds__(256) void k_add_interior(Geo g, T *__restrict__ u, const Td *__restrict__ d)
{
    ALL3_IDX
    if (!inner(g, i, j, k)) return;
    const long c = lin(g, i, j, k);
    u[c] = u[c] + (T)d[c];
}

// ============================================================================
// transposed operators for the adjoint solves
// ============================================================================
// r = b - (alpha x + beta A^T x) on the interior, formed in FP64 and stored
// in To (FP32 for the refinement passes of a mixed context); x is zero off
// the interior.
// A^T x (c) = Ap(k) x(c) + Aw x(c+1) + Ae x(c-1) + As x(jp1) + An x(jm1)
//           + Ab(k+1) x(c+sk) + At(k-1) x(c-sk)
template <class To>
__global__ __launch_bounds__(256) void k_opT_res(Geo g, const double *__restrict__ b, const double *__restrict__ x,
                                                 double alpha, double beta, To *__restrict__ r)
{
    ALL3_IDX
    if (!inner(g, i, j, k)) return;
    const long c = lin(g, i, j, k), sk = g.sZ;
    const long cp = lin(g, i, jp1_of(j, g.nY), k), cm = lin(g, i, jm1_of(j, g.nY), k);
    const double ax = g.Ap[k] * x[c] + g.Aw * x[c + 1] + g.Ae * x[c - 1] + g.As * x[cp] + g.An * x[cm] +
                      g.Ab[k + 1] * x[c + sk] + g.At[k - 1] * x[c - sk];
    r[c] = (To)(b[c] - (alpha * x[c] + beta * ax));
}
// out = in on the interior, 0 elsewhere
__global__ __launch_bounds__(256) void k_copy_interior(Geo g, const double *__restrict__ in, double *__restrict__ out)
{
    ALL3_IDX
    const long c = lin(g, i, j, k);
    out[c] = inner(g, i, j, k) ? in[c] : 0.0;
}

// ============================================================================
// objective: J = cJ * sum_{i in [i0,i1]} tau(i), tau of dnstbl/dnstbl.f90:1620-1668
// ============================================================================
// lamU += dJ/dU.  Thread per wall point (i, j), i in [0,nX-1], j in [1,nY-2].
__global__ void k_adj_tau(Geo g, double *__restrict__ lamU, int i0, int i1, double cJ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1;
    if (i >= g.nX) return;
    const long N = g.N, sk = g.sZ, c = lin(g, i, j, 0);
    const double s = cJ / g.Re, df1 = g.df1, df2 = g.df2;
    double *U = lamU, *W = lamU + 2 * N;
    if (i >= i0 && i <= i1) {
        // fder = ((U(k=2) - U(k=0)) df1^2 + (U(k=0) - U(k=1)) df2^2) / (df2 df1^2 - df2^2 df1)
        const double den = (df2 * (df1 * df1)) - ((df2 * df2) * df1);
        U[c] += s * ((df2 * df2) - (df1 * df1)) / den;
        U[c + sk] += s * (-(df2 * df2)) / den;
        U[c + 2 * sk] += s * (df1 * df1) / den;
    }
    // dw/dx at the wall of the stations within 3 of i
    double a = 0.0;
    for (int o = -3; o <= 3; o++) {
        const int st = i - o;                       // station st reads W(st + o) = W(i)
        if (o == 0 || st < i0 || st > i1) continue;
        if (st <= 2 || st >= g.nX - 3) {
            if (o == 1) a += 1.0 / (2.0 * g.dx);
            else if (o == -1) a -= 1.0 / (2.0 * g.dx);
        } else {
            const double w6 = (o == -3) ? -1.0 : (o == -2) ? 9.0 : (o == -1) ? -45.0 : (o == 1) ? 45.0 : (o == 2) ? -9.0 : 1.0;
            a += w6 / (60.0 * g.dx);
        }
    }
    W[c] += s * a;
}

// ============================================================================
// pressure boundary conditions, reversed (dnssetbl.f90:738-797)
// ============================================================================
// span copies P(i,0,k) = P(i,nY-2,k), P(i,nY-1,k) = P(i,1,k); thread per (i,k)
__global__ void k_adj_span(Geo g, double *__restrict__ lam)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= g.nX) return;
    const long g0 = lin(g, i, 0, k), g1 = lin(g, i, g.nY - 1, k);
    lam[lin(g, i, g.nY - 2, k)] += lam[g0];
    lam[lin(g, i, 1, k)] += lam[g1];
    lam[g0] = 0.0;
    lam[g1] = 0.0;
}
// inflow pressure P(0,j,k) (:773-794); with the inflow plane fixed it depends
// on U(1), U(2), P(1), P(2) only.  Thread per interior (j, k).
__global__ void k_adj_pin(Geo g, const double *__restrict__ Uend, double *__restrict__ lamU, double *__restrict__ lamP)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y + 1;
    if (j > g.nY - 2 || k > g.nZ - 2) return;
    const long c = lin(g, 0, j, k);
    const double L = lamP[c];
    lamP[c] = 0.0;
    const double dx = g.dx, Re = g.Re, u1 = Uend[c];
    const double f = 2.0 * dx / 3.0 * L;
    lamU[c + 1] += f * (u1 / dx + 2.0 / (dx * dx * Re));
    lamU[c + 2] += f * (-1.0 / (dx * dx * Re));
    lamP[c + 1] += 4.0 * L / 3.0;
    lamP[c + 2] += -L / 3.0;
}
// outflow pressure P(nX-1,j,k) = (3U(nX-1) - 4U(nX-2) + U(nX-3)) / (2 dx Re) (:767-771)
__global__ void k_adj_pout(Geo g, double *__restrict__ lamU, double *__restrict__ lamP)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y + 1;
    if (j > g.nY - 2 || k > g.nZ - 2) return;
    const long c = lin(g, g.nX - 1, j, k);
    const double f = lamP[c] / (2.0 * g.dx * g.Re);
    lamP[c] = 0.0;
    lamU[c] += 3.0 * f;
    lamU[c - 1] += -4.0 * f;
    lamU[c - 2] += f;
}
// top pressure P(i,j,nZ-1) = (11W(nZ-1) - 18W(nZ-2) + 9W(nZ-3) - 2W(nZ-4)) JacInv(nZ-1) / (6 dXi Re) (:757-765)
__global__ void k_adj_ptop(Geo g, double *__restrict__ lamW, double *__restrict__ lamP)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
    if (i > g.nX - 2) return;
    const int kt = g.nZ - 1;
    const long c = lin(g, i, j, kt), sk = g.sZ;
    const double f = lamP[c] * g.JacInv[kt] / (6.0 * g.dXi * g.Re);
    lamP[c] = 0.0;
    lamW[c] += 11.0 * f;
    lamW[c - sk] += -18.0 * f;
    lamW[c - 2 * sk] += 9.0 * f;
    lamW[c - 3 * sk] += -2.0 * f;
}
// bottom pressure (:738-755): P(i,j,0) = P(i,j,1) - rt12 df1,
//   rt12 = -w1 (w2 - w1)/df1 + (rt1 + rt2 + rt3)/Re,
//   rt1 = (W(i-1,j,0) - 2 w1 + W(i+1,j,0))/dx^2, rt2 likewise in raw j +- 1,
//   rt3 = (2 df1 w3 - 2 df2 w2 + 2 (df2 - df1) w1)/(df1 df2^2 - df2 df1^2).
// (1) gather into lamW(.,.,0) over every wall point incl. span ghosts; reads
// the targets' lamP(.,.,0) without changing it
__global__ void k_adj_pbot_w0(Geo g, const double *__restrict__ Uend, const double *__restrict__ lamP,
                              double *__restrict__ lamW)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= g.nX || j >= g.nY) return;
    const double df1 = g.df1, df2 = g.df2, Re = g.Re, dx = g.dx, dy = g.dy;
    auto tgt = [&](int ti, int tj) { return ti >= 1 && ti <= g.nX - 2 && tj >= 1 && tj <= g.nY - 2; };
    double acc = 0.0;
    if (tgt(i - 1, j)) acc += lamP[lin(g, i - 1, j, 0)] / (dx * dx * Re);
    if (tgt(i + 1, j)) acc += lamP[lin(g, i + 1, j, 0)] / (dx * dx * Re);
    if (tgt(i, j - 1)) acc += lamP[lin(g, i, j - 1, 0)] / (dy * dy * Re);
    if (tgt(i, j + 1)) acc += lamP[lin(g, i, j + 1, 0)] / (dy * dy * Re);
    acc = -df1 * acc;
    if (tgt(i, j)) {
        const long c = lin(g, i, j, 0), sk = g.sZ, N = g.N;
        const double w1 = Uend[2 * N + c], w2 = Uend[2 * N + c + sk];
        const double den = (df1 * (df2 * df2) - (df2 * (df1 * df1)));
        const double d = -(w2 - 2.0 * w1) / df1 + (-2.0 / (dx * dx) - 2.0 / (dy * dy) + 2.0 * (df2 - df1) / den) / Re;
        acc += -df1 * lamP[c] * d;
    }
    lamW[lin(g, i, j, 0)] += acc;
}
// (2) per target: P(i,j,1) and W(i,j,1..2) terms, then the target is consumed
__global__ void k_adj_pbot_loc(Geo g, const double *__restrict__ Uend, double *__restrict__ lamP,
                               double *__restrict__ lamW)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1;
    if (i > g.nX - 2) return;
    const double df1 = g.df1, df2 = g.df2, Re = g.Re;
    const long c = lin(g, i, j, 0), sk = g.sZ, N = g.N;
    const double L = lamP[c];
    lamP[c] = 0.0;
    lamP[c + sk] += L;
    const double w1 = Uend[2 * N + c];
    const double den = (df1 * (df2 * df2) - (df2 * (df1 * df1)));
    lamW[c + sk] += -df1 * L * (-w1 / df1 + (-2.0 * df2 / den) / Re);
    lamW[c + 2 * sk] += -df1 * L * ((2.0 * df1 / den) / Re);
}

// ============================================================================
// velocity boundary conditions, reversed (dnssetbl.f90:729-736), one component
// ============================================================================
// outflow u(nX-1,j,k) = (4u(nX-2) - u(nX-3))/3, k in [0,nZ-1]; thread per (j,k)
__global__ void k_adj_vout(Geo g, double *__restrict__ lam)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y;
    if (j > g.nY - 2) return;
    const long c = lin(g, g.nX - 1, j, k);
    const double L = lam[c];
    lam[c] = 0.0;
    lam[c - 1] += 4.0 * L / 3.0;
    lam[c - 2] += -L / 3.0;
}
// top u(i,j,nZ-1) = (4u(nZ-2) - u(nZ-3))/3, then the wall value u(i,j,0)
// (w = wall_w, the control: its adjoint is the gradient); thread per (i,j)
__global__ void k_adj_vtop_wall(Geo g, double *__restrict__ lam, double *__restrict__ grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1;
    if (i >= g.nX) return;
    const long sk = g.sZ, ct = lin(g, i, j, g.nZ - 1), cw = lin(g, i, j, 0);
    const double L = lam[ct];
    lam[ct] = 0.0;
    lam[ct - sk] += 4.0 * L / 3.0;
    lam[ct - 2 * sk] += -L / 3.0;
    if (grad) grad[(long)i + (long)g.nX * j] += lam[cw];
    lam[cw] = 0.0;
}

// ============================================================================
// projection + pressure update, reversed (dnssetbl.f90:697-726)
// ============================================================================
// lamPhi(s) = [s interior] lamP(s) - dt (G^T lamU)(s), every point s
__global__ __launch_bounds__(256) void k_adj_proj(Geo g, const double *__restrict__ lamU, const double *__restrict__ lamP,
                                                 double *__restrict__ lamPhi)
{
    ALL3_IDX
    const long c = lin(g, i, j, k), N = g.N, sk = g.sZ;
    const double *lu = lamU, *lv = lamU + N, *lw = lamU + 2 * N;
    double gt = 0.0;
    if (j >= 1 && j <= g.nY - 2 && k >= 1 && k <= g.nZ - 2) {
        if (i - 1 >= 1 && i - 1 <= g.nX - 2) gt += 0.5 * g.invdx * lu[c - 1];
        if (i + 1 >= 1 && i + 1 <= g.nX - 2) gt -= 0.5 * g.invdx * lu[c + 1];
    }
    if (i >= 1 && i <= g.nX - 2 && k >= 1 && k <= g.nZ - 2 && j >= 1 && j <= g.nY - 2) {
        gt += 0.5 * g.invdy * (lv[lin(g, i, jm1_of(j, g.nY), k)] - lv[lin(g, i, jp1_of(j, g.nY), k)]);
    }
    if (i >= 1 && i <= g.nX - 2 && j >= 1 && j <= g.nY - 2) {
        if (k - 1 >= 1 && k - 1 <= g.nZ - 2) gt += 0.5 * g.invdXi * g.JacInv[k - 1] * lw[c - sk];
        if (k + 1 >= 1 && k + 1 <= g.nZ - 2) gt -= 0.5 * g.invdXi * g.JacInv[k + 1] * lw[c + sk];
    }
    lamPhi[c] = (inner(g, i, j, k) ? lamP[c] : 0.0) - g.dt * gt;
}
// face extrapolations of PCorr, reversed: x faces (k in [0,nZ-1]) then z faces
__global__ void k_adj_phi_x(Geo g, double *__restrict__ lp)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, k = blockIdx.y;
    if (j > g.nY - 2) return;
    const long c0 = lin(g, 0, j, k), cn = lin(g, g.nX - 1, j, k);
    const double L0 = lp[c0];
    lp[c0] = 0.0;
    lp[c0 + 1] += 4.0 * L0 / 3.0;
    lp[c0 + 2] += -L0 / 3.0;
    const double Ln = lp[cn];
    lp[cn] = 0.0;
    lp[cn - 1] += 5.0 * Ln / 2.0;
    lp[cn - 2] += -2.0 * Ln;
    lp[cn - 3] += Ln / 2.0;
}
__global__ void k_adj_phi_z(Geo g, double *__restrict__ lp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1;
    if (i >= g.nX) return;
    const long sk = g.sZ, c0 = lin(g, i, j, 0), cn = lin(g, i, j, g.nZ - 1);
    const double L0 = lp[c0];
    lp[c0] = 0.0;
    lp[c0 + sk] += 4.0 * L0 / 3.0;
    lp[c0 + 2 * sk] += -L0 / 3.0;
    const double Ln = lp[cn];
    lp[cn] = 0.0;
    lp[cn - sk] += 5.0 * Ln / 2.0;
    lp[cn - 2 * sk] += -2.0 * Ln;
    lp[cn - 3 * sk] += Ln / 2.0;
}

// ============================================================================
// PPE right-hand side, reversed (dnssetbl.f90:437-447): lamU += D^T lamB / dt
// ============================================================================
__global__ __launch_bounds__(256) void k_adj_div(Geo g, const double *__restrict__ lamB, double *__restrict__ lamU)
{
    ALL3_IDX
    const long c = lin(g, i, j, k), N = g.N, sk = g.sZ;
    const double rdt = 1.0 / g.dt;
    if (j >= 1 && j <= g.nY - 2 && k >= 1 && k <= g.nZ - 2) {
        double a = 0.0;
        if (i - 1 >= 1 && i - 1 <= g.nX - 2) a += lamB[c - 1];
        if (i + 1 >= 1 && i + 1 <= g.nX - 2) a -= lamB[c + 1];
        lamU[c] += 0.5 * g.invdx * rdt * a;
    }
    if (inner(g, i, j, k)) {
        const double a = lamB[lin(g, i, jm1_of(j, g.nY), k)] - lamB[lin(g, i, jp1_of(j, g.nY), k)];
        lamU[N + c] += 0.5 * g.invdy * rdt * a;
    }
    if (i >= 1 && i <= g.nX - 2 && j >= 1 && j <= g.nY - 2) {
        double a = 0.0;
        if (k - 1 >= 1 && k - 1 <= g.nZ - 2) a += lamB[c - sk] / (g.x3[k] - g.x3[k - 2]);
        if (k + 1 >= 1 && k + 1 <= g.nZ - 2) a -= lamB[c + sk] / (g.x3[k + 2] - g.x3[k]);
        lamU[2 * N + c] += rdt * a;
    }
}

// ============================================================================
// momentum RHS + Helmholtz boundary data + inflow, reversed
// (dnssetbl.f90:257-369,371-433)
// ============================================================================
// Coefficient with which the convection of a target t reads F at offset o of
// one direction, given the target's advecting velocity a, branch flags and
// grid spacing factor inv (invdx, invdy or invdXi).
__device__ __forceinline__ double conv_coef(int o, double a, bool edge, bool kk, double inv, double c12)
{
    if (edge) return (o == 1) ? a * 0.5 * inv : (o == -1) ? -a * 0.5 * inv : 0.0;
    const double wc = (o == -2) ? 1.0 : (o == -1) ? -8.0 : (o == 1) ? 8.0 : (o == 2) ? -1.0 : 0.0;
    double v = a * wc * (c12 * inv);
    if (kk) {
        const double w4 = (o == -2 || o == 2) ? 1.0 : (o == -1 || o == 1) ? -4.0 : 6.0;
        v += fabs(a) * w4 * (0.25 * inv);
    }
    return v;
}
// derivative of one direction's term with respect to the advecting velocity a,
// given the five stencil values f[-2..2] (f[0] at index 2)
__device__ __forceinline__ double conv_dadv(const double *f, double a, bool edge, bool kk, double inv, double c12)
{
    if (edge) return (f[3] - f[1]) * 0.5 * inv;
    double v = (f[0] - 8.0 * f[1] + 8.0 * f[3] - f[4]) * (c12 * inv);
    if (kk) v += sgn(a) * (f[4] - 4.0 * f[3] + 6.0 * f[2] - 4.0 * f[1] + f[0]) * (0.25 * inv);
    return v;
}

// Thread per point s.  U is the step's linearisation point (after the inflow
// overwrite), lamQ the adjoint of Q (interior, zero elsewhere), lamU the
// adjoint of the Helmholtz output on entry.  On exit lamU holds the adjoint of
// the step's input velocity; lamP receives -dt Gp^T lamQ.
__global__ __launch_bounds__(256) void k_adj_mom(Geo g, const double *__restrict__ U, const double *__restrict__ lamQ,
                                                double *__restrict__ lamU, double *__restrict__ lamP)
{
    ALL3_IDX
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const long N = g.N, sk = g.sZ, c = lin(g, i, j, k);
    const double dt = g.dt, Re = g.Re, c12 = g.c12, h = g.dt / g.Re;
    const double *Uu = U, *Uv = U + N, *Uw = U + 2 * N;
    const bool si = inner(g, i, j, k);
    double out[3];
    // base: the Helmholtz solve overwrote the interior (its input adjoint is
    // lamQ through Q's U term); faces pass through and feed the Helmholtz
    // right-hand side h B u(face)
    for (int p = 0; p < 3; p++) {
        const double *lq = lamQ + p * N;
        if (si) {
            out[p] = lq[c];
        } else {
            double o = lamU[p * N + c];
            if (j >= 1 && j <= nY - 2 && k >= 1 && k <= nZ - 2) {
                if (i == 0 && nX - 2 >= 1) o += h * g.Aw * lq[c + 1];
                if (i == nX - 1) o += h * g.Ae * lq[c - 1];
            }
            if (i >= 1 && i <= nX - 2 && j >= 1 && j <= nY - 2) {
                if (k == 0) o += h * g.Ab[1] * lq[c + sk];
                if (k == nZ - 1) o += h * g.At[nZ - 2] * lq[c - sk];
            }
            out[p] = o;
        }
    }
    // stencil part of the convection: targets t in the x, y, z stencils of s
    double gp = 0.0;                  // -dt Gp^T lamQ at s
    // x direction
    if (j >= 1 && j <= nY - 2 && k >= 1 && k <= nZ - 2) {
        for (int o = -2; o <= 2; o++) {
            const int ti = i - o;     // target reads F(ti + o) = F(s)
            if (ti < 1 || ti > nX - 2) continue;
            const long t = c - o;
            const double a = Uu[t];
            const bool edge = (ti <= 1 || ti >= nX - 2);
            const bool kk = fabs(Re * a * g.dx) > 2.0;
            const double cf = conv_coef(o, a, edge, kk, g.invdx, c12);
            for (int p = 0; p < 3; p++) out[p] += -dt * cf * lamQ[p * N + t];
        }
        if (i - 1 >= 1 && i - 1 <= nX - 2) gp += 0.5 * g.invdx * lamQ[c - 1];
        if (i + 1 >= 1 && i + 1 <= nX - 2) gp -= 0.5 * g.invdx * lamQ[c + 1];
    }
    // y direction (periodic tables: the target whose jp_o neighbour is j)
    if (si) {
        const int tj[5] = {jp2_of(j, nY), jp1_of(j, nY), j, jm1_of(j, nY), jm2_of(j, nY)};   // o = -2..2
        for (int q = 0; q < 5; q++) {
            const int o = q - 2;
            const long t = lin(g, i, tj[q], k);
            const double a = Uv[t];
            const bool edge = (tj[q] == 1 || tj[q] == nY - 2);
            const bool kk = fabs(Re * a * g.dy) > 2.0;
            const double cf = conv_coef(o, a, edge, kk, g.invdy, c12);
            for (int p = 0; p < 3; p++) out[p] += -dt * cf * lamQ[p * N + t];
        }
        gp += 0.5 * g.invdy * (lamQ[N + lin(g, i, jm1_of(j, nY), k)] - lamQ[N + lin(g, i, jp1_of(j, nY), k)]);
    }
    // z direction (the term is scaled by JacInv of the target)
    if (i >= 1 && i <= nX - 2 && j >= 1 && j <= nY - 2) {
        for (int o = -2; o <= 2; o++) {
            const int tk = k - o;
            if (tk < 1 || tk > nZ - 2) continue;
            const long t = c - (long)o * sk;
            const double a = Uw[t];
            const bool edge = (tk <= 1 || tk >= nZ - 2);
            const bool kk = fabs(Re * a * (g.x3[tk + 1] - g.x3[tk])) > 2.0;
            const double cf = conv_coef(o, a, edge, kk, g.invdXi, c12) * g.JacInv[tk];
            for (int p = 0; p < 3; p++) out[p] += -dt * cf * lamQ[p * N + t];
        }
        if (k - 1 >= 1 && k - 1 <= nZ - 2) gp += 0.5 * g.invdXi * g.JacInv[k - 1] * lamQ[2 * N + c - sk];
        if (k + 1 >= 1 && k + 1 <= nZ - 2) gp -= 0.5 * g.invdXi * g.JacInv[k + 1] * lamQ[2 * N + c + sk];
    }
    // advecting-velocity part at s itself
    if (si) {
        const double a = Uu[c], b = Uv[c], w = Uw[c];
        const bool xe = (i <= 1 || i >= nX - 2), ye = (j == 1 || j == nY - 2), ze = (k <= 1 || k >= nZ - 2);
        const bool xk = fabs(Re * a * g.dx) > 2.0, yk = fabs(Re * b * g.dy) > 2.0;
        const bool zk = fabs(Re * w * (g.x3[k + 1] - g.x3[k])) > 2.0;
        const long cy[5] = {lin(g, i, jm2_of(j, nY), k), lin(g, i, jm1_of(j, nY), k), c, lin(g, i, jp1_of(j, nY), k),
                            lin(g, i, jp2_of(j, nY), k)};
        double du = 0.0, dv = 0.0, dw = 0.0;
        for (int p = 0; p < 3; p++) {
            const double *F = U + p * N;
            const double ap = -dt * lamQ[p * N + c];
            double f[5];
            for (int q = 0; q < 5; q++) f[q] = (xe && (q == 0 || q == 4)) ? 0.0 : F[c + (q - 2)];
            du += ap * conv_dadv(f, a, xe, xk, g.invdx, c12);
            for (int q = 0; q < 5; q++) f[q] = (ye && (q == 0 || q == 4)) ? 0.0 : F[cy[q]];
            dv += ap * conv_dadv(f, b, ye, yk, g.invdy, c12);
            for (int q = 0; q < 5; q++) f[q] = (ze && (q == 0 || q == 4)) ? 0.0 : F[c + (long)(q - 2) * sk];
            dw += ap * conv_dadv(f, w, ze, zk, g.invdXi, c12) * g.JacInv[k];
        }
        out[0] += du;
        out[1] += dv;
        out[2] += dw;
    }
    // the inflow plane is overwritten by the (frozen) recycling inflow
    if (i == 0 && j >= 1 && j <= nY - 2 && k >= 1 && k <= nZ - 2) out[0] = out[1] = out[2] = 0.0;
    for (int p = 0; p < 3; p++) lamU[p * N + c] = out[p];
    lamP[c] += -dt * gp;
}

// J += cJ * sum_{i in [i0,i1]} tau(i), one thread, fixed order
__global__ void k_tau_acc(const double *__restrict__ tau, int i0, int i1, double cJ, double *__restrict__ J)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double a = 0.0;
    for (int i = i0; i <= i1; i++) a += tau[i];
    J[0] = J[0] + cJ * a;
}

// ============================================================================
// host launchers
// ============================================================================
void launch_tau_acc(const Geo &g, const double *tau, int i0, int i1, double cJ, double *J, hipStream_t s)
{
    hipLaunchKernelGGL(k_tau_acc, dim3(1), dim3(64), 0, s, tau, i0, i1, cJ, J);
}
template <class T>
void launch_helm_rfield(const Geo &g, const T *Q, const T *u, T *r, hipStream_t s)
{
    hipLaunchKernelGGL(k_helm_rfield<T>, all3(g), dim3(256), 0, s, g, Q, u, r);
}
template <class T, class Td>
void launch_add_interior(const Geo &g, T *u, const Td *d, hipStream_t s)
{
    hipLaunchKernelGGL((k_add_interior<T, Td>), all3(g), dim3(256), 0, s, g, u, d);
}
template void launch_helm_rfield<double>(const Geo &, const double *, const double *, double *, hipStream_t);
template void launch_helm_rfield<float>(const Geo &, const float *, const float *, float *, hipStream_t);
template void launch_add_interior<double, double>(const Geo &, double *, const double *, hipStream_t);
template void launch_add_interior<float, float>(const Geo &, float *, const float *, hipStream_t);
template void launch_add_interior<double, float>(const Geo &, double *, const float *, hipStream_t);

void launch_opT_res(const Geo &g, const double *b, const double *x, double alpha, double beta, double *r, hipStream_t s)
{
    hipLaunchKernelGGL(k_opT_res<double>, all3(g), dim3(256), 0, s, g, b, x, alpha, beta, r);
}
void launch_opT_res(const Geo &g, const double *b, const double *x, double alpha, double beta, float *r, hipStream_t s)
{
    hipLaunchKernelGGL(k_opT_res<float>, all3(g), dim3(256), 0, s, g, b, x, alpha, beta, r);
}
void launch_copy_interior(const Geo &g, const double *in, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_interior, all3(g), dim3(256), 0, s, g, in, out);
}
void launch_adj_tau(const Geo &g, double *lamU, int i0, int i1, double cJ, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_tau, dim3((g.nX + 255) / 256, g.nY - 2), dim3(256), 0, s, g, lamU, i0, i1, cJ);
}
// pressure then velocity boundary conditions, reversed; grad += adjoint of wall_w
void launch_adj_bcs(const Geo &g, const double *Uend, double *lamU, double *lamP, double *grad, hipStream_t s)
{
    const long N = g.N;
    double *lamW = lamU + 2 * N;
    const dim3 b256(256);
    hipLaunchKernelGGL(k_adj_span, dim3((g.nX + 255) / 256, g.nZ), b256, 0, s, g, lamP);
    hipLaunchKernelGGL(k_adj_pin, dim3((g.nY - 2 + 255) / 256, g.nZ - 2), b256, 0, s, g, Uend, lamU, lamP);
    hipLaunchKernelGGL(k_adj_pout, dim3((g.nY - 2 + 255) / 256, g.nZ - 2), b256, 0, s, g, lamU, lamP);
    hipLaunchKernelGGL(k_adj_ptop, dim3((g.nX - 2 + 255) / 256, g.nY - 2), b256, 0, s, g, lamW, lamP);
    hipLaunchKernelGGL(k_adj_pbot_w0, dim3((g.nX + 255) / 256, g.nY), b256, 0, s, g, Uend, (const double *)lamP, lamW);
    hipLaunchKernelGGL(k_adj_pbot_loc, dim3((g.nX - 2 + 255) / 256, g.nY - 2), b256, 0, s, g, Uend, lamP, lamW);
    for (int p = 0; p < 3; p++) {
        double *l = lamU + p * N;
        hipLaunchKernelGGL(k_adj_span, dim3((g.nX + 255) / 256, g.nZ), b256, 0, s, g, l);
        hipLaunchKernelGGL(k_adj_vout, dim3((g.nY - 2 + 255) / 256, g.nZ), b256, 0, s, g, l);
        hipLaunchKernelGGL(k_adj_vtop_wall, dim3((g.nX + 255) / 256, g.nY - 2), b256, 0, s, g, l, p == 2 ? grad : nullptr);
    }
}
void launch_adj_proj(const Geo &g, const double *lamU, const double *lamP, double *lamPhi, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_proj, all3(g), dim3(256), 0, s, g, lamU, lamP, lamPhi);
    hipLaunchKernelGGL(k_adj_phi_x, dim3((g.nY - 2 + 255) / 256, g.nZ), dim3(256), 0, s, g, lamPhi);
    hipLaunchKernelGGL(k_adj_phi_z, dim3((g.nX + 255) / 256, g.nY - 2), dim3(256), 0, s, g, lamPhi);
}
void launch_adj_div(const Geo &g, const double *lamB, double *lamU, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_div, all3(g), dim3(256), 0, s, g, lamB, lamU);
}
void launch_adj_mom(const Geo &g, const double *U, const double *lamQ, double *lamU, double *lamP, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_mom, all3(g), dim3(256), 0, s, g, U, lamQ, lamU, lamP);
}

}  // namespace tbl
