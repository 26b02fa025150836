// tbl_launch.h -- host-side launchers of the device kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "tbl_dev.h"

namespace tbl {

enum { PC_SSOR_RB = 1, PC_NONE = 2, PC_JACOBI = 3, PC_FAST = 4, PC_SSOR_BLOCK = 5 };
struct FastPC;

// stencil kernels (tbl_stencil.hip)
unsigned resid_blocks(const Geo &g);
unsigned rhs_blocks(const Geo &g, bool single);
// hpart (optional): the momentum kernel also forms the Helmholtz solve's first
// residual pass (block partials in k_helm_resid's layout, rhs_blocks of them)
template <class T> void launch_momentum_rhs(const Geo &g, const T *Ut, const T *P, T *Q, hipStream_t s,
                                            const unsigned char *brin = nullptr, unsigned char *brout = nullptr,
                                            double *hpart = nullptr);
template <class T> void launch_helm_resid(const Geo &g, const T *Q, const T *Ut, const void *ctrl, double *part, int first, hipStream_t s);
void launch_helm_fin(const Geo &g, void *ctrl, const double *part, int it, double errorstop, double floor_coef, hipStream_t s,
                     int nb = 0);
template <class T> void launch_helm_sweep(const Geo &g, const T *Q, T *Ut, const void *ctrl, int colour, hipStream_t s);
template <class T> void launch_helm_zline(const Geo &g, const T *Q, T *Ut, const T *zf, const void *ctrl, int colour, hipStream_t s);
// one pass of the z-line zebra sweep: x neighbours from cur, span rows from
// ysrc, results to dst (see k_helm_zebra)
template <class T> void launch_helm_zebra(const Geo &g, const T *Q, const T *cur, const T *ysrc, T *dst, const T *zf,
                                          const void *ctrl, int colour, hipStream_t s);
// boundary / ghost values of the three velocity components, src -> dst
template <class T> void launch_copy_faces3(const Geo &g, const T *src, T *dst, hipStream_t s);
// per-slab lexicographic Gauss-Seidel sweep (snapshot of the slabs' ghost
// planes into halo, then the sweep); threads a slab kernel needs, 0 if the
// slabs are too thick for one workgroup
template <class T> void launch_helm_lex(const Geo &g, const T *Q, T *Ut, T *halo, const void *ctrl, int nkmax,
                                        hipStream_t s);
int slab_threads(const Geo &g, int nkmax);
size_t gs_ctrl_bytes();
void gs_ctrl_init(void *host, int active_mask);
void gs_ctrl_read(const void *host, int *active, int *sweeps, double *res);
// gate (optional): the Gauss-Seidel control block; the launch does nothing
// while any of its components is active
template <class T> void launch_ppe_rhs(const Geo &g, const T *Ut, double *pRhs, hipStream_t s, const void *gate = nullptr);
unsigned poisson_blocks(const Geo &g);   // partials written by launch_poisson with part
// out = A x (scaled); with part: also the block partials of <out, v0> (returns the block count)
template <class T>
unsigned launch_poisson(const Geo &g, const T *x, T *out, const int *active, const double *scale, hipStream_t s,
                        const T *v0 = nullptr, double *part = nullptr);
// bpart (optional): block partials of ||pRhs||^2, as many as part
template <class Tr> void launch_resid(const Geo &g, const double *pRhs, const double *PC, Tr *r, double *part, hipStream_t s,
                                      double *bpart = nullptr);
template <class Tr>
// gated: does nothing while gc's cycle is active (queued ahead of a host poll)
void launch_resid_upd(const Geo &g, const double *pRhs, const double *PC, const Tr *Z, const GmCtrl *gc, double *PCn,
                      Tr *r, double *part, hipStream_t s, int gated = 0);
// part (FPC_*) applies to the fast preconditioner only
template <class T> void launch_precond(const Geo &g, int pc, const FastPC *fast, const T *r, T *scratch, T *out, const int *active, hipStream_t s,
                                       int part = 0);
template <class T> void launch_zero_faces(const Geo &g, T *x, hipStream_t s);
template <class T> void launch_save_inflow(const Geo &g, const T *Ut, T *Utold, hipStream_t s);
template <class T> void launch_project(const Geo &g, double *PC, T *Press, T *Ut, hipStream_t s);
template <class T> void launch_bcs(const Geo &g, T *Ut, const T *Utold, T *Press, const double *wall_w, hipStream_t s);
template <class T> void launch_recycle(const Geo &g, T *Ut, int i_rec0, double eps, const double *lamda_dev,
                                       const double *avg, const double *Wt, const int *newk, const double *Ume,
                                       const double *uf, hipStream_t s);

// fast separable Poisson preconditioner (tbl_fastpc.hip)
struct FastPC;
int fastpc_create(const Geo &g, const double *Ab, const double *At, const double *Ap, hipStream_t s, bool single,
                  FastPC **out, char *err, int errlen);
void fastpc_destroy(FastPC *f);
// parts of one application: all of it, its first transform, or the rest
enum { FPC_ALL = 0, FPC_HEAD = 1, FPC_TAIL = 2 };
template <class T> int run_fastpc(const FastPC *f, const Geo &g, const T *r, T *out, const int *active, hipStream_t s,
                                  int part = FPC_ALL);
// out = (alpha I + beta A)^{-1} r, or its transpose (double only)
template <class T> int run_fastpc_ext(const FastPC *f, const Geo &g, const T *r, T *out, const int *active, double alpha,
                                      double beta, bool transpose, hipStream_t s, int part = FPC_ALL);

// Krylov kernels (tbl_krylov.hip)
unsigned vec_blocks(long N);
// x[0..n) = v (a kernel, not a memset: safe inside captured graphs)
void launch_fill(double *x, double v, long n, hipStream_t s);
template <class T> void launch_norm2(const T *x, long N, double *part, const int *active, hipStream_t s);
void launch_fin_start(GmCtrl *c, const double *part_r, int nb_r, const double *part_z, int nb_z, int first, hipStream_t s);
void launch_gm_init(GmCtrl *c, double errtol, double inner_tol, int kmax, int mmax, int mode, hipStream_t s);
void launch_fin_bnorm(GmCtrl *c, const double *part, int nb, hipStream_t s);
template <class T> void launch_dots(const T *V, long ldv, int j, long N, const GmCtrl *c, double *part, hipStream_t s);
void launch_fin_dots(GmCtrl *c, const double *part, int nb, int j, hipStream_t s);
template <class T> void launch_update(const T *V, long ldv, int j, long N, GmCtrl *c, double *part, hipStream_t s);
void launch_fin_update(GmCtrl *c, const double *part, int nb, int j, hipStream_t s);
// gated: does nothing while the cycle is active (queued ahead of a host poll)
void launch_fin_cycle(GmCtrl *c, hipStream_t s, int gated = 0);
template <class T> void launch_xupd(const T *V, long ldv, long N, const GmCtrl *c, double *PC, hipStream_t s);
template <class Td, class Ts> void launch_convert(const Ts *src, Td *dst, long n, hipStream_t s);

// diagnostics (tbl_diag.hip)
struct RecDev {            // device-resident scalars of the recycling state
    double lamda, lamda03;
    int ksep;
};
template <class T> void launch_div_check(const Geo &g, const T *Ut, double *part, hipStream_t s);
void launch_fin_sum(const double *part, int nb, double *out, hipStream_t s);
// out2[0] = |sum of part| / nint, out2[1] = the sum
void launch_fin_div(const double *part, int nb, double nint, double *out2, hipStream_t s);
#define MF_BLK 64   // mass-flux partial blocks per plane
template <class T> void launch_mass_flux(const Geo &g, const T *Ut, const int *planes, double denom, double *out5,
                                         double *part, hipStream_t s);
template <class T> void launch_stats(const Geo &g, const T *Ut, double *sums, hipStream_t s);
template <class T> void launch_rescale(const Geo &g, int i_rec0, int nsamples, const T *Ut, double *sums,
                                       double *avg, RecDev *rec, double *Wt, int *newk, double *bltkns,
                                       double *deltastar, double *theta, double *uf, double *scan_scratch,
                                       hipStream_t s);
template <class T> void launch_ke(const Geo &g, int i_rec0, const T *Ut, const double *avg, double *part,
                                  int *guard, double *uf, hipStream_t s);
// acc (optional): the running window sum, acc(i) += tau(i)
template <class T> void launch_wall_shear(const Geo &g, const T *Ut, double *tau, hipStream_t s, double *acc = nullptr);
void launch_tau_window(const Geo &g, int window, double *acc, double *mean, double *cf, hipStream_t s);

// direct Helmholtz solve and the discrete adjoint (tbl_adjoint.hip)
template <class T> void launch_helm_rfield(const Geo &g, const T *Q, const T *u, T *r, hipStream_t s);
// u += d on the interior (d may be narrower than u: an FP32 correction of FP64 u)
template <class T, class Td = T> void launch_add_interior(const Geo &g, T *u, const Td *d, hipStream_t s);
void launch_opT_res(const Geo &g, const double *b, const double *x, double alpha, double beta, double *r, hipStream_t s);
void launch_opT_res(const Geo &g, const double *b, const double *x, double alpha, double beta, float *r, hipStream_t s);
void launch_copy_interior(const Geo &g, const double *in, double *out, hipStream_t s);
void launch_adj_tau(const Geo &g, double *lamU, int i0, int i1, double cJ, hipStream_t s);
void launch_adj_bcs(const Geo &g, const double *Uend, double *lamU, double *lamP, double *grad, hipStream_t s);
void launch_adj_proj(const Geo &g, const double *lamU, const double *lamP, double *lamPhi, hipStream_t s);
void launch_adj_div(const Geo &g, const double *lamB, double *lamU, hipStream_t s);
void launch_adj_mom(const Geo &g, const double *U, const double *lamQ, double *lamU, double *lamP, hipStream_t s);
void launch_pc_extrap(double *pc, double *prev, long N, hipStream_t s);
void launch_tau_acc(const Geo &g, const double *tau, int i0, int i1, double cJ, double *J, hipStream_t s);

}  // namespace tbl
