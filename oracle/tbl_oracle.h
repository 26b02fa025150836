/*
 * tbl_oracle.h -- CPU FP64 restatement of the reference's fractional-step
 * turbulent-boundary-layer time step (dnssetbl/dnssetbl.f90).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: it is called by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nothing in the shipped library (turbulentbl_amd/, include/) links or calls
 * it.  Every function cites the reference lines it follows.
 *
 * Arrays are Fortran-ordered exactly like the reference: Ut(nX,nY,nZ,3),
 * Press/PCorr/pRhs(nX,nY,nZ), stats (nX,nZ); i fastest; the macros below take
 * the reference's 1-based indices.
 */
#ifndef TBL_ORACLE_H
#define TBL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* most slabs (emulated MPI ranks) a grid may have */
#define ORC_MAX_SLABS 256
/* threads of the OpenMP loops.  Slabs run in parallel (per-slab Gauss-Seidel,
 * per-slab SSOR, per-slab partial sums) and stencil loops over k planes, so
 * results depend on the slab count P but never on the thread count. */
int orc_threads(void);

/* preconditioner choice for the Poisson GMRES */
enum { ORC_PC_SSOR_BLOCK = 0,  /* reference: per-slab lexicographic SSOR (dnssetbl.f90:509-532) */
       ORC_PC_SSOR_RB    = 1,  /* red-black ordered SSOR over the whole domain                 */
       ORC_PC_NONE       = 2,
       ORC_PC_JACOBI     = 3 };

/* Helmholtz Gauss-Seidel ordering */
enum { ORC_GS_LEX_BLOCK = 0,   /* reference: per-slab lexicographic, halo after each sweep (dnssetbl.f90:371-433) */
       ORC_GS_RB        = 1,   /* red-black (colour = parity of i+j+k) over the whole domain */
       ORC_GS_ZLINE_RB  = 2,   /* red-black over (i,j) columns (parity of i+j), each wall-normal
                                  column solved exactly (partitioned Thomas, orc_zline_factors) with
                                  x/y neighbours at current values */
       ORC_GS_ZEBRA     = 3 }; /* span-row zebra: pass c (c = 0, 1) solves every wall-normal column
                                  of the rows with parity(1-based j) == c exactly (same column solve
                                  as ORC_GS_ZLINE_RB), x neighbours at their values from the start
                                  of the pass, span neighbours (the other rows) at current values */

typedef struct orc_grid {
    /* sizes / parameters */
    int nX, nY, nZ;
    int P;                      /* number of k-slabs ("MPI ranks") emulated */
    double Lx, Ly, Lz, delta;   /* x1(nX), x2(nY), x3(nZ), grid condensation */
    double dt, Re, Ue;
    double omega;               /* SSOR relaxation (reference: REAL(4) 1.835) */
    int i_rec;                  /* recycling plane (reference: 68) */
    int freq_tim_avg;           /* reference: 500 */
    int planes[5];              /* mass-flux planes (reference: 5,20,40,60,85) */
    /* derived mesh (dnssetbl.f90:183-247) */
    double dx, dy, dXi, invdx, invdy, invdXi, onetwelfth;
    double Aw, Ae, As, An;
    double df1, df2;
    double *x1, *x2, *x3, *Xi, *Jac, *Jac2, *JacInv;
    double *Ab, *At, *Ap, *rAp;
    int *jm1, *jm2, *jp1, *jp2;         /* indexed by 1-based j */
    int *b, *e, *bn, *en;               /* slab bounds, indexed by rank */
    int scheme;                 /* 0: dnssetbl convection; 1: dnstbl (orc_grid_set_scheme) */
    double forc[3];             /* dnstbl body force */
    /* work arrays kept across calls, as the reference keeps v and rhat static
     * (:47,60-61): zero-initialised once, and only interior entries are ever
     * written, so their boundary entries stay zero */
    double *ws_v;               /* GMRES basis, ws_nv vectors of N */
    int ws_nv;
    double *ws_r, *ws_rh;       /* GMRES residual, preconditioner work (N each) */
} orc_grid;

orc_grid *orc_grid_create(int nX, int nY, int nZ, int P,
                          double Lx, double Ly, double Lz, double delta,
                          double dt, double Re, double Ue, double omega,
                          int i_rec, int freq_tim_avg, const int *planes);
void orc_grid_destroy(orc_grid *g);
/* convection scheme of orc_momentum_rhs: 0 = dnssetbl/dnssetbl.f90:306-369
 * (default), 1 = dnstbl/dnstbl.f90:620-766 with body force forc[3] */
void orc_grid_set_scheme(orc_grid *g, int scheme, const double *forc);
/* copy 1-D mesh arrays out: x1(nX) x2(nY) x3(nZ) Xi Jac Jac2 JacInv Ab At Ap rAp (nZ each) */
void orc_grid_mesh(const orc_grid *g, double *x1, double *x2, double *x3, double *Xi,
                   double *Jac, double *Jac2, double *JacInv,
                   double *Ab, double *At, double *Ap, double *rAp);
void orc_grid_scalars(const orc_grid *g, double *out /* dx dy dXi Aw Ae As An df1 df2 */);

/* ---- a2: recycling inflow (dnssetbl.f90:257-303) ---- */
void orc_recycle_inflow(const orc_grid *g, int nstep, double lamda,
                        const double *usum2, const double *vsum2, const double *wsum2,
                        const double *Wt, const int *newk, const double *Ume,
                        const double *Ufluc_rec /* (nY,nZ,3) at i_rec */,
                        double *Ut);

/* ---- a3+a4: convection + old-pressure gradient + Q (dnssetbl.f90:306-369) ---- */
void orc_momentum_rhs(const orc_grid *g, const double *Ut, const double *Press,
                      double *conv, double *Q);

/* ---- a5: implicit diffusion, Gauss-Seidel (dnssetbl.f90:371-433) ---- */
/* returns number of sweeps per component in sweeps[3], final L1 residual in res[3] */
void orc_diffusion_solve(const orc_grid *g, const double *Q, double *Ut,
                         int order, double errorstop, int mmax,
                         int *sweeps, double *res);
/* L1 residual of component p (1..3) as the reference computes it (dnssetbl.f90:378-386) */
double orc_helm_residual(const orc_grid *g, const double *Q, const double *Ut, int p);
/* one sweep of component p (1..3) */
/* z-line partition and elimination factors (coefficient parts, per k) */
typedef struct orc_zline {
    int C;              /* chunks */
    double h;           /* dt/Re */
    double *a, *R, *cf, *ap2, *cp2;   /* [nZ+1], 1-based k */
    double *c0;         /* [C]  F-row rhs coupling */
    double *rA, *rg, *re;             /* [2C] reduced-system Thomas factors */
} orc_zline;
int  orc_zline_chunks(int nz);
int  orc_zline_begin(int ch, int C, int nz);
int  orc_zline_factors(const orc_grid *g, orc_zline *z);
void orc_zline_free(orc_zline *z);
void orc_helm_sweep(const orc_grid *g, const double *Q, double *Ut, int p, int order);

/* ---- a6: PPE right-hand side (dnssetbl.f90:437-447) ---- */
void orc_ppe_rhs(const orc_grid *g, const double *Ut, double *pRhs);

/* ---- a7: Poisson operator on the interior; out = A x (boundary of out untouched) ---- */
void orc_poisson_apply(const orc_grid *g, const double *x, double *out);

/* ---- a8: preconditioner  out = M^{-1} r (interior; boundary of out set to 0) ---- */
void orc_precond_apply(const orc_grid *g, int pc, const double *r, double *out);

/* ---- a9: restarted left-preconditioned GMRES (dnssetbl.f90:451-671) ---- */
/* stats: [0]=cycles m, [1]=total inner iterations, [2]=last kit, [3]=final sumav,
 *        [4]=sumnor, [5]=true residual ||pRhs-A PCorr||_2 (interior) after the solve */
void orc_ppe_gmres(const orc_grid *g, const double *pRhs, double *PCorr,
                   int pc, double errtol, int kmax, int mmax, double *stats);

/* ---- a10: PCorr extrapolation + pressure update + projection (dnssetbl.f90:673-726) ---- */
void orc_project(const orc_grid *g, double *PCorr, double *Press, double *Ut);

/* ---- a11: velocity + pressure boundary conditions (dnssetbl.f90:729-797) ---- */
/* wall_w: optional (nX,nY) wall-normal wall velocity (control hook at k=1), NULL = 0 */
void orc_apply_bcs(const orc_grid *g, double *Ut, const double *Utold, double *Press,
                   const double *wall_w);

/* ---- a12: divergence check, returns |sum div| / Ni (dnssetbl.f90:851-882) ---- */
double orc_div_check(const orc_grid *g, const double *Ut, double *signed_sum);

/* ---- a13: mass flux at 5 planes (dnssetbl.f90:885-904) ---- */
void orc_mass_flux(const orc_grid *g, const double *Ut, double *mass5);

/* ---- a14: statistics accumulation (dnssetbl.f90:936-984); sums = 9 x (nX,nZ) ---- */
void orc_stats_accumulate(const orc_grid *g, const double *Ut, double *sums);

/* ---- a15: time average + rescaling update (dnssetbl.f90:1015-1153) ---- */
/* sums in: 9x(nX,nZ); outputs: avg = 9x(nX,nZ) (usum2..vwsum2), lamda, lamda03,
 * Wt(nX,nZ), newk(nZ), bltkns/deltastar/theta(nX), Ufluc_rec (nY,nZ,3) from Ut at i_rec */
void orc_rescale(const orc_grid *g, const double *sums, int nsamples, const double *Ut,
                 double *avg, double *lamda, double *lamda03, double *Wt, int *newk,
                 double *bltkns, double *deltastar, double *theta, double *Ufluc_rec);

/* ---- a16: kinetic energy of fluctuations + blow-up guard (dnssetbl.f90:1223-1261) ---- */
/* returns KE; *guard = 1 if max component fluctuation > 25; Ufluc_rec updated at i_rec */
double orc_ke_guard(const orc_grid *g, const double *Ut, const double *usum2,
                    const double *vsum2, const double *wsum2, int *guard, double *Ufluc_rec);

/* ---- a17: wall shear per x (dnstbl/dnstbl.f90:1620-1668); tau(nX) = sum_j (dw/dx+du/dz)/Re ---- */
void orc_wall_shear(const orc_grid *g, const double *Ut, double *tau);

/* The whole step (dnssetbl.f90:253-1262) is composed from these blocks by
 * the Python wrapper (oracle/orc.py, State.step), the one composition the
 * tests check. */

#ifdef __cplusplus
}
#endif
#endif
