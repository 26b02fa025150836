/*
 * tbl.h -- C ABI of libtbl_hip.so, the MI355X (gfx950) implementation of the
 * reference's fractional-step time step for the spatially developing
 * turbulent boundary layer (theparvezahmad/turbulentBL, dnssetbl/dnssetbl.f90).
 *
 * The reference has no subroutines or FFI: every block of its time loop is an
 * inline loop nest over global arrays (dnssetbl.f90:253-1262).  Each entry
 * point below replaces one of those blocks (cited per function) and keeps the
 * reference's array shapes and layout for every host buffer it takes:
 * Fortran order, i (streamwise) fastest, 1-based on the Fortran side,
 * Ut(nX,nY,nZ,3), Press/PCorr(nX,nY,nZ), statistics (nX,nZ), FP64.
 *
 * Ownership: host arrays belong to the caller and are never retained; device
 * fields, Krylov workspace and statistics belong to the context.  One host
 * thread and one HIP stream per context; calls are stream-ordered and every
 * call that returns host data synchronises.  Errors are returned, never
 * thrown: 0 = OK, >0 = HIP error code, <0 = one of the TBL_E* values below;
 * tbl_last_error() gives the message.
 */
#ifndef TBL_H
#define TBL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tbl_ctx tbl_ctx;

enum {
    TBL_OK = 0,
    TBL_EARG = -1,       /* invalid argument / unsupported size */
    TBL_EBLOWUP = -2,    /* fluctuation > 25 (dnssetbl.f90:1251-1261) */
    TBL_ENAN = -3,       /* non-finite value in a reduction */
    TBL_ENOCONV = -4,    /* solver hit its iteration cap (only with fail_on_noconv) */
    TBL_ESTATE = -5      /* call out of order (e.g. before tbl_upload) */
};

enum { TBL_FP64 = 0, TBL_MIXED = 1 };                 /* storage/compute precision      */
enum { TBL_PC_SSOR_RB = 1, TBL_PC_NONE = 2, TBL_PC_JACOBI = 3,  /* Poisson preconditioner */
       TBL_PC_FAST = 4,     /* separable fast-diagonalisation (exact inverse of A in exact arithmetic) */
       TBL_PC_SSOR_BLOCK = 5 };  /* the reference's per-rank block SSOR (dnssetbl.f90:509-532,586-610):
                                    lexicographic forward and backward sweeps inside each of `slabs`
                                    wall-normal slabs, zero couplings across slab and span faces */
enum { TBL_HELM_POINT_RB = 0, TBL_HELM_ZLINE_RB = 1,  /* Gauss-Seidel ordering for the diffusion solve */
       TBL_HELM_DIRECT = 2,    /* exact solve by fast diagonalisation + one refinement (no sweeps) */
       TBL_HELM_ZEBRA = 3,     /* wall-normal lines solved exactly, span rows in zebra order (two
                                  passes by the parity of j), x neighbours from the start of the
                                  pass; whole-row memory access, one extra velocity buffer */
       TBL_HELM_LEX_SLAB = 4 };/* the reference's ordering (dnssetbl.f90:395-430): lexicographic
                                  point Gauss-Seidel inside each of `slabs` wall-normal slabs, the
                                  ghost planes from before the sweep (one workgroup per slab) */
enum { TBL_TOL_REF = 0, TBL_TOL_TRUE = 1 };           /* GMRES stopping rule            */
/* convection scheme of the momentum right-hand side */
enum { TBL_SCHEME_SETBL = 0,   /* dnssetbl/dnssetbl.f90:306-369: Kawamura-Kuwahara where |Re u dx| > 2,
                                  2nd-order central at the near-boundary nodes */
       TBL_SCHEME_DNSTBL = 1 };/* dnstbl/dnstbl.f90:620-766: Kawamura-Kuwahara only inside the cell-Peclet
                                  window 2 < Pe < 50 (2 < Pe < 5 wall-normal), Pe signed by
                                  sign(1, ceiling(u)); first-order upwind at the span-edge rows
                                  j = 2, nY-1; constant body force forc in Q */

/* Parameters: the reference's compile-time PARAMETERs (dnssetbl.f90:12-34)
 * plus the choices the reference hard-codes (domain, recycling plane, mass
 * flux planes) and this implementation's precision / solver options. */
typedef struct tbl_params {
    int nX, nY, nZ;            /* grid incl. boundary/ghost nodes; nY even (red-black) */
    double Lx, Ly, Lz;         /* x1(nX), x2(nY), x3(nZ): 12, 4, 3.4 in the reference  */
    double delta;              /* tanh condensation parameter (1.95)                   */
    double dt, Re, Ue;         /* 1e-4, 150.92, 19.798                                 */
    double omega;              /* SSOR relaxation (REAL(4) 1.835)                      */
    int i_rec;                 /* recycling plane, 1-based (68)                        */
    int freq_tim_avg;          /* averaging window in steps (500)                      */
    int planes[5];             /* mass-flux planes, 1-based (5,20,40,60,85)            */
    int precision;             /* TBL_FP64 or TBL_MIXED                                */
    /* implicit diffusion (Gauss-Seidel) */
    double errorstop;          /* 1e-6  */
    int gs_mmax;               /* 1000  */
    int helm_mode;             /* TBL_HELM_POINT_RB, TBL_HELM_ZLINE_RB, TBL_HELM_ZEBRA (wall-normal lines),
                                  TBL_HELM_LEX_SLAB or TBL_HELM_DIRECT */
    /* pressure Poisson GMRES */
    int tol_mode;              /* TBL_TOL_REF: preconditioned residual relative to the
                                  first cycle (the reference's rule, :541-542);
                                  TBL_TOL_TRUE: ||pRhs - A PCorr||_2 <= errtol ||pRhs||_2, both
                                  norms over the interior nodes */
    double errtol;             /* 1e-5 */
    int kmax;                  /* restart length (50), <= 64 */
    int mmax;                  /* max cycles (1000) */
    int precond;               /* TBL_PC_* */
    double inner_tol;          /* TOL_TRUE: per-cycle reduction of the inner solve */
    int fail_on_noconv;        /* return TBL_ENOCONV when a solver hits its cap */
    int wall_shear_window;     /* 0 = off; n > 0: tbl_step adds the wall shear tau(i) of every
                                  step to a running sum and, at each step with nstep % n == 0,
                                  forms the window mean tau(i)/(nY-2) = sum/n/(nY-2) and
                                  C_f(i) = 2 (sum/n)/(Ue^2 (nY-2)), then clears the sum (the
                                  reference's 500-step window, dnstbl/dnstbl.f90:1666-1682);
                                  read them with tbl_wall_shear_mean */
    int pc_extrap;             /* GMRES initial guess: 0 = the previous step's PCorr (the
                                  reference's, :462-480); 1 = linear extrapolation in time
                                  2 PCorr(n-1) - PCorr(n-2) (pays off only where the pressure
                                  evolves smoothly enough to cut refinement cycles) */
    int scheme;                /* TBL_SCHEME_SETBL (default) or TBL_SCHEME_DNSTBL */
    double forc[3];            /* body force per component (dnstbl.f90:51,133-135,755-761; 0) */
    int slabs;                 /* wall-normal slabs of the per-slab orderings (TBL_HELM_LEX_SLAB,
                                  TBL_PC_SSOR_BLOCK): the reference's MPI ranks, bounds as
                                  dnssetbl.f90:76-92 (remainder planes to the high ranks); 1..256,
                                  every slab at least one interior plane, and
                                  ceil((nY-2)/64)*64 * ceil(planes/32) <= 512 threads per slab */
} tbl_params;

/* Recycling / statistics state (dnssetbl.f90:36-62 arrays that persist
 * between steps).  Any pointer may be NULL in tbl_get_state (not copied) and
 * in tbl_set_state (left as is). */
typedef struct tbl_state {
    double *sums;        /* 9*(nX,nZ) running sums usum vsum wsum uusum vvsum wwsum uvsum uwsum vwsum */
    double *avg;         /* 9*(nX,nZ) last averages usum2 ... vwsum2 */
    double *Wt;          /* (nX,nZ) Lund weight */
    int *newk;           /* (nZ) recycling index map, 1-based */
    double *Ume;         /* (nZ) inflow mean profile (Input/initProf.dat) */
    double *Ufluc_rec;   /* (nY,nZ,3) fluctuations at the recycling plane */
    double *bltkns;      /* (nX) delta99 */
    double *deltastar;   /* (nX) */
    double *theta;       /* (nX) */
    double *wall_w;      /* (nX,nY) wall-normal wall velocity (blowing/suction hook) */
    double lamda, lamda03;
    double time;
    int nstep, cnt;
} tbl_state;

/* Per-step diagnostics (dnssetbl.f90:851-905,1223-1261; C_f from dnstbl/dnstbl.f90:1620-1684) */
typedef struct tbl_diag {
    double div;          /* |sum div4| / ((nX-2)(nY-2)(nZ-2)), the reference's printed metric */
    double div_signed;   /* the signed sum */
    double mass[5];      /* mass flux at the 5 planes */
    double ke;           /* fluctuation kinetic energy (0 until cnt > freq_tim_avg) */
    int guard;           /* bit set: 1 a fluctuation exceeded 25, 2 a non-finite value */
    int gs_sweeps[3];    /* Gauss-Seidel sweeps per component */
    int gm_cycles, gm_iters;
    double gm_true_res;  /* ||pRhs - A PCorr||_2 after the solve */
    double gm_sumav;     /* final stopping measure */
    int rescaled;        /* a statistics rescale happened in this step */
    int nstep;
    int cf_ready;        /* a wall-shear window (wall_shear_window) ended in this step */
} tbl_diag;

void tbl_default_params(tbl_params *p);
int tbl_create(const tbl_params *p, tbl_ctx **out);
int tbl_destroy(tbl_ctx *ctx);
/* the message of the last failed call on ctx; with ctx = NULL, the message
 * of the last tbl_create that failed on the calling thread (empty after one
 * that succeeded) */
int tbl_last_error(const tbl_ctx *ctx, char *buf, int len);
int tbl_sync(tbl_ctx *ctx);
/* bytes of device memory the context holds */
double tbl_device_bytes(const tbl_ctx *ctx);

/* Fields: host FP64 arrays in the reference layout; NULL = leave / skip.
 * Replaces the restart read of Input/3D.dat (dnssetbl.f90:118-135). */
int tbl_upload(tbl_ctx *ctx, const double *Ut, const double *Press, const double *PCorr);
int tbl_download(tbl_ctx *ctx, double *Ut, double *Press, double *PCorr);
/* named work fields for block-level calls: see TBL_F_*.  TBL_F_UTOLD: the step
 * reads Utold only at the inflow plane of u (dnssetbl.f90:791), so
 * tbl_begin_step saves only Ut(1,:,:,1) there; its other entries keep what
 * was last set. */
enum { TBL_F_UT = 0, TBL_F_UTOLD = 1, TBL_F_PRESS = 2, TBL_F_PCORR = 3, TBL_F_PRHS = 4,
       TBL_F_Q = 5, TBL_F_WORK = 6 };
int tbl_field_set(tbl_ctx *ctx, int field, const double *host);
int tbl_field_get(tbl_ctx *ctx, int field, double *host);

int tbl_set_state(tbl_ctx *ctx, const tbl_state *s);
int tbl_get_state(tbl_ctx *ctx, tbl_state *s);

/* Binary checkpoint and bitwise resume: the fields in their storage precision,
 * the pressure-correction warm start(s), statistics, recycling state,
 * wall_w, the wall-shear window and the clock.  A context created with the
 * same parameters and loaded from the file steps bit for bit as the saved
 * one would.  Replaces the reference's restart files for resuming
 * (dnssetbl.f90:118-150 read, :906-934 and :1193-1197 written), whose
 * E22.15 / F15.7 text fields do not round-trip FP64 values; those formats
 * stay available through the Fortran driver.  TBL_EARG if the file's grid or
 * precision differs from the context's. */
int tbl_checkpoint_save(tbl_ctx *ctx, const char *path);
int tbl_checkpoint_load(tbl_ctx *ctx, const char *path);

/* Mesh and coefficients computed on the host at create time
 * (dnssetbl.f90:183-247); arrays as in the reference, NULL to skip. */
int tbl_get_mesh(const tbl_ctx *ctx, double *x1, double *x2, double *x3, double *Jac,
                 double *Jac2, double *Ab, double *At, double *Ap);

/* ---- blocks of one step, in the reference's order ---- */
/* time += dt; nstep++; cnt++; Utold = Ut  (dnssetbl.f90:254-257) */
int tbl_begin_step(tbl_ctx *ctx);
/* a2: recycling inflow, active when nstep > freq_tim_avg (dnssetbl.f90:259-303) */
int tbl_recycle_inflow(tbl_ctx *ctx);
/* a3+a4: convection + old-pressure gradient -> Q (dnssetbl.f90:306-369) */
int tbl_momentum_rhs(tbl_ctx *ctx);
/* a5: implicit diffusion, Gauss-Seidel (red-black) to errorstop (dnssetbl.f90:371-433) */
int tbl_diffusion_solve(tbl_ctx *ctx, double errorstop, int mmax, int *sweeps3, double *res3);
/* a6: PPE right-hand side (dnssetbl.f90:437-447) */
int tbl_ppe_rhs(tbl_ctx *ctx);
/* a7-a9: pressure Poisson GMRES with device-resident Arnoldi/Givens (dnssetbl.f90:449-671) */
int tbl_ppe_solve(tbl_ctx *ctx, double errtol, int kmax, int mmax, int *cycles, int *iters,
                  double *true_res);
/* a10: PCorr extrapolation + pressure update + projection (dnssetbl.f90:673-726) */
int tbl_project(tbl_ctx *ctx);
/* a11: velocity + pressure boundary conditions (dnssetbl.f90:728-797) */
int tbl_apply_bcs(tbl_ctx *ctx);
/* a12: divergence metric (dnssetbl.f90:851-882) */
int tbl_div_check(tbl_ctx *ctx, double *div, double *signed_sum);
/* tbl_div_check without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Reads the
 * caller's device velocity Ut_dev (nX,nY,nZ,3) in the context's storage type
 * (double in FP64 contexts, float in mixed ones) and writes the caller's
 * device array out_dev[2] = {|sum div4| / ((nX-2)(nY-2)(nZ-2)), signed sum},
 * the values tbl_div_check returns, ordered on the context's stream after its
 * earlier work; the host is not synchronised (tbl_sync before reading
 * out_dev).  The block partials live in one static device array per GPU of
 * 32768 doubles (256 KiB, held until the process ends, set up by
 * tbl_create): grids needing more partials return TBL_EARG.  Every call on a
 * GPU uses that GPU's array, so calls that overlap on one GPU (contexts on
 * different streams or threads) race. */
int tbl_div_check_noalloc(tbl_ctx *ctx, const void *Ut_dev, double *out_dev);
/* a13: mass flux at the 5 planes (dnssetbl.f90:885-904) */
int tbl_mass_flux(tbl_ctx *ctx, double *mass5);
/* a14: spanwise statistics accumulation (dnssetbl.f90:936-984) */
int tbl_stats_accumulate(tbl_ctx *ctx);
/* a15: time average + recycling rescale, resets the sums (dnssetbl.f90:986-1220) */
int tbl_rescale(tbl_ctx *ctx);
/* a16: fluctuation kinetic energy + blow-up guard (dnssetbl.f90:1223-1261);
 * *guard is the bit set of tbl_diag.guard */
int tbl_ke_guard(tbl_ctx *ctx, double *ke, int *guard);
/* a17: wall shear per streamwise station, tau(nX) (dnstbl/dnstbl.f90:1620-1668) */
int tbl_wall_shear(tbl_ctx *ctx, double *tau);
/* a17: the last completed wall_shear_window of tbl_step (dnstbl/dnstbl.f90:1666-1682):
 * tau_mean(i) = (sum_n tau_n(i) / window) / (nY-2) and
 * cf(i) = 2 (sum_n tau_n(i) / window) / (Ue^2 (nY-2)), i = 1..nX (0 at i = 1 and nX);
 * nwindows = windows completed since tbl_create (0: the arrays are zero).  NULL
 * pointers are skipped.  TBL_ESTATE when wall_shear_window is 0. */
int tbl_wall_shear_mean(tbl_ctx *ctx, double *tau_mean, double *cf, int *nwindows);

/* whole steps (dnssetbl.f90:253-1262); diag may be NULL or hold nsteps records */
int tbl_step(tbl_ctx *ctx, int nsteps, tbl_diag *diag);

/* operator hooks on the work fields, for checking a7/a8 in isolation:
 * out = A in  (op 0)  or  out = M^{-1} in  (op 1) */
int tbl_apply_operator(tbl_ctx *ctx, int op, const double *in, double *out);

/* ---- discrete adjoint ----
 * No counterpart in the reference: the gradient of a downstream wall-shear
 * objective with respect to the wall-normal wall velocity (the blowing /
 * suction hook w(i,j,1) = wall_w(i,j) of dnssetbl.f90:729).  Over a window of
 * nsteps steps that crosses no statistics rescale (:986-1220; the recycling
 * inflow is then frozen),
 *   J = (1/nsteps) sum_n mean_{i in [i0,i1]} tau_n(i) / (nY-2),
 * tau_n the wall shear of dnstbl/dnstbl.f90:1620-1668 after step n.  The
 * window's steps solve the Helmholtz and Poisson systems exactly (fast
 * diagonalisation + refinement) and update no statistics; the adjoint applies
 * the transposed blocks in reverse with binomial checkpointing (nsnap
 * snapshots).  In an FP64 context everything is FP64.  In a mixed context the
 * forward steps run in mixed precision (FP32 fields and fast solves, the
 * pressure by FP32 solves inside FP64 refinement), checkpoints hold FP32
 * states, and the adjoint is accumulated in FP64 about the mixed trajectory
 * (states converted to double, transposed solves in FP64).  On return the
 * context holds the state at the end of the window and wall_w = the given
 * profile. */
typedef struct tbl_adjoint_opts {
    int nsteps;          /* window length (steps) */
    int nsnap;           /* checkpoints for the reverse sweep (>= 1) */
    int i0, i1;          /* objective stations, 1-based, 2 <= i0 <= i1 <= nX-1 */
    int forward_only;    /* 1: evaluate J only */
    int branch_mode;     /* Peclet switches of the convection (dnssetbl.f90:316-350) are
                            discontinuous in the velocity: 0 = evaluate them, 1 = evaluate and
                            record them for this window, 2 = replay the recorded pattern
                            (forward_only; same start step), so that finite differences
                            see the piecewise-smooth map the adjoint differentiates */
} tbl_adjoint_opts;
/* wall_w: (nX,nY) wall-normal wall velocity, NULL = keep the current one;
 * dJdw: (nX,nY) gradient, may be NULL; nfwd/nadj: forward / adjoint steps run */
int tbl_adjoint_window(tbl_ctx *ctx, const tbl_adjoint_opts *opts, const double *wall_w, double *J, double *dJdw,
                       int *nfwd, int *nadj);

/* kernel timers (HIP events on the context's stream).  Classes: see
 * TBL_K_*; tbl_timer_enable(ctx, mask) times the classes in the bit mask. */
enum { TBL_K_RHS = 0, TBL_K_HELM = 1, TBL_K_HELM_RES = 2, TBL_K_DIV = 3, TBL_K_POISSON = 4,
       TBL_K_PRECOND = 5, TBL_K_DOTS = 6, TBL_K_UPDATE = 7, TBL_K_XUPD = 8, TBL_K_RESID = 9,
       TBL_K_PROJECT = 10, TBL_K_BCS = 11, TBL_K_DIAG = 12, TBL_K_RECYCLE = 13, TBL_K_RESCALE = 14,
       TBL_K_NCLASS = 15 };
int tbl_timer_enable(tbl_ctx *ctx, unsigned mask);
/* per class: launches counted, total milliseconds and total algorithmic bytes
 * (compulsory HBM traffic at storage precision); resets the counters */
int tbl_timer_read(tbl_ctx *ctx, int *count, double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
