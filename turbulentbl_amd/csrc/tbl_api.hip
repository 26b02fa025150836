// tbl_api.hip -- the extern "C" entry points of libtbl_hip.so (see include/tbl.h).
//
// One context = one device, one HIP stream, all fields resident in HBM in the
// reference's layout (Fortran order, i fastest).  FP64 mode stores every
// field in double; mixed mode stores velocity, pressure, Q and the Krylov
// basis in float and keeps PCorr, pRhs, the outer residual, the Gram-Schmidt
// coefficients, Hessenberg/Givens, divergence, statistics and the recycling
// rescale in double.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <sched.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/tbl.h"
#include "tbl_dev.h"
#include "tbl_launch.h"

using namespace tbl;

struct TimerRec { hipEvent_t a, b; int cls; double bytes; int n; };   // n launches between a and b

struct tbl_ctx {
    tbl_params p;
    Geo g;
    hipStream_t s = nullptr;
    int dev = 0;
    bool mixed = false;
    size_t es = 8;                       // bytes per stored field element
    // host mesh (a1)
    std::vector<double> x1, x2, x3, Xi, Jac, Jac2, JacInv, Ab, At, Ap, rAp;
    double mass_denom = 1.0;
    // device buffers
    double *mesh = nullptr;
    void *Ut = nullptr, *Utold = nullptr, *Q = nullptr, *Press = nullptr;
    void *Ualt = nullptr;                // second velocity buffer of the zebra sweep (Ut and Ualt swap)
    int *slab = nullptr;                 // per-slab orderings: 0-based first/last interior plane per slab
    void *halo = nullptr;                // per-slab Gauss-Seidel: ghost-plane snapshots (3 x 2 slabs planes)
    double *PCorr = nullptr, *pRhs = nullptr;
    double *PCprev = nullptr;            // PCorr of the step before (extrapolated initial guess)
    double *PCalt = nullptr;             // second PCorr buffer: the fused update + residual writes it
    int pc_hist = 0;                     // previous solutions available for the extrapolation
    void *V = nullptr, *W = nullptr, *scr = nullptr, *r = nullptr;
    void *Z = nullptr;                   // flexible-GMRES preconditioned vectors (true-residual mode)
    long ldv = 0;
    double *part = nullptr;
    size_t npart = 0;
    GmCtrl *gm = nullptr;
    GmStatus *gm_h = nullptr;            // pinned
    void *gs = nullptr, *gs_h = nullptr; // Gauss-Seidel control (device / pinned)
    double *sums = nullptr, *avg = nullptr, *Wt = nullptr, *Ume = nullptr, *uf = nullptr;
    double *bltkns = nullptr, *dstar = nullptr, *theta = nullptr, *wall_w = nullptr, *tau = nullptr;
    double *tausum = nullptr, *taumean = nullptr, *cf = nullptr;   // wall-shear window (a17)
    int nwin = 0;                        // wall-shear windows completed
    double *scan = nullptr;
    FastPC *fast = nullptr;
    void *zf = nullptr;                  // z-line Gauss-Seidel elimination factors (storage type)
    int *newk = nullptr;
    RecDev *rec = nullptr;
    int *guard = nullptr;
    double *dd = nullptr;                // device diagnostics record
    double *noalloc = nullptr;           // this device's copy of `scratch` (tbl_div_check_noalloc)
    double *dh = nullptr;                // pinned copy
    bool have_wall_w = false;
    // Peclet-switch patterns of an adjoint window (record / replay): one byte
    // per cell and step, window of br_steps steps starting after step br_nstep0
    unsigned char *br = nullptr;
    int br_steps = 0, br_nstep0 = -1;
    const unsigned char *br_in = nullptr;   // momentum RHS: replay this pattern
    unsigned char *br_out = nullptr;        // momentum RHS: record into this
    // clock
    double time = 0.0;
    int nstep = 0, cnt = 0;
    bool uploaded = false;
    // timers
    unsigned tmask = 0;
    std::vector<TimerRec> pend;
    std::vector<hipEvent_t> pool;
    hipEvent_t poll_ev = nullptr;        // marks a stopping-test read-back the host spins on
    int tcount[TBL_K_NCLASS] = {0};
    double tms[TBL_K_NCLASS] = {0};
    double tbytes[TBL_K_NCLASS] = {0};
    std::string err;
};

// device diagnostics record layout
// device diagnostics record; the mass-flux partials (5 * MF_BLK) follow it
enum { DD_DIV = 0, DD_MASS = 1, DD_KE = 6, DD_BNORM = 7, DD_NCOUNT = 8 };

#define HIPCHK(ctx, call)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return (int)e_;                                                                  \
        }                                                                                    \
    } while (0)
#define CHK(expr)                \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != 0) return rc_;\
    } while (0)

static int fail(tbl_ctx *c, int code, const char *msg)
{
    if (c) c->err = msg;
    return code;
}

// ------------------------------------------------------------------ timers
static void timer_flush(tbl_ctx *c)
{
    if (c->pend.empty()) return;
    hipStreamSynchronize(c->s);
    for (auto &t : c->pend) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, t.a, t.b);
        c->tcount[t.cls] += t.n;
        c->tms[t.cls] += ms;
        c->tbytes[t.cls] += t.bytes;
        c->pool.push_back(t.a);
        c->pool.push_back(t.b);
    }
    c->pend.clear();
}
static hipEvent_t ev_get(tbl_ctx *c)
{
    if (c->pool.empty()) {
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
}
// Times the launches in its scope when the class is enabled; `bytes` is the
// algorithmic (compulsory) HBM traffic of those launches: each distinct array
// element read once and written once at its storage precision.
struct Tscope {
    tbl_ctx *c;
    int cls;
    double bytes;
    int n;                                 // launches in the scope (back to back, one pair of events)
    hipEvent_t a = nullptr;
    Tscope(tbl_ctx *c_, int cls_, double bytes_ = 0.0, int n_ = 1) : c(c_), cls(cls_), bytes(bytes_), n(n_)
    {
        if (c->tmask & (1u << cls)) {
            a = ev_get(c);
            hipEventRecord(a, c->s);
        }
    }
    ~Tscope()
    {
        if (a) {
            hipEvent_t b = ev_get(c);
            hipEventRecord(b, c->s);
            c->pend.push_back({a, b, cls, bytes, n});
            if (c->pend.size() > 4096) timer_flush(c);
        }
    }
};

// ------------------------------------------------------------------ a1: mesh
// dnssetbl.f90:183-247, same arithmetic as the reference (REAL(4) literals kept)
static void build_mesh(tbl_ctx *c)
{
    const tbl_params &p = c->p;
    const int nX = p.nX, nY = p.nY, nZ = p.nZ;
    c->x1.assign(nX, 0.0); c->x2.assign(nY, 0.0); c->x3.assign(nZ, 0.0); c->Xi.assign(nZ, 0.0);
    c->Jac.assign(nZ, 0.0); c->Jac2.assign(nZ, 0.0); c->JacInv.assign(nZ, 0.0);
    c->Ab.assign(nZ, 0.0); c->At.assign(nZ, 0.0); c->Ap.assign(nZ, 0.0); c->rAp.assign(nZ, 0.0);
    const double dx = p.Lx / (double)(nX - 1), dy = p.Ly / (double)(nY - 1), dXi = p.Lz / (double)(nZ - 1);
    for (int i = 0; i < nX - 1; i++) c->x1[i + 1] = c->x1[i] + dx;
    for (int i = 0; i < nY - 1; i++) c->x2[i + 1] = c->x2[i] + dy;
    for (int k = 0; k < nZ - 1; k++) c->Xi[k + 1] = c->Xi[k] + dXi;
    const double delta = p.delta, Lzd = p.Lz, Lhd = p.Lz * 0.5;
    const double Lzf = (double)(float)p.Lz, Lhf = (double)(float)(p.Lz * 0.5);
    for (int k = 0; k < nZ; k++) {
        const double a = delta * 0.5 * (c->Xi[k] - Lzf);
        const double ch = cosh(0.5 * delta * (c->Xi[k] - Lzf));
        c->x3[k] = Lzd + Lzd * tanh(a) / (tanh(Lhd * delta));
        c->Jac[k] = Lhd * delta / (tanh(Lhf * delta) * (ch * ch));
        c->Jac2[k] = -(Lhd * (delta * delta) * tanh(0.5 * delta * (c->Xi[k] - Lzf)) / (tanh(Lhf * delta) * (ch * ch)));
    }
    for (int k = 0; k < nZ; k++) {
        const double J = c->Jac[k];
        c->Ab[k] = 1.0 / (dXi * dXi * J * J) + c->Jac2[k] / (2.0 * dXi * J * J * J);
        c->At[k] = 1.0 / (dXi * dXi * J * J) - c->Jac2[k] / (2.0 * dXi * J * J * J);
        c->Ap[k] = -2.0 * ((1.0 / (dx * dx)) + (1.0 / (dy * dy)) + (1.0 / (dXi * dXi * J * J)));
        c->rAp[k] = 1.0 / c->Ap[k];
        c->JacInv[k] = 1.0 / J;
    }
    Geo &g = c->g;
    g.nX = nX; g.nY = nY; g.nZ = nZ;
    g.sZ = (long)nX * nY;
    g.N = (long)nX * nY * nZ;
    g.dx = dx; g.dy = dy; g.dXi = dXi;
    g.invdx = 1.0 / dx; g.invdy = 1.0 / dy; g.invdXi = 1.0 / dXi;
    g.c12 = 1.0 / 12.0;
    g.Aw = 1.0 / (dx * dx); g.Ae = 1.0 / (dx * dx); g.An = 1.0 / (dy * dy); g.As = 1.0 / (dy * dy);
    g.df1 = c->x3[1] - c->x3[0];
    g.df2 = c->x3[2] - c->x3[0];
    g.dt = p.dt; g.Re = p.Re; g.Ue = p.Ue; g.omega = p.omega;
    g.scheme = p.scheme;
    for (int q = 0; q < 3; q++) g.forc[q] = p.forc[q];
    c->mass_denom = (c->x3[nZ - 2] - c->x3[1]) * (c->x2[nY - 2] - c->x2[1]);
}

extern "C" void tbl_default_params(tbl_params *p)
{
    memset(p, 0, sizeof(*p));
    p->nX = 90; p->nY = 80; p->nZ = 50;
    p->Lx = 12.0; p->Ly = 4.0; p->Lz = 3.4; p->delta = 1.95;
    p->dt = 1.0e-4; p->Re = 150.92; p->Ue = 19.798;
    p->omega = (double)1.835f;
    p->i_rec = 68; p->freq_tim_avg = 500;
    const int pl[5] = {5, 20, 40, 60, 85};
    for (int q = 0; q < 5; q++) p->planes[q] = pl[q];
    p->precision = TBL_FP64;
    p->errorstop = 1.0e-6; p->gs_mmax = 1000;
    p->tol_mode = TBL_TOL_REF; p->errtol = 1.0e-5; p->kmax = 50; p->mmax = 1000;
    p->precond = TBL_PC_FAST; p->inner_tol = 1.0e-3;
    p->helm_mode = TBL_HELM_ZEBRA;
    p->fail_on_noconv = 0; p->wall_shear_window = 0; p->pc_extrap = 0;
    p->scheme = TBL_SCHEME_SETBL;
    p->forc[0] = p->forc[1] = p->forc[2] = 0.0;
    p->slabs = 1;
}

static int h2d(tbl_ctx *c, void *d, const void *h, size_t b);

template <class X> static int dalloc(tbl_ctx *c, X **ptr, size_t bytes)
{
    HIPCHK(c, hipMalloc((void **)ptr, bytes ? bytes : 8));
    HIPCHK(c, hipMemsetAsync(*ptr, 0, bytes ? bytes : 8, c->s));
    return 0;
}

// ------------------------------------------------------------------ static scratch
// Scratch of tbl_div_check_noalloc: the block partials of the divergence sum,
// in a file-scope device array so that the call allocates nothing.  Every
// device holds its own copy until the process ends; host code uses the
// device address hipGetSymbolAddress returns for it (never the array's name,
// which on the host is a stand-in and not a device address).
#define TBL_NOALLOC_PARTS 32768
#define TBL_MAX_DEVICES 64
static __device__ __align__(256) double scratch[TBL_NOALLOC_PARTS];
static std::mutex scratch_mu;
static double *scratch_dev[TBL_MAX_DEVICES];

// once per device, with that device current (called from tbl_create)
static int scratch_setup(tbl_ctx *c)
{
    if (c->dev < 0 || c->dev >= TBL_MAX_DEVICES) return fail(c, TBL_EARG, "device index above 63");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[c->dev]) {
        void *ptr = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        scratch_dev[c->dev] = (double *)ptr;
    }
    c->noalloc = scratch_dev[c->dev];
    return 0;
}

static int create_impl(tbl_ctx *c)
{
    const tbl_params &p = c->p;
    const long N = (long)p.nX * p.nY * p.nZ;
    HIPCHK(c, hipGetDevice(&c->dev));
    HIPCHK(c, hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    CHK(scratch_setup(c));
    build_mesh(c);
    const int nZ = p.nZ;
    CHK(dalloc(c, &c->mesh, sizeof(double) * 7 * nZ));
    std::vector<double> m;
    for (auto *v : {&c->x3, &c->Jac2, &c->JacInv, &c->Ab, &c->At, &c->Ap, &c->rAp}) m.insert(m.end(), v->begin(), v->end());
    HIPCHK(c, hipMemcpyAsync(c->mesh, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, c->s));
    c->g.x3 = c->mesh; c->g.Jac2 = c->mesh + nZ; c->g.JacInv = c->mesh + 2 * nZ; c->g.Ab = c->mesh + 3 * nZ;
    c->g.At = c->mesh + 4 * nZ; c->g.Ap = c->mesh + 5 * nZ; c->g.rAp = c->mesh + 6 * nZ;

    c->es = c->mixed ? sizeof(float) : sizeof(double);
    CHK(dalloc(c, &c->Ut, 3 * N * c->es));
    CHK(dalloc(c, &c->Utold, 3 * N * c->es));
    CHK(dalloc(c, &c->Q, 3 * N * c->es));
    CHK(dalloc(c, &c->Press, N * c->es));
    CHK(dalloc(c, &c->PCorr, N * sizeof(double)));
    CHK(dalloc(c, &c->pRhs, N * sizeof(double)));
    c->ldv = (N + 63) & ~63L;
    CHK(dalloc(c, &c->V, (size_t)(p.kmax + 1) * c->ldv * c->es));
    CHK(dalloc(c, &c->W, N * c->es));
    CHK(dalloc(c, &c->scr, N * c->es));
    CHK(dalloc(c, &c->r, N * c->es));
    if (p.tol_mode == TBL_TOL_TRUE) CHK(dalloc(c, &c->Z, (size_t)p.kmax * c->ldv * c->es));
    if (p.helm_mode == TBL_HELM_ZEBRA) CHK(dalloc(c, &c->Ualt, 3 * N * c->es));
    c->g.nslab = 0; c->g.nkmax = 0; c->g.slab = nullptr;
    if (p.helm_mode == TBL_HELM_LEX_SLAB || p.precond == TBL_PC_SSOR_BLOCK) {
        // the reference's slab bounds (dnssetbl.f90:76-92): nZ/P planes per
        // rank, the remainder to the high ranks; the first and last rank
        // exclude the wall and top boundary planes
        const int P = p.slabs, nZ = p.nZ;
        std::vector<int> b(P + 1), e(P), sl(2 * P);
        int it1 = nZ / P;
        const int it2 = P - nZ % P;
        b[0] = 1;
        for (int r = 0; r < P; r++) {
            if (r == it2) it1 = it1 + 1;
            e[r] = b[r] + it1 - 1;
            b[r + 1] = e[r] + 1;
        }
        int nkmax = 0;
        for (int r = 0; r < P; r++) {
            const int bn = (r == 0) ? b[r] + 1 : b[r], en = (r == P - 1) ? e[r] - 1 : e[r];
            if (en < bn) return fail(c, TBL_EARG, "slabs: every slab needs at least one interior plane");
            sl[2 * r] = bn - 1;
            sl[2 * r + 1] = en - 1;
            nkmax = std::max(nkmax, en - bn + 1);
        }
        c->g.nslab = P;
        c->g.nkmax = nkmax;
        if (!slab_threads(c->g, nkmax))
            return fail(c, TBL_EARG, "slabs: a slab needs more than 512 threads (ceil((nY-2)/64)*64 * "
                                     "ceil(planes/32)); use more slabs");
        CHK(dalloc(c, &c->slab, sizeof(int) * 2 * P));
        HIPCHK(c, hipMemcpyAsync(c->slab, sl.data(), sizeof(int) * 2 * P, hipMemcpyHostToDevice, c->s));
        c->g.slab = c->slab;
        if (p.helm_mode == TBL_HELM_LEX_SLAB) CHK(dalloc(c, &c->halo, 3 * 2 * (size_t)P * c->g.sZ * c->es));
    }
    if (p.helm_mode == TBL_HELM_ZLINE_RB || p.helm_mode == TBL_HELM_ZEBRA) {
        // elimination factors of the partitioned wall-normal line solve (the
        // line matrix a = -h Ab, b = 1 - h Ap, c = -h At is the same for every
        // column): interior planes k = 1..nZ-2 in C balanced chunks of at most
        // LC planes; per k the forward multiplier a and reciprocal pivot R,
        // the forward c', and the back-substituted a'', c''; per chunk the
        // F-row rhs coupling c0; per first/last unknown the reduced-system
        // Thomas factors.  Computed in double, stored in the storage type.
        const int nZ = p.nZ, nz = nZ - 2;
        if (nz < 2 || nz > 512) return fail(c, TBL_EARG, "z-line Gauss-Seidel needs 4 <= nZ <= 514");
        const int LC = nz <= 256 ? 16 : 32, C = (nz + LC - 1) / LC;
        const double h = p.dt / p.Re;
        std::vector<double> zd(5 * (size_t)nZ + 7 * (size_t)C, 0.0);
        double *za = zd.data(), *zR = za + nZ, *zC = za + 2 * nZ, *zA2 = za + 3 * nZ, *zC2 = za + 4 * nZ;
        double *zc0 = za + 5 * nZ, *rA = zc0 + C, *rg = rA + 2 * C, *re = rg + 2 * C;
        std::vector<double> ap(nZ, 0.0), cp(nZ, 0.0), B(2 * C), Cc(2 * C);
        for (int ch = 0; ch < C; ch++) {
            const int kb = 1 + (int)(((long)ch * nz) / C), ke = 1 + (int)(((long)(ch + 1) * nz) / C) - 1;
            const int M = ke - kb + 1;
            const double a0 = (kb == 1) ? 0.0 : -(h * c->Ab[kb]);
            const double b0 = 1.0 - h * c->Ap[kb];
            const double c0 = (kb == nZ - 2) ? 0.0 : -(h * c->At[kb]);
            for (int k = kb + 1; k <= ke; k++) {
                const double a = -(h * c->Ab[k]), b = 1.0 - h * c->Ap[k];
                const double cc = (k == nZ - 2) ? 0.0 : -(h * c->At[k]);
                if (k == kb + 1) {
                    const double R = 1.0 / b;
                    ap[k] = a * R; cp[k] = cc * R;
                    za[k] = 0.0; zR[k] = R;
                } else {
                    const double R = 1.0 / (b - a * cp[k - 1]);
                    ap[k] = -(a * ap[k - 1]) * R; cp[k] = cc * R;
                    za[k] = a; zR[k] = R;
                }
                zC[k] = cp[k];
            }
            if (M >= 3) {
                zA2[ke - 1] = ap[ke - 1];
                zC2[ke - 1] = cp[ke - 1];
                for (int k = ke - 2; k >= kb + 1; k--) {
                    zA2[k] = ap[k] - cp[k] * zA2[k + 1];
                    zC2[k] = -(cp[k] * zC2[k + 1]);
                }
            }
            rA[2 * ch] = a0;
            if (M >= 3) {
                B[2 * ch] = b0 - c0 * zA2[kb + 1];
                Cc[2 * ch] = -(c0 * zC2[kb + 1]);
                zc0[ch] = c0;
            } else {
                B[2 * ch] = b0;
                Cc[2 * ch] = c0;
                zc0[ch] = 0.0;
            }
            rA[2 * ch + 1] = ap[ke];
            B[2 * ch + 1] = 1.0;
            Cc[2 * ch + 1] = cp[ke];
        }
        rg[0] = 1.0 / B[0];
        re[0] = Cc[0] * rg[0];
        for (int s2 = 1; s2 < 2 * C; s2++) {
            rg[s2] = 1.0 / (B[s2] - rA[s2] * re[s2 - 1]);
            re[s2] = Cc[s2] * rg[s2];
        }
        CHK(dalloc(c, &c->zf, zd.size() * c->es));
        if (c->mixed) {
            std::vector<float> zs(zd.begin(), zd.end());
            CHK(h2d(c, c->zf, zs.data(), zs.size() * sizeof(float)));
        } else {
            CHK(h2d(c, c->zf, zd.data(), zd.size() * sizeof(double)));
        }
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    if (p.precond == TBL_PC_FAST || p.helm_mode == TBL_HELM_DIRECT) {
        char msg[256];
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (fastpc_create(c->g, c->Ab.data(), c->At.data(), c->Ap.data(), c->s, c->mixed, &c->fast, msg, sizeof(msg)))
            return fail(c, TBL_EARG, msg);
    }
    // partials: [0, 6*nbr+8) for tile reductions (Gauss-Seidel uses 6 per tile);
    // GMRES puts the residual partials at 0 and the vector partials at nbr+8,
    // (kmax+1) groups of vec_blocks(N) each
    const size_t nbr = resid_blocks(c->g), nbv = vec_blocks(N);
    c->npart = 6 * nbr + 8 + (size_t)(p.kmax + 2) * (nbv + 8);
    c->npart = std::max(c->npart, nbr + 8 + (size_t)poisson_blocks(c->g));   // fused first dot fits
    c->npart = std::max(c->npart, 6 * (size_t)rhs_blocks(c->g, c->mixed));             // fused first residual pass fits
    CHK(dalloc(c, &c->part, c->npart * sizeof(double)));
    CHK(dalloc(c, &c->gm, sizeof(GmCtrl)));
    HIPCHK(c, hipHostMalloc((void **)&c->gm_h, sizeof(GmStatus), hipHostMallocDefault));
    CHK(dalloc(c, &c->gs, gs_ctrl_bytes()));
    HIPCHK(c, hipHostMalloc(&c->gs_h, gs_ctrl_bytes(), hipHostMallocDefault));
    const long M = (long)p.nX * p.nZ;
    CHK(dalloc(c, &c->sums, 9 * M * sizeof(double)));
    CHK(dalloc(c, &c->avg, 9 * M * sizeof(double)));
    CHK(dalloc(c, &c->Wt, M * sizeof(double)));
    CHK(dalloc(c, &c->Ume, p.nZ * sizeof(double)));
    CHK(dalloc(c, &c->uf, 3L * p.nY * p.nZ * sizeof(double)));
    CHK(dalloc(c, &c->bltkns, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->dstar, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->theta, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->tau, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->tausum, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->taumean, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->cf, p.nX * sizeof(double)));
    CHK(dalloc(c, &c->wall_w, (size_t)p.nX * p.nY * sizeof(double)));
    CHK(dalloc(c, &c->scan, 2 * (p.nZ + 2) * sizeof(double)));
    CHK(dalloc(c, &c->newk, p.nZ * sizeof(int)));
    CHK(dalloc(c, &c->rec, sizeof(RecDev)));
    CHK(dalloc(c, &c->guard, sizeof(int)));
    CHK(dalloc(c, &c->dd, (DD_NCOUNT + 5 * MF_BLK) * sizeof(double)));
    HIPCHK(c, hipHostMalloc((void **)&c->dh, DD_NCOUNT * sizeof(double), hipHostMallocDefault));
    RecDev r0 = {1.0, 1.0, 0};
    HIPCHK(c, hipMemcpyAsync(c->rec, &r0, sizeof(r0), hipMemcpyHostToDevice, c->s));
    std::vector<int> nk(p.nZ);
    for (int k = 0; k < p.nZ; k++) nk[k] = (k < p.nZ - 1) ? k + 1 : p.nZ - 1;
    HIPCHK(c, hipMemcpyAsync(c->newk, nk.data(), sizeof(int) * p.nZ, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

// message of the last tbl_create that failed on this thread: no context
// exists then, so tbl_last_error(NULL, ...) returns it
static thread_local std::string create_err;

// the parameter checks of tbl_create; nullptr if p is valid
static const char *params_error(const tbl_params *p)
{
    if (p->nX < 8 || p->nY < 8 || p->nZ < 8) return "nX, nY and nZ must each be at least 8";
    if (p->nY & 1) return "nY must be even (the periodic span's nY-2 rows are coloured red-black)";
    if ((double)p->nX * p->nY * p->nZ * 3.0 > 2.0e9) return "nX*nY*nZ*3 must not exceed 2e9 (32-bit in-field offsets)";
    if (!(p->dt > 0.0) || !(p->Re > 0.0)) return "dt and Re must be positive";
    if (p->kmax < 1 || p->kmax > TBL_KMAX_LIMIT) return "kmax must be in 1..64";
    if (p->i_rec < 2 || p->i_rec > p->nX - 1) return "i_rec must be in 2..nX-1";
    if (p->freq_tim_avg < 1) return "freq_tim_avg must be at least 1";
    if (p->wall_shear_window < 0) return "wall_shear_window must be 0 (off) or positive";
    for (int q = 0; q < 5; q++)
        if (p->planes[q] < 1 || p->planes[q] > p->nX) return "mass-flux planes must be in 1..nX";
    if (p->precision != TBL_FP64 && p->precision != TBL_MIXED) return "unknown precision";
    if (p->precond != TBL_PC_SSOR_RB && p->precond != TBL_PC_NONE && p->precond != TBL_PC_JACOBI &&
        p->precond != TBL_PC_FAST && p->precond != TBL_PC_SSOR_BLOCK)
        return "unknown preconditioner";
    if (p->helm_mode != TBL_HELM_POINT_RB && p->helm_mode != TBL_HELM_ZLINE_RB && p->helm_mode != TBL_HELM_DIRECT &&
        p->helm_mode != TBL_HELM_ZEBRA && p->helm_mode != TBL_HELM_LEX_SLAB)
        return "unknown helm_mode";
    if (p->slabs < 1 || p->slabs > 256) return "slabs must be in 1..256";
    if (p->tol_mode != TBL_TOL_REF && p->tol_mode != TBL_TOL_TRUE) return "unknown tol_mode";
    if (p->scheme != TBL_SCHEME_SETBL && p->scheme != TBL_SCHEME_DNSTBL) return "unknown convection scheme";
    return nullptr;
}

extern "C" int tbl_create(const tbl_params *p, tbl_ctx **out)
{
    if (!p || !out) {
        create_err = "tbl_create: NULL params or output pointer";
        return TBL_EARG;
    }
    *out = nullptr;
    if (const char *e = params_error(p)) {
        create_err = std::string("tbl_create: ") + e;
        return TBL_EARG;
    }
    tbl_ctx *c = new tbl_ctx();
    c->p = *p;
    c->mixed = (p->precision == TBL_MIXED);
    int rc = create_impl(c);
    if (rc) {
        create_err = "tbl_create: " + c->err;
        tbl_destroy(c);
        return rc;
    }
    create_err.clear();
    *out = c;
    return 0;
}

extern "C" int tbl_destroy(tbl_ctx *c)
{
    if (!c) return 0;
    if (c->s) hipStreamSynchronize(c->s);
    timer_flush(c);
    for (auto e : c->pool) hipEventDestroy(e);
    if (c->poll_ev) hipEventDestroy(c->poll_ev);
    void *dev[] = {c->mesh, c->Ut, c->Ualt, c->Utold, c->Q, c->Press, c->PCorr, c->pRhs, c->V, c->W, c->scr, c->r, c->Z,
                   c->part, c->gm, c->gs, c->sums, c->avg, c->Wt, c->Ume, c->uf, c->bltkns, c->dstar,
                   c->theta, c->tau, c->tausum, c->taumean, c->cf, c->wall_w, c->scan, c->newk, c->rec, c->guard, c->dd, c->zf, c->br, c->PCprev,
                   c->PCalt, c->slab, c->halo};
    for (void *d : dev)
        if (d) hipFree(d);
    if (c->fast) fastpc_destroy(c->fast);
    if (c->gm_h) hipHostFree(c->gm_h);
    if (c->gs_h) hipHostFree(c->gs_h);
    if (c->dh) hipHostFree(c->dh);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
    return 0;
}

extern "C" int tbl_last_error(const tbl_ctx *c, char *buf, int len)
{
    if (!buf || len <= 0) return TBL_EARG;
    snprintf(buf, len, "%s", c ? c->err.c_str() : create_err.c_str());
    return 0;
}

extern "C" int tbl_sync(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" double tbl_device_bytes(const tbl_ctx *c)
{
    if (!c) return 0.0;
    const tbl_params &p = c->p;
    const double N = (double)p.nX * p.nY * p.nZ, M = (double)p.nX * p.nZ;
    double b = (3 * 3 + 1) * N * c->es + 2 * N * 8 + ((double)(p.kmax + 1) * c->ldv + 3 * N) * c->es;
    if (c->Ualt) b += 3 * N * c->es;
    b += (18 * M + M) * 8 + c->npart * 8 + sizeof(GmCtrl);
    return b;
}

// ------------------------------------------------------------------ fields
static size_t field_len(const tbl_ctx *c, int f)
{
    const size_t N = (size_t)c->g.N;
    return (f == TBL_F_UT || f == TBL_F_UTOLD || f == TBL_F_Q) ? 3 * N : N;
}
static void *field_ptr(tbl_ctx *c, int f, bool *is_double)
{
    *is_double = !c->mixed;
    switch (f) {
    case TBL_F_UT: return c->Ut;
    case TBL_F_UTOLD: return c->Utold;
    case TBL_F_Q: return c->Q;
    case TBL_F_PRESS: return c->Press;
    case TBL_F_WORK: return c->r;        // work vector whose boundary entries are never read
    case TBL_F_PCORR: *is_double = true; return c->PCorr;
    case TBL_F_PRHS: *is_double = true; return c->pRhs;
    }
    return nullptr;
}

static int put_field(tbl_ctx *c, void *dst, bool dbl, const double *host, size_t n)
{
    if (dbl) {
        HIPCHK(c, hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
    } else {
        std::vector<float> tmp(n);
        for (size_t e = 0; e < n; e++) tmp[e] = (float)host[e];
        HIPCHK(c, hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    return 0;
}
static int get_field(tbl_ctx *c, const void *src, bool dbl, double *host, size_t n)
{
    if (dbl) {
        HIPCHK(c, hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        for (size_t e = 0; e < n; e++) host[e] = tmp[e];
    }
    return 0;
}

extern "C" int tbl_field_set(tbl_ctx *c, int f, const double *host)
{
    if (c && f == TBL_F_PCORR) c->pc_hist = 0;
    if (!c || !host || f < 0 || f > TBL_F_WORK) return TBL_EARG;
    bool dbl;
    void *d = field_ptr(c, f, &dbl);
    CHK(put_field(c, d, dbl, host, field_len(c, f)));
    if (f == TBL_F_UT) c->uploaded = true;
    if (f == TBL_F_WORK) {   // keep the work vector's faces zero (see op_impl)
        if (c->mixed) launch_zero_faces<float>(c->g, (float *)d, c->s);
        else launch_zero_faces<double>(c->g, (double *)d, c->s);
        HIPCHK(c, hipStreamSynchronize(c->s));
    }
    return 0;
}
extern "C" int tbl_field_get(tbl_ctx *c, int f, double *host)
{
    if (!c || !host || f < 0 || f > TBL_F_WORK) return TBL_EARG;
    bool dbl;
    void *d = field_ptr(c, f, &dbl);
    return get_field(c, d, dbl, host, field_len(c, f));
}
extern "C" int tbl_upload(tbl_ctx *c, const double *Ut, const double *Press, const double *PCorr)
{
    if (!c) return TBL_EARG;
    if (Ut) CHK(tbl_field_set(c, TBL_F_UT, Ut));
    if (Press) CHK(tbl_field_set(c, TBL_F_PRESS, Press));
    if (PCorr) CHK(tbl_field_set(c, TBL_F_PCORR, PCorr));
    if (PCorr) c->pc_hist = 0;
    return 0;
}
extern "C" int tbl_download(tbl_ctx *c, double *Ut, double *Press, double *PCorr)
{
    if (!c) return TBL_EARG;
    if (Ut) CHK(tbl_field_get(c, TBL_F_UT, Ut));
    if (Press) CHK(tbl_field_get(c, TBL_F_PRESS, Press));
    if (PCorr) CHK(tbl_field_get(c, TBL_F_PCORR, PCorr));
    return 0;
}

static int h2d(tbl_ctx *c, void *d, const void *h, size_t b)
{
    HIPCHK(c, hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, c->s));
    return 0;
}
static int d2h(tbl_ctx *c, void *h, const void *d, size_t b)
{
    HIPCHK(c, hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, c->s));
    return 0;
}

// wait for the stream before the host reads a stopping test: spin on an event
// rather than block in a stream synchronise, since the next launch waits on
// the host's wake-up.  The first POLL_SPIN_US microseconds spin tightly; after
// that every query yields the core (sched_yield), so a long wait does not keep
// a host core from the caller's other threads.
// poll_mark / poll_spin split the wait so that work gated on the device's
// own stopping flags can be queued behind the read-back: the GPU then runs it
// while the host waits for the flags.
#define POLL_SPIN_US 20
static int poll_mark(tbl_ctx *c)
{
    if (!c->poll_ev) HIPCHK(c, hipEventCreateWithFlags(&c->poll_ev, hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(c->poll_ev, c->s));
    return 0;
}
static int poll_spin(tbl_ctx *c)
{
    hipError_t e;
    const auto t0 = std::chrono::steady_clock::now();
    bool spin = true;
    while ((e = hipEventQuery(c->poll_ev)) == hipErrorNotReady) {
        if (spin) spin = std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(POLL_SPIN_US);
        else sched_yield();
    }
    HIPCHK(c, e);
    return 0;
}
static int poll_wait(tbl_ctx *c)
{
    CHK(poll_mark(c));
    return poll_spin(c);
}

extern "C" int tbl_set_state(tbl_ctx *c, const tbl_state *st)
{
    if (!c || !st) return TBL_EARG;
    const tbl_params &p = c->p;
    const size_t M = (size_t)p.nX * p.nZ;
    if (st->sums) CHK(h2d(c, c->sums, st->sums, 9 * M * 8));
    if (st->avg) CHK(h2d(c, c->avg, st->avg, 9 * M * 8));
    if (st->Wt) CHK(h2d(c, c->Wt, st->Wt, M * 8));
    if (st->newk) {
        for (int k = 0; k < p.nZ; k++)
            if (st->newk[k] < 1 || st->newk[k] > p.nZ) return fail(c, TBL_EARG, "newk out of range");
        CHK(h2d(c, c->newk, st->newk, p.nZ * sizeof(int)));
    }
    if (st->Ume) CHK(h2d(c, c->Ume, st->Ume, p.nZ * 8));
    if (st->Ufluc_rec) CHK(h2d(c, c->uf, st->Ufluc_rec, 3L * p.nY * p.nZ * 8));
    if (st->bltkns) CHK(h2d(c, c->bltkns, st->bltkns, p.nX * 8));
    if (st->deltastar) CHK(h2d(c, c->dstar, st->deltastar, p.nX * 8));
    if (st->theta) CHK(h2d(c, c->theta, st->theta, p.nX * 8));
    if (st->wall_w) {
        CHK(h2d(c, c->wall_w, st->wall_w, (size_t)p.nX * p.nY * 8));
        c->have_wall_w = true;
    }
    RecDev r = {st->lamda, st->lamda03, 0};
    CHK(h2d(c, c->rec, &r, sizeof(r)));
    c->time = st->time;
    c->nstep = st->nstep;
    c->cnt = st->cnt;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
}

extern "C" int tbl_get_state(tbl_ctx *c, tbl_state *st)
{
    if (!c || !st) return TBL_EARG;
    const tbl_params &p = c->p;
    const size_t M = (size_t)p.nX * p.nZ;
    if (st->sums) CHK(d2h(c, st->sums, c->sums, 9 * M * 8));
    if (st->avg) CHK(d2h(c, st->avg, c->avg, 9 * M * 8));
    if (st->Wt) CHK(d2h(c, st->Wt, c->Wt, M * 8));
    if (st->newk) CHK(d2h(c, st->newk, c->newk, p.nZ * sizeof(int)));
    if (st->Ume) CHK(d2h(c, st->Ume, c->Ume, p.nZ * 8));
    if (st->Ufluc_rec) CHK(d2h(c, st->Ufluc_rec, c->uf, 3L * p.nY * p.nZ * 8));
    if (st->bltkns) CHK(d2h(c, st->bltkns, c->bltkns, p.nX * 8));
    if (st->deltastar) CHK(d2h(c, st->deltastar, c->dstar, p.nX * 8));
    if (st->theta) CHK(d2h(c, st->theta, c->theta, p.nX * 8));
    if (st->wall_w) CHK(d2h(c, st->wall_w, c->wall_w, (size_t)p.nX * p.nY * 8));
    RecDev r;
    CHK(d2h(c, &r, c->rec, sizeof(r)));
    HIPCHK(c, hipStreamSynchronize(c->s));
    st->lamda = r.lamda;
    st->lamda03 = r.lamda03;
    st->time = c->time;
    st->nstep = c->nstep;
    st->cnt = c->cnt;
    return 0;
}

extern "C" int tbl_get_mesh(const tbl_ctx *c, double *x1, double *x2, double *x3, double *Jac, double *Jac2,
                            double *Ab, double *At, double *Ap)
{
    if (!c) return TBL_EARG;
    auto cp = [](double *dst, const std::vector<double> &v) { if (dst) memcpy(dst, v.data(), v.size() * sizeof(double)); };
    cp(x1, c->x1); cp(x2, c->x2); cp(x3, c->x3); cp(Jac, c->Jac); cp(Jac2, c->Jac2);
    cp(Ab, c->Ab); cp(At, c->At); cp(Ap, c->Ap);
    return 0;
}

// ------------------------------------------------------------------ checkpoint
// Binary checkpoint of everything a step reads from earlier steps: the
// fields in their storage precision (bit for bit), the pressure-correction
// warm start (and the previous one of the extrapolated guess), the
// statistics and recycling state, the wall-shear window and the clock.
// Layout: 8-byte magic, the int32 header below, the double time, then one
// record per buffer (uint64 byte count, then the bytes), in a fixed order.
static const char CKPT_MAGIC[8] = {'T', 'B', 'L', 'C', 'K', 'P', 'T', '1'};
enum { CKPT_VERSION = 1 };
struct CkptHead {
    int version, nX, nY, nZ, precision, pc_hist, have_wall_w, nstep, cnt, nwin;
};
struct CkptRec {
    void *dev;
    size_t bytes;
};
static std::vector<CkptRec> ckpt_records(tbl_ctx *c)
{
    const tbl_params &p = c->p;
    const size_t N = (size_t)c->g.N, M = (size_t)p.nX * p.nZ;
    return {{c->Ut, 3 * N * c->es}, {c->Press, N * c->es}, {c->PCorr, N * 8},
            {c->PCprev, c->PCprev ? N * 8 : 0}, {c->sums, 9 * M * 8}, {c->avg, 9 * M * 8}, {c->Wt, M * 8},
            {c->Ume, (size_t)p.nZ * 8}, {c->uf, 3 * (size_t)p.nY * p.nZ * 8}, {c->bltkns, (size_t)p.nX * 8},
            {c->dstar, (size_t)p.nX * 8}, {c->theta, (size_t)p.nX * 8}, {c->wall_w, (size_t)p.nX * p.nY * 8},
            {c->newk, (size_t)p.nZ * sizeof(int)}, {c->rec, sizeof(RecDev)}, {c->tausum, (size_t)p.nX * 8},
            {c->taumean, (size_t)p.nX * 8}, {c->cf, (size_t)p.nX * 8}};
}
#define CKPT_CHUNK ((size_t)1 << 28)

extern "C" int tbl_checkpoint_save(tbl_ctx *c, const char *path)
{
    if (!c || !path) return TBL_EARG;
    if (!c->uploaded) return fail(c, TBL_ESTATE, "tbl_checkpoint_save before tbl_upload");
    HIPCHK(c, hipStreamSynchronize(c->s));
    FILE *f = fopen(path, "wb");
    if (!f) return fail(c, TBL_EARG, "tbl_checkpoint_save: cannot open the file for writing");
    CkptHead h = {CKPT_VERSION, c->p.nX, c->p.nY, c->p.nZ, c->p.precision, c->pc_hist, c->have_wall_w ? 1 : 0,
                  c->nstep, c->cnt, c->nwin};
    bool ok = fwrite(CKPT_MAGIC, 8, 1, f) == 1 && fwrite(&h, sizeof(h), 1, f) == 1 &&
              fwrite(&c->time, sizeof(double), 1, f) == 1;
    std::vector<char> buf;
    for (const CkptRec &r : ckpt_records(c)) {
        const uint64_t nb = r.bytes;
        ok = ok && fwrite(&nb, sizeof(nb), 1, f) == 1;
        for (size_t o = 0; ok && o < r.bytes; o += CKPT_CHUNK) {
            const size_t n = std::min(CKPT_CHUNK, r.bytes - o);
            buf.resize(n);
            hipError_t e = hipMemcpy(buf.data(), (char *)r.dev + o, n, hipMemcpyDeviceToHost);
            if (e != hipSuccess) {
                fclose(f);
                c->err = std::string("tbl_checkpoint_save: ") + hipGetErrorString(e);
                return (int)e;
            }
            ok = fwrite(buf.data(), 1, n, f) == n;
        }
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : fail(c, TBL_EARG, "tbl_checkpoint_save: write failed");
}

extern "C" int tbl_checkpoint_load(tbl_ctx *c, const char *path)
{
    if (!c || !path) return TBL_EARG;
    FILE *f = fopen(path, "rb");
    if (!f) return fail(c, TBL_EARG, "tbl_checkpoint_load: cannot open the file");
    char magic[8];
    CkptHead h;
    double t = 0.0;
    if (fread(magic, 8, 1, f) != 1 || memcmp(magic, CKPT_MAGIC, 8) != 0 || fread(&h, sizeof(h), 1, f) != 1 ||
        fread(&t, sizeof(t), 1, f) != 1 || h.version != CKPT_VERSION) {
        fclose(f);
        return fail(c, TBL_EARG, "tbl_checkpoint_load: not a checkpoint of this version");
    }
    if (h.nX != c->p.nX || h.nY != c->p.nY || h.nZ != c->p.nZ || h.precision != c->p.precision) {
        fclose(f);
        return fail(c, TBL_EARG, "tbl_checkpoint_load: grid or precision differs from the context's");
    }
    HIPCHK(c, hipStreamSynchronize(c->s));
    std::vector<char> buf;
    const size_t N = (size_t)c->g.N;
    for (int q = 0; q < (int)ckpt_records(c).size(); q++) {
        uint64_t nb = 0;
        if (fread(&nb, sizeof(nb), 1, f) != 1) { fclose(f); return fail(c, TBL_EARG, "tbl_checkpoint_load: truncated file"); }
        if (q == 3 && nb > 0 && !c->PCprev) {          // the extrapolation history
            int rc = dalloc(c, &c->PCprev, N * sizeof(double));
            if (rc) { fclose(f); return rc; }
            HIPCHK(c, hipStreamSynchronize(c->s));
        }
        const CkptRec r = ckpt_records(c)[q];
        if (nb != r.bytes) { fclose(f); return fail(c, TBL_EARG, "tbl_checkpoint_load: record size mismatch"); }
        for (size_t o = 0; o < r.bytes; o += CKPT_CHUNK) {
            const size_t n = std::min(CKPT_CHUNK, r.bytes - o);
            buf.resize(n);
            if (fread(buf.data(), 1, n, f) != n) { fclose(f); return fail(c, TBL_EARG, "tbl_checkpoint_load: truncated file"); }
            hipError_t e = hipMemcpy((char *)r.dev + o, buf.data(), n, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                fclose(f);
                c->err = std::string("tbl_checkpoint_load: ") + hipGetErrorString(e);
                return (int)e;
            }
        }
    }
    fclose(f);
    c->pc_hist = h.pc_hist;
    c->have_wall_w = h.have_wall_w != 0;
    c->nstep = h.nstep;
    c->cnt = h.cnt;
    c->nwin = h.nwin;
    c->time = t;
    c->uploaded = true;
    return 0;
}

// ------------------------------------------------------------------ blocks
template <class T> static T *F(void *p) { return (T *)p; }

extern "C" int tbl_begin_step(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    if (!c->uploaded) return fail(c, TBL_ESTATE, "tbl_begin_step before tbl_upload");
    c->time = c->time + c->p.dt;
    c->nstep += 1;
    c->cnt += 1;
    // Utold = Ut (:257); the step reads Utold only at the inflow plane of u
    // (the (u - u_old)/dt term of the inflow pressure, :791), so only that
    // plane is saved
    if (c->mixed) launch_save_inflow<float>(c->g, F<float>(c->Ut), F<float>(c->Utold), c->s);
    else launch_save_inflow<double>(c->g, F<double>(c->Ut), F<double>(c->Utold), c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}

template <class T> static int recycle_impl(tbl_ctx *c)
{
    double eps = c->nstep * (5.0e-5);
    if (eps > 1.0) eps = 1.0;
    if (!(c->nstep > c->p.freq_tim_avg)) return 0;
    Tscope ts(c, TBL_K_RECYCLE, 6.0 * c->p.nY * c->p.nZ * 8);
    launch_recycle<T>(c->g, F<T>(c->Ut), c->p.i_rec - 1, eps, &c->rec->lamda, c->avg, c->Wt, c->newk, c->Ume,
                      c->uf, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_recycle_inflow(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? recycle_impl<float>(c) : recycle_impl<double>(c);
}

template <class T> static int rhs_impl(tbl_ctx *c, double *hpart = nullptr)
{
    Tscope ts(c, TBL_K_RHS, 7.0 * (double)c->g.N * sizeof(T));
    launch_momentum_rhs<T>(c->g, F<T>(c->Ut), F<T>(c->Press), F<T>(c->Q), c->s, c->br_in, c->br_out, hpart);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_momentum_rhs(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? rhs_impl<float>(c) : rhs_impl<double>(c);
}

// exact Helmholtz solve per component: u += (I - hA)^{-1} (Q - H u), by fast
// diagonalisation; in FP64 twice (one refinement step on the residual).  In
// FP32 storage once: I - hA is well conditioned (the stretched grid's
// largest h lambda is about 2), so one FP32 fast solve is already at the
// FP32 rounding of the stored field, which a refinement cannot improve
template <class T> static int helm_direct(tbl_ctx *c, int *sweeps, double *res)
{
    if (!c->fast) return fail(c, TBL_ESTATE, "direct Helmholtz solve without the fast solver");
    const Geo &g = c->g;
    const long N = g.N;
    T *r = F<T>(c->r), *d = F<T>(c->scr);
    const double h = c->p.dt / c->p.Re;
    const int passes = sizeof(T) == sizeof(double) ? 2 : 1;
    launch_zero_faces<T>(g, r, c->s);
    for (int p = 0; p < 3; p++) {
        T *u = F<T>(c->Ut) + p * N;
        const T *q = F<T>(c->Q) + p * N;
        for (int it = 0; it < passes; it++) {
            Tscope ts(c, TBL_K_HELM, 4.0 * (double)N * sizeof(T));
            launch_helm_rfield<T>(g, q, u, r, c->s);
            if (run_fastpc_ext<T>(c->fast, g, r, d, nullptr, 1.0, -h, false, c->s))
                return fail(c, TBL_EARG, "fast solver failure in the direct Helmholtz solve");
            launch_add_interior<T>(g, u, d, c->s);
        }
    }
    HIPCHK(c, hipGetLastError());
    if (sweeps) for (int p = 0; p < 3; p++) sweeps[p] = passes;
    if (res) for (int p = 0; p < 3; p++) res[p] = 0.0;
    return 0;
}

// pre_nb > 0: the first residual pass was formed by the momentum kernel
// (pre_nb block partials in c->part)
// ppe_done (whole step only): the PPE right-hand side of the solution is
// queued behind each stopping-test read-back, gated on the device's active
// flags, and *ppe_done says whether the gated launch that ran was the final
// one (off while the PPE right-hand side's timer class is timed)
template <class T> static int helm_impl(tbl_ctx *c, double errorstop, int mmax, int *sweeps, double *res, int pre_nb = 0,
                                        bool *ppe_done = nullptr)
{
    if (ppe_done) *ppe_done = false;
    if (c->p.helm_mode == TBL_HELM_DIRECT) return helm_direct<T>(c, sweeps, res);
    const bool ppe_ahead = ppe_done && !(c->tmask & (1u << TBL_K_DIV));
    gs_ctrl_init(c->gs_h, 7);
    // stream-ordered copy from the pinned control block; the host next
    // touches gs_h after a synchronised read-back
    CHK(h2d(c, c->gs, c->gs_h, gs_ctrl_bytes()));
    // the z-line sweep converges in a few iterations: poll the stopping test
    // of every residual pass (odd it >= 3) before launching the sweeps, so
    // that no sweep is launched for components that have all stopped; point
    // sweeps poll every 8 iterations on small grids
    const bool zebra = c->p.helm_mode == TBL_HELM_ZEBRA;
    const bool every = (c->g.N > 2000000 || c->p.helm_mode == TBL_HELM_ZLINE_RB || zebra ||
                        c->p.helm_mode == TBL_HELM_LEX_SLAB);
    // zebra: sweeps read cur and write alt, which then swap; the boundary and
    // ghost values (not changed by the solve) are copied into alt first
    T *cur = F<T>(c->Ut), *alt = zebra ? F<T>(c->Ualt) : nullptr;
    if (zebra) launch_copy_faces3<T>(c->g, cur, alt, c->s);
    int act[3] = {1, 1, 1}, sw[3] = {0, 0, 0};
    double rs[3] = {0, 0, 0};
    bool fresh = false;                        // gs_h holds the final control block
    int it;
    for (it = 1; it <= mmax; it++) {
        if (it % 2 == 1) {
            const bool pre = (it == 1 && pre_nb > 0);
            if (!pre) {
                Tscope ts(c, TBL_K_HELM_RES, 2.0 * (act[0] + act[1] + act[2]) * (double)c->g.N * sizeof(T));
                launch_helm_resid<T>(c->g, F<T>(c->Q), zebra ? cur : F<T>(c->Ut), c->gs, c->part, it == 1 ? 1 : 0,
                                     c->s);
            }
            launch_helm_fin(c->g, c->gs, c->part, it, errorstop, c->mixed ? 32.0 * 1.1920928955078125e-7 : 0.0, c->s,
                            pre ? pre_nb : 0);
            if (it >= 3 && (every || it % 8 == 1)) {
                CHK(d2h(c, c->gs_h, c->gs, gs_ctrl_bytes()));
                CHK(poll_mark(c));
                if (ppe_ahead) launch_ppe_rhs<T>(c->g, zebra ? cur : F<T>(c->Ut), c->pRhs, c->s, c->gs);
                CHK(poll_spin(c));
                gs_ctrl_read(c->gs_h, act, sw, rs);
                if (!act[0] && !act[1] && !act[2]) {
                    fresh = true;
                    if (ppe_ahead) *ppe_done = true;
                    break;
                }
            }
        }
        // one timed launch per colour; algorithmic bytes of a colour pass: read
        // Q and the other colour's u, write this colour's u, i.e. 1.5 N values
        // per active component (the host's view of the active set is exact
        // when it polls every residual pass, an upper bound otherwise)
        if (zebra) {
            // algorithmic bytes of a zebra pass: Q, the pass's rows (x
            // neighbours), the other rows (span neighbours) and the new rows,
            // each half the grid: 2 N values per active component; an ended
            // component is copied through (N values)
            const int na = act[0] + act[1] + act[2];
            {
                // one timed scope for the sweep's two passes (launched back to
                // back): each event record leaves a gap of a few microseconds
                Tscope ts(c, TBL_K_HELM, 2.0 * (2.0 * na + 1.0 * (3 - na)) * (double)c->g.N * sizeof(T), 2);
                for (int colour = 0; colour < 2; colour++)
                    launch_helm_zebra<T>(c->g, F<T>(c->Q), cur, colour == 0 ? cur : alt, alt, F<T>(c->zf), c->gs,
                                         colour, c->s);
            }
            std::swap(cur, alt);
            HIPCHK(c, hipGetLastError());
            continue;
        }
        if (c->p.helm_mode == TBL_HELM_LEX_SLAB) {
            // one lexicographic sweep per slab: read Q, read and write u (the
            // neighbours come from registers, LDS and cache): 3 N values per
            // active component
            Tscope ts(c, TBL_K_HELM, 3.0 * (act[0] + act[1] + act[2]) * (double)c->g.N * sizeof(T));
            launch_helm_lex<T>(c->g, F<T>(c->Q), F<T>(c->Ut), F<T>(c->halo), c->gs, c->g.nkmax, c->s);
            HIPCHK(c, hipGetLastError());
            continue;
        }
        for (int colour = 0; colour < 2; colour++) {
            Tscope ts(c, TBL_K_HELM, 1.5 * (act[0] + act[1] + act[2]) * (double)c->g.N * sizeof(T));
            if (c->p.helm_mode == TBL_HELM_ZLINE_RB)
                launch_helm_zline<T>(c->g, F<T>(c->Q), F<T>(c->Ut), F<T>(c->zf), c->gs, colour, c->s);
            else
                launch_helm_sweep<T>(c->g, F<T>(c->Q), F<T>(c->Ut), c->gs, colour, c->s);
        }
        HIPCHK(c, hipGetLastError());
    }
    if (!fresh) {
        CHK(d2h(c, c->gs_h, c->gs, gs_ctrl_bytes()));
        CHK(poll_wait(c));
        gs_ctrl_read(c->gs_h, act, sw, rs);
    }
    if (zebra && cur != F<T>(c->Ut)) std::swap(c->Ut, c->Ualt);   // the solution is in cur
    bool capped = false;
    for (int p = 0; p < 3; p++)
        if (act[p]) { sw[p] = mmax; capped = true; }
    if (sweeps) for (int p = 0; p < 3; p++) sweeps[p] = sw[p];
    if (res) for (int p = 0; p < 3; p++) res[p] = rs[p];
    if (capped && c->p.fail_on_noconv) return fail(c, TBL_ENOCONV, "Gauss-Seidel reached gs_mmax");
    return 0;
}
extern "C" int tbl_diffusion_solve(tbl_ctx *c, double errorstop, int mmax, int *sweeps3, double *res3)
{
    if (!c || mmax < 1) return TBL_EARG;
    return c->mixed ? helm_impl<float>(c, errorstop, mmax, sweeps3, res3)
                    : helm_impl<double>(c, errorstop, mmax, sweeps3, res3);
}

template <class T> static int prhs_impl(tbl_ctx *c)
{
    Tscope ts(c, TBL_K_DIV, (double)c->g.N * (3.0 * sizeof(T) + 8.0));
    launch_ppe_rhs<T>(c->g, F<T>(c->Ut), c->pRhs, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_ppe_rhs(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? prhs_impl<float>(c) : prhs_impl<double>(c);
}

// algorithmic HBM bytes of one preconditioner application in storage type T:
// fast diagonalisation = four transforms and the line solve, each reading and
// writing the field once (10 N); block SSOR = forward (read r, write) and
// backward (read, write) sweeps (4 N); red-black SSOR = forward (read r, write
// the half sweep) and backward (read r and the half sweep, write) (5 N);
// Jacobi / none = read and write (2 N)
static double precond_bytes(const tbl_ctx *c, size_t es)
{
    const double N = (double)c->g.N;
    switch (c->p.precond) {
    case TBL_PC_FAST: return 10.0 * N * es;
    case TBL_PC_SSOR_BLOCK: return 4.0 * N * es;
    case TBL_PC_SSOR_RB: return 5.0 * N * es;
    default: return 2.0 * N * es;
    }
}

// restarted GMRES; each cycle forms the true residual in double (outer
// refinement) and runs at most kmax Arnoldi steps in the storage type T.
template <class T>
static int gmres_impl(tbl_ctx *c, double errtol, int kmax, int mmax, int *cycles, int *iters, double *true_res)
{
    const Geo &g = c->g;
    const long N = g.N;
    const unsigned nbv = vec_blocks(N), nbr = resid_blocks(g);
    T *V = F<T>(c->V), *W = F<T>(c->W), *scr = F<T>(c->scr), *r = F<T>(c->r);
    const long ldv = c->ldv;
    if (kmax > c->p.kmax) return fail(c, TBL_EARG, "kmax above the allocated restart length");
    // control block, set up on the device
    const int mode = (c->p.tol_mode == TBL_TOL_TRUE) ? GM_MODE_TRUE : GM_MODE_REF;
    launch_gm_init(c->gm, errtol, c->p.inner_tol, kmax, mmax, mode, c->s);
    launch_zero_faces(g, c->PCorr, c->s);
    double *part_r = c->part;                  // residual partials
    double *part_z = c->part + nbr + 8;        // norm partials (and dot partials)
    const bool fuse_dot = poisson_blocks(g) <= c->npart - nbr - 8;
    if (c->p.pc_extrap) {
        // initial guess 2 PCorr(n-1) - PCorr(n-2); the face values are zeroed below
        if (!c->PCprev) CHK(dalloc(c, &c->PCprev, (size_t)N * sizeof(double)));
        if (c->pc_hist >= 2) launch_pc_extrap(c->PCorr, c->PCprev, N, c->s);
        else HIPCHK(c, hipMemcpyAsync(c->PCprev, c->PCorr, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, c->s));
        c->pc_hist = c->pc_hist < 2 ? c->pc_hist + 1 : 2;
    }
    const int *act = &c->gm->st.active;
    const int poll = (N > 2000000 || c->p.precond == TBL_PC_FAST) ? 1 : 4;
    bool final_res = false;                    // st.true_res is the residual of the final PCorr
    if (mode == GM_MODE_TRUE) {
        // Flexible (right-preconditioned) GMRES: the Givens estimate is the true
        // residual, one preconditioner application per Arnoldi step, and the
        // correction x += sum y_i z_i uses the stored z_i = M^{-1} v_i.  Each
        // cycle starts from the double-precision residual (iterative refinement).
        T *Z = F<T>(c->Z);
        // A cycle that ends after one Arnoldi step (known exactly when the
        // host polls every step) leaves its correction PCorr += y z0 pending:
        // the next cycle's residual kernel forms it on the fly and writes the
        // updated PCorr to the second buffer, which then becomes PCorr.
        const bool fuse_upd = (poll == 1);
        // queue-ahead of the first transform (off while the preconditioner
        // class is timed, so that its timer brackets whole applications)
        const bool ahead = c->p.precond == TBL_PC_FAST && !(c->tmask & (1u << TBL_K_PRECOND));
        // a cycle that stops after its first Arnoldi step goes on with
        // fin_cycle and the next cycle's fused residual pass: both are queued
        // behind that step's status read-back, gated on the cycle having ended
        // (off while the residual class is timed)
        const bool ahead_res = fuse_upd && !(c->tmask & (1u << TBL_K_RESID));
        if (fuse_upd) {
            if (!c->PCalt) CHK(dalloc(c, &c->PCalt, (size_t)N * sizeof(double)));
            launch_zero_faces(g, c->PCalt, c->s);
        }
        bool pend = false;
        bool queued = false;                   // the cycle's residual pass ran behind the last read-back
        for (int m = 0; m < mmax; m++) {
            if (queued) {
                queued = false;
            } else if (pend) {
                Tscope ts(c, TBL_K_RESID, (double)N * (24.0 + 2.0 * sizeof(T)));
                launch_resid_upd<T>(g, c->pRhs, c->PCorr, Z, c->gm, c->PCalt, V, part_r, c->s);
                std::swap(c->PCorr, c->PCalt);
                pend = false;
            } else {
                // the first residual also gives ||pRhs|| (partials after part_r)
                Tscope ts(c, TBL_K_RESID, (double)N * (16.0 + sizeof(T)));
                launch_resid<T>(g, c->pRhs, c->PCorr, V, part_r, c->s, m == 0 ? part_z : nullptr);
            }
            if (m == 0) launch_fin_bnorm(c->gm, part_z, (int)nbr, c->s);
            launch_fin_start(c->gm, part_r, (int)nbr, part_r, (int)nbr, m == 0 ? 1 : 0, c->s);
            CHK(d2h(c, c->gm_h, &c->gm->st, sizeof(GmStatus)));
            // the first Arnoldi step's first transform is queued behind the
            // status read-back, gated on st.active (0 exactly when the solve
            // has ended), so the GPU works while the host waits for the status
            CHK(poll_mark(c));
            if (ahead) launch_precond<T>(g, c->p.precond, c->fast, V, scr, Z, act, c->s, FPC_HEAD);
            CHK(poll_spin(c));
            if (c->gm_h->outer_done) { final_res = true; break; }
            bool one_step = false;             // the cycle stopped after its first Arnoldi step
            for (int j = 0; j < kmax; j++) {
                T *vj = V + (long)j * ldv, *vj1 = V + (long)(j + 1) * ldv, *zj = Z + (long)j * ldv;
                {
                    Tscope ts(c, TBL_K_PRECOND, precond_bytes(c, sizeof(T)));
                    launch_precond<T>(g, c->p.precond, c->fast, vj, scr, zj, act, c->s,
                                      (ahead && j == 0) ? FPC_TAIL : FPC_ALL);
                }
                if (j == 0 && fuse_dot) {
                    // first Arnoldi step: the Poisson apply also forms <A z_0, v_0>
                    unsigned nbp;
                    {
                        Tscope ts(c, TBL_K_POISSON, 3.0 * (double)N * sizeof(T));
                        nbp = launch_poisson<T>(g, zj, vj1, act, &c->gm->sigma[j], c->s, V, part_z);
                    }
                    launch_fin_dots(c->gm, part_z, (int)nbp, j, c->s);
                } else {
                    {
                        Tscope ts(c, TBL_K_POISSON, 2.0 * (double)N * sizeof(T));
                        launch_poisson<T>(g, zj, vj1, act, &c->gm->sigma[j], c->s);
                    }
                    {
                        Tscope ts(c, TBL_K_DOTS, (j + 2.0) * (double)N * sizeof(T));
                        launch_dots<T>(V, ldv, j, N, c->gm, part_z, c->s);
                    }
                    launch_fin_dots(c->gm, part_z, (int)nbv, j, c->s);
                }
                {
                    Tscope ts(c, TBL_K_UPDATE, (j + 3.0) * (double)N * sizeof(T));
                    launch_update<T>(V, ldv, j, N, c->gm, part_z, c->s);
                }
                launch_fin_update(c->gm, part_z, (int)nbv, j, c->s);
                HIPCHK(c, hipGetLastError());
                if ((j + 1) % poll == 0 && j + 1 < kmax) {
                    CHK(d2h(c, c->gm_h, &c->gm->st, sizeof(GmStatus)));
                    const bool qa = ahead_res && j == 0 && m + 1 < mmax;
                    CHK(poll_mark(c));
                    if (qa) {
                        launch_fin_cycle(c->gm, c->s, 1);
                        launch_resid_upd<T>(g, c->pRhs, c->PCorr, Z, c->gm, c->PCalt, V, part_r, c->s, 1);
                    }
                    CHK(poll_spin(c));
                    if (!c->gm_h->active) {
                        one_step = (j == 0);
                        if (qa) {
                            std::swap(c->PCorr, c->PCalt);
                            queued = true;
                        }
                        break;
                    }
                }
            }
            if (!queued) {                     // (queued: fin_cycle and PCorr += y z0 have run)
                launch_fin_cycle(c->gm, c->s);
                if (fuse_upd && one_step) {
                    pend = true;
                } else {
                    Tscope ts(c, TBL_K_XUPD);
                    launch_xupd<T>(Z, ldv, N, c->gm, c->PCorr, c->s);
                }
            }
            HIPCHK(c, hipGetLastError());
            // no read-back here: the next cycle's residual decides (its
            // synchronised status read covers this cycle), and the cycle cap
            // is the loop bound
        }
        if (pend) {                            // the cycle cap ended the loop
            Tscope ts(c, TBL_K_XUPD);
            launch_xupd<T>(Z, ldv, N, c->gm, c->PCorr, c->s);
        }
        if (!final_res) {
            CHK(d2h(c, c->gm_h, &c->gm->st, sizeof(GmStatus)));
            CHK(poll_wait(c));
        }
    }
    for (int m = 0; m < mmax && mode == GM_MODE_REF; m++) {
        {
            Tscope ts(c, TBL_K_RESID, (double)N * (16.0 + sizeof(T)));
            launch_resid<T>(g, c->pRhs, c->PCorr, r, part_r, c->s);
        }
        {
            Tscope ts(c, TBL_K_PRECOND, precond_bytes(c, sizeof(T)));
            launch_precond<T>(g, c->p.precond, c->fast, r, scr, V, nullptr, c->s);
        }
        launch_norm2<T>(V, N, part_z, nullptr, c->s);
        launch_fin_start(c->gm, part_r, (int)nbr, part_z, (int)nbv, m == 0 ? 1 : 0, c->s);
        HIPCHK(c, hipGetLastError());
        for (int j = 0; j < kmax; j++) {
            T *vj = V + (long)j * ldv, *vj1 = V + (long)(j + 1) * ldv;
            {
                Tscope ts(c, TBL_K_POISSON, 2.0 * (double)N * sizeof(T));
                launch_poisson<T>(g, vj, W, act, &c->gm->sigma[j], c->s);
            }
            {
                Tscope ts(c, TBL_K_PRECOND, precond_bytes(c, sizeof(T)));
                launch_precond<T>(g, c->p.precond, c->fast, W, scr, vj1, act, c->s);
            }
            {
                Tscope ts(c, TBL_K_DOTS, (j + 2.0) * (double)N * sizeof(T));
                launch_dots<T>(V, ldv, j, N, c->gm, part_z, c->s);
            }
            launch_fin_dots(c->gm, part_z, (int)nbv, j, c->s);
            {
                Tscope ts(c, TBL_K_UPDATE, (j + 3.0) * (double)N * sizeof(T));
                launch_update<T>(V, ldv, j, N, c->gm, part_z, c->s);
            }
            launch_fin_update(c->gm, part_z, (int)nbv, j, c->s);
            HIPCHK(c, hipGetLastError());
            if ((j + 1) % poll == 0 && j + 1 < kmax) {
                CHK(d2h(c, c->gm_h, &c->gm->st, sizeof(GmStatus)));
                CHK(poll_wait(c));
                if (!c->gm_h->active) break;
            }
        }
        launch_fin_cycle(c->gm, c->s);
        {
            Tscope ts(c, TBL_K_XUPD);
            launch_xupd<T>(V, ldv, N, c->gm, c->PCorr, c->s);
        }
        HIPCHK(c, hipGetLastError());
        CHK(d2h(c, c->gm_h, &c->gm->st, sizeof(GmStatus)));
        CHK(poll_wait(c));
        if (c->gm_h->outer_done) break;
    }
    GmStatus st = *c->gm_h;
    if (st.nan) return fail(c, TBL_ENAN, "non-finite residual in the pressure solve");
    if (st.zero_rhs) {
        // ||pRhs|| = 0 (true-residual mode): A is nonsingular, so PCorr = 0
        launch_fill(c->PCorr, 0.0, N, c->s);
        HIPCHK(c, hipGetLastError());
        final_res = false;
    }
    // true residual after the solve (the reference reports none; logged here).
    // A flexible solve that stopped at a cycle start has just formed it from
    // the same kernel and partials.
    double tr = st.true_res;
    if (!final_res) {
        launch_resid<double>(g, c->pRhs, c->PCorr, nullptr, part_r, c->s);
        launch_fin_sum(part_r, (int)nbr, c->dd + DD_BNORM, c->s);
        double r2 = 0.0;
        CHK(d2h(c, &r2, c->dd + DD_BNORM, sizeof(double)));
        CHK(poll_wait(c));
        tr = sqrt(r2);
    }
    if (cycles) *cycles = st.cycle;
    if (iters) *iters = st.total;
    if (true_res) *true_res = tr;
    c->gm_h->true_res = tr;
    if (!st.conv && c->p.fail_on_noconv) return fail(c, TBL_ENOCONV, "GMRES reached mmax");
    return 0;
}
extern "C" int tbl_ppe_solve(tbl_ctx *c, double errtol, int kmax, int mmax, int *cycles, int *iters, double *true_res)
{
    if (!c || kmax < 1 || mmax < 1) return TBL_EARG;
    return c->mixed ? gmres_impl<float>(c, errtol, kmax, mmax, cycles, iters, true_res)
                    : gmres_impl<double>(c, errtol, kmax, mmax, cycles, iters, true_res);
}

template <class T> static int project_impl(tbl_ctx *c)
{
    Tscope ts(c, TBL_K_PROJECT, (double)c->g.N * (8.0 + 8.0 * sizeof(T)));
    launch_project<T>(c->g, c->PCorr, F<T>(c->Press), F<T>(c->Ut), c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_project(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? project_impl<float>(c) : project_impl<double>(c);
}

template <class T> static int bcs_impl(tbl_ctx *c)
{
    // the boundary faces of the three velocity components and the pressure,
    // read and written
    const double faces = 2.0 * ((double)c->g.nX * c->g.nY + (double)c->g.nY * c->g.nZ + (double)c->g.nX * c->g.nZ);
    Tscope ts(c, TBL_K_BCS, 8.0 * faces * sizeof(T));
    launch_bcs<T>(c->g, F<T>(c->Ut), F<T>(c->Utold), F<T>(c->Press), c->have_wall_w ? c->wall_w : nullptr, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_apply_bcs(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? bcs_impl<float>(c) : bcs_impl<double>(c);
}

template <class T> static int div_impl(tbl_ctx *c)
{
    Tscope ts(c, TBL_K_DIAG, 3.0 * (double)c->g.N * sizeof(T));      // the three velocity components
    launch_div_check<T>(c->g, F<T>(c->Ut), c->part, c->s);
    launch_fin_sum(c->part, (int)resid_blocks(c->g), c->dd + DD_DIV, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
static double div_metric(const tbl_ctx *c, double s)
{
    const tbl_params &p = c->p;
    return fabs(s) / ((p.nY - 2) * (p.nX - 2) * (p.nZ - 2));
}
extern "C" int tbl_div_check(tbl_ctx *c, double *div, double *signed_sum)
{
    if (!c) return TBL_EARG;
    CHK(c->mixed ? div_impl<float>(c) : div_impl<double>(c));
    double s = 0.0;
    CHK(d2h(c, &s, c->dd + DD_DIV, sizeof(double)));
    CHK(poll_wait(c));
    if (signed_sum) *signed_sum = s;
    if (div) *div = div_metric(c, s);
    return 0;
}

// tbl_div_check on caller-owned device memory with the block partials in the
// static scratch: no allocation, no host synchronisation
extern "C" int tbl_div_check_noalloc(tbl_ctx *c, const void *Ut_dev, double *out_dev)
{
    if (!c) return TBL_EARG;
    if (!Ut_dev || !out_dev) return fail(c, TBL_EARG, "tbl_div_check_noalloc: NULL device pointer");
    if (!c->noalloc) return fail(c, TBL_ESTATE, "tbl_div_check_noalloc: static scratch not set up");
    const unsigned nb = resid_blocks(c->g);
    if (nb > TBL_NOALLOC_PARTS) {
        char m[160];
        snprintf(m, sizeof(m), "tbl_div_check_noalloc: this grid needs %u partials, the static scratch holds %d",
                 nb, TBL_NOALLOC_PARTS);
        return fail(c, TBL_EARG, m);
    }
    const tbl_params &p = c->p;
    const double nint = (double)((p.nY - 2) * (p.nX - 2) * (p.nZ - 2));
    if (c->mixed) launch_div_check<float>(c->g, (const float *)Ut_dev, c->noalloc, c->s);
    else launch_div_check<double>(c->g, (const double *)Ut_dev, c->noalloc, c->s);
    launch_fin_div(c->noalloc, (int)nb, nint, out_dev, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}

template <class T> static int mass_impl(tbl_ctx *c)
{
    Tscope ts(c, TBL_K_DIAG, 5.0 * (double)c->g.nY * c->g.nZ * sizeof(T));   // u on the five planes
    launch_mass_flux<T>(c->g, F<T>(c->Ut), c->p.planes, c->mass_denom, c->dd + DD_MASS, c->dd + DD_NCOUNT, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_mass_flux(tbl_ctx *c, double *mass5)
{
    if (!c) return TBL_EARG;
    CHK(c->mixed ? mass_impl<float>(c) : mass_impl<double>(c));
    if (mass5) {
        CHK(d2h(c, mass5, c->dd + DD_MASS, 5 * sizeof(double)));
        CHK(poll_wait(c));
    }
    return 0;
}

template <class T> static int stats_impl(tbl_ctx *c)
{
    // the velocity, and the nine (nX, nZ) sums read and written
    Tscope ts(c, TBL_K_DIAG, 3.0 * (double)c->g.N * sizeof(T) + 18.0 * (double)c->g.nX * c->g.nZ * 8.0);
    launch_stats<T>(c->g, F<T>(c->Ut), c->sums, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_stats_accumulate(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? stats_impl<float>(c) : stats_impl<double>(c);
}

template <class T> static int rescale_impl(tbl_ctx *c)
{
    Tscope ts(c, TBL_K_RESCALE);
    launch_rescale<T>(c->g, c->p.i_rec - 1, c->p.freq_tim_avg, F<T>(c->Ut), c->sums, c->avg, c->rec, c->Wt,
                      c->newk, c->bltkns, c->dstar, c->theta, c->uf, c->scan, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_rescale(tbl_ctx *c)
{
    if (!c) return TBL_EARG;
    return c->mixed ? rescale_impl<float>(c) : rescale_impl<double>(c);
}

template <class T> static int ke_impl(tbl_ctx *c)
{
    Tscope ts(c, TBL_K_DIAG, 3.0 * (double)c->g.N * sizeof(T));      // the velocity against the means
    launch_ke<T>(c->g, c->p.i_rec - 1, F<T>(c->Ut), c->avg, c->part, c->guard, c->uf, c->s);
    launch_fin_sum(c->part, (int)resid_blocks(c->g), c->dd + DD_KE, c->s);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_ke_guard(tbl_ctx *c, double *ke, int *guard)
{
    if (!c) return TBL_EARG;
    HIPCHK(c, hipMemsetAsync(c->guard, 0, sizeof(int), c->s));
    CHK(c->mixed ? ke_impl<float>(c) : ke_impl<double>(c));
    double s = 0.0;
    int gd = 0;
    CHK(d2h(c, &s, c->dd + DD_KE, sizeof(double)));
    CHK(d2h(c, &gd, c->guard, sizeof(int)));
    CHK(poll_wait(c));
    const tbl_params &p = c->p;
    if (ke) *ke = s / (double)(p.nX * (p.nY - 2) * (p.nZ - 2));
    if (guard) *guard = gd;
    return 0;
}

template <class T> static int tau_impl(tbl_ctx *c, double *acc = nullptr)
{
    Tscope ts(c, TBL_K_DIAG, 3.0 * (double)c->g.sZ * sizeof(T));     // u on the three wall planes
    launch_wall_shear<T>(c->g, F<T>(c->Ut), c->tau, c->s, acc);
    HIPCHK(c, hipGetLastError());
    return 0;
}
extern "C" int tbl_wall_shear(tbl_ctx *c, double *tau)
{
    if (!c) return TBL_EARG;
    CHK(c->mixed ? tau_impl<float>(c) : tau_impl<double>(c));
    if (tau) {
        CHK(d2h(c, tau, c->tau, c->p.nX * sizeof(double)));
        CHK(poll_wait(c));
    }
    return 0;
}

extern "C" int tbl_wall_shear_mean(tbl_ctx *c, double *tau_mean, double *cf, int *nwindows)
{
    if (!c) return TBL_EARG;
    if (c->p.wall_shear_window <= 0) return fail(c, TBL_ESTATE, "tbl_wall_shear_mean: wall_shear_window is 0 (off)");
    if (tau_mean) CHK(d2h(c, tau_mean, c->taumean, c->p.nX * sizeof(double)));
    if (cf) CHK(d2h(c, cf, c->cf, c->p.nX * sizeof(double)));
    CHK(poll_wait(c));
    if (nwindows) *nwindows = c->nwin;
    return 0;
}

// ------------------------------------------------------------------ whole step
extern "C" int tbl_step(tbl_ctx *c, int nsteps, tbl_diag *diag)
{
    if (!c || nsteps < 0) return TBL_EARG;
    const tbl_params &p = c->p;
    for (int n = 0; n < nsteps; n++) {
        tbl_diag d;
        memset(&d, 0, sizeof(d));
        bool ppe_done = false;                 // the diffusion solve queued the PPE right-hand side
        CHK(tbl_begin_step(c));
        CHK(tbl_recycle_inflow(c));
        if (p.helm_mode != TBL_HELM_DIRECT) {
            // the momentum kernel also forms the Gauss-Seidel solve's first
            // residual pass (same values, one read of the fields)
            const int nb = (int)rhs_blocks(c->g, c->mixed);
            CHK(c->mixed ? rhs_impl<float>(c, c->part) : rhs_impl<double>(c, c->part));
            HIPCHK(c, hipGetLastError());
            CHK(c->mixed ? helm_impl<float>(c, p.errorstop, p.gs_mmax, d.gs_sweeps, nullptr, nb, &ppe_done)
                         : helm_impl<double>(c, p.errorstop, p.gs_mmax, d.gs_sweeps, nullptr, nb, &ppe_done));
        } else {
            CHK(tbl_momentum_rhs(c));
            CHK(tbl_diffusion_solve(c, p.errorstop, p.gs_mmax, d.gs_sweeps, nullptr));
        }
        if (!ppe_done) CHK(tbl_ppe_rhs(c));
        CHK(tbl_ppe_solve(c, p.errtol, p.kmax, p.mmax, &d.gm_cycles, &d.gm_iters, &d.gm_true_res));
        d.gm_sumav = c->gm_h->sumav;
        CHK(tbl_project(c));
        CHK(tbl_apply_bcs(c));
        CHK(c->mixed ? div_impl<float>(c) : div_impl<double>(c));
        CHK(c->mixed ? mass_impl<float>(c) : mass_impl<double>(c));
        if (p.wall_shear_window > 0) {
            // a17 window: tau of every step into the running sum; mean and C_f
            // at the window's last step (dnstbl/dnstbl.f90:1666-1682)
            CHK(c->mixed ? tau_impl<float>(c, c->tausum) : tau_impl<double>(c, c->tausum));
            if (c->nstep % p.wall_shear_window == 0) {
                Tscope ts(c, TBL_K_DIAG);
                launch_tau_window(c->g, p.wall_shear_window, c->tausum, c->taumean, c->cf, c->s);
                HIPCHK(c, hipGetLastError());
                c->nwin += 1;
                d.cf_ready = 1;
            }
        }
        CHK(tbl_stats_accumulate(c));
        if (c->nstep % p.freq_tim_avg == 0) {
            CHK(tbl_rescale(c));
            d.rescaled = 1;
        }
        HIPCHK(c, hipMemsetAsync(c->guard, 0, sizeof(int), c->s));
        const bool ke_on = c->cnt > p.freq_tim_avg;
        if (ke_on) CHK(c->mixed ? ke_impl<float>(c) : ke_impl<double>(c));
        CHK(d2h(c, c->dh, c->dd, DD_NCOUNT * sizeof(double)));
        int gd = 0;
        CHK(d2h(c, &gd, c->guard, sizeof(int)));
        CHK(poll_wait(c));
        d.div_signed = c->dh[DD_DIV];
        d.div = div_metric(c, d.div_signed);
        for (int q = 0; q < 5; q++) d.mass[q] = c->dh[DD_MASS + q];
        d.ke = ke_on ? c->dh[DD_KE] / (double)(p.nX * (p.nY - 2) * (p.nZ - 2)) : 0.0;
        d.guard = gd;
        d.nstep = c->nstep;
        if (diag) diag[n] = d;
        // the guard is a bit set (1: fluctuation > 25, 2: non-finite), both
        // bits may be set by different nodes
        if ((gd & 2) || !(d.div_signed == d.div_signed)) return fail(c, TBL_ENAN, "non-finite field");
        if (gd & 1) return fail(c, TBL_EBLOWUP, "very large fluctuations (> 25)");
    }
    return 0;
}

// ------------------------------------------------------------------ discrete adjoint
// Work buffers of one adjoint window (allocated per call).
struct AdjWin {
    double *lamU = nullptr, *lamP = nullptr, *lamQ = nullptr, *lamPhi = nullptr, *lamB = nullptr;
    double *rhs = nullptr, *r = nullptr, *d = nullptr, *Uend = nullptr, *Utmp = nullptr, *grad = nullptr;
    double *Jd = nullptr;
    char *snapbuf = nullptr;             // nslots states: Ut, Press in the storage type, PCorr in double
    double *Ud = nullptr;                // mixed contexts: the FP64 copy of the linearisation velocity
    float *rf = nullptr, *df = nullptr;  // mixed contexts: FP32 residual and correction of the transposed solves
    size_t slot_bytes = 0;
    int nslots = 0;
    std::vector<int> sn, sc;
    std::vector<double> st;
    int i0 = 0, i1 = 0;
    double cJ = 0.0;
    int nfwd = 0, nadj = 0;
    int nstep0 = 0;
    int brmode = 0;          // 1 record the Peclet switches, 2 replay them (first sweep only)
};
static void adj_free(AdjWin &A)
{
    void *p[] = {A.lamU, A.lamP, A.lamQ, A.lamPhi, A.lamB, A.rhs, A.r, A.d, A.Uend, A.Utmp, A.grad, A.Jd, A.snapbuf, A.Ud, A.rf, A.df};
    for (void *q : p)
        if (q) hipFree(q);
}
// snapshot = Ut (3N), Press (N) in the storage type, PCorr (N, double) and
// the step counters
static int snap_store(tbl_ctx *c, AdjWin &A, int slot)
{
    const size_t N = (size_t)c->g.N, es = c->es;
    char *b = A.snapbuf + (size_t)slot * A.slot_bytes;
    HIPCHK(c, hipMemcpyAsync(b, c->Ut, 3 * N * es, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(b + 3 * N * es, c->Press, N * es, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(b + 4 * N * es, c->PCorr, N * 8, hipMemcpyDeviceToDevice, c->s));
    A.sn[slot] = c->nstep;
    A.sc[slot] = c->cnt;
    A.st[slot] = c->time;
    return 0;
}
static int snap_restore(tbl_ctx *c, AdjWin &A, int slot)
{
    const size_t N = (size_t)c->g.N, es = c->es;
    const char *b = A.snapbuf + (size_t)slot * A.slot_bytes;
    HIPCHK(c, hipMemcpyAsync(c->Ut, b, 3 * N * es, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->Press, b + 3 * N * es, N * es, hipMemcpyDeviceToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(c->PCorr, b + 4 * N * es, N * 8, hipMemcpyDeviceToDevice, c->s));
    c->nstep = A.sn[slot];
    c->cnt = A.sc[slot];
    c->time = A.st[slot];
    c->pc_hist = 0;
    return 0;
}
// exact pressure solve: PCorr += A^{-1} (pRhs - A PCorr), twice; in a mixed
// context with the residual formed in double and the inverse applied by the
// FP32 fast solver (iterative refinement: each pass gains about the FP32
// solver's accuracy, 1e-6, so two passes leave ~1e-12)
template <class T> static int ppe_direct(tbl_ctx *c)
{
    const Geo &g = c->g;
    T *r = F<T>(c->r), *d = F<T>(c->scr);
    launch_zero_faces<double>(g, c->PCorr, c->s);
    launch_zero_faces<T>(g, r, c->s);
    const int passes = 2;
    for (int it = 0; it < passes; it++) {
        launch_resid<T>(g, c->pRhs, c->PCorr, r, c->part, c->s);
        if (run_fastpc_ext<T>(c->fast, g, r, d, nullptr, 0.0, 1.0, false, c->s))
            return fail(c, TBL_EARG, "fast solver failure in the direct pressure solve");
        launch_add_interior<double, T>(g, c->PCorr, d, c->s);
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}
// one step of the window: the reference's step with exact solves, no statistics
static int win_step(tbl_ctx *c, AdjWin &A, bool accJ)
{
    CHK(tbl_begin_step(c));
    CHK(tbl_recycle_inflow(c));
    if (accJ && A.brmode) {
        unsigned char *b = c->br + (size_t)(c->nstep - A.nstep0 - 1) * (size_t)c->g.N;
        if (A.brmode == 1) c->br_out = b;
        else c->br_in = b;
    }
    const int rc = tbl_momentum_rhs(c);
    c->br_in = nullptr;
    c->br_out = nullptr;
    CHK(rc);
    CHK(c->mixed ? helm_direct<float>(c, nullptr, nullptr) : helm_direct<double>(c, nullptr, nullptr));
    CHK(tbl_ppe_rhs(c));
    CHK(c->mixed ? ppe_direct<float>(c) : ppe_direct<double>(c));
    CHK(tbl_project(c));
    CHK(tbl_apply_bcs(c));
    if (accJ) {
        if (c->mixed) launch_wall_shear<float>(c->g, (const float *)c->Ut, c->tau, c->s);
        else launch_wall_shear<double>(c->g, (const double *)c->Ut, c->tau, c->s);
        launch_tau_acc(c->g, c->tau, A.i0, A.i1, A.cJ, A.Jd, c->s);
    }
    HIPCHK(c, hipGetLastError());
    A.nfwd++;
    return 0;
}
// x = (alpha I + beta A)^{-T} b on the interior planes, by iterative
// refinement in FP64: x = M^{-T} b, then x += M^{-T} (b - (alpha I + beta A)^T x),
// M^{-T} the transposed fast solve.  FP64 contexts: the FP64 fast solver,
// two passes.  Mixed contexts: the FP32 fast solver on the FP64 residual
// rounded to FP32, also two passes (each gains about the FP32 solver's
// accuracy, 1e-6), x and the residual accumulated in FP64.  b is zero off the interior;
// x is written on planes 1..nZ-2 (zero on their face rows).
static int solve_T(tbl_ctx *c, AdjWin &A, const double *b, double *x, double alpha, double beta)
{
    const Geo &g = c->g;
    const FastPC *f = c->fast;
    if (!c->mixed) {
        if (run_fastpc_ext<double>(f, g, b, x, nullptr, alpha, beta, true, c->s)) return fail(c, TBL_EARG, "fast solver failure");
        launch_opT_res(g, b, x, alpha, beta, A.r, c->s);
        if (run_fastpc_ext<double>(f, g, A.r, A.d, nullptr, alpha, beta, true, c->s)) return fail(c, TBL_EARG, "fast solver failure");
        launch_add_interior<double>(g, x, A.d, c->s);
        return 0;
    }
    const long pst = g.sZ, nzp = (long)(g.nZ - 2) * pst;
    launch_convert<float, double>(b + pst, A.rf + pst, nzp, c->s);
    if (run_fastpc_ext<float>(f, g, A.rf, A.df, nullptr, alpha, beta, true, c->s)) return fail(c, TBL_EARG, "fast solver failure");
    launch_convert<double, float>(A.df + pst, x + pst, nzp, c->s);
    for (int it = 1; it < 2; it++) {
        launch_opT_res(g, b, x, alpha, beta, A.rf, c->s);
        if (run_fastpc_ext<float>(f, g, A.rf, A.df, nullptr, alpha, beta, true, c->s)) return fail(c, TBL_EARG, "fast solver failure");
        launch_add_interior<double, float>(g, x, A.df, c->s);
    }
    return 0;
}
// adjoint of step n: on entry the fields hold the state after step n-1,
// A.Uend the velocity after step n, A.lamU / A.lamP the adjoint of the state
// after step n; on exit they hold the adjoint of the state after step n-1
// (including dJ_{n-1}/dU) and A.Uend the velocity after step n-1.
// In a mixed context the states are FP32; the adjoint kernels read them
// converted to double (A.Ud, A.Uend), so the adjoint is accumulated in FP64
// about the mixed-precision trajectory.
static int adj_step(tbl_ctx *c, AdjWin &A, int n)
{
    const Geo &g = c->g;
    const size_t N = (size_t)g.N;
    const double h = c->p.dt / c->p.Re;
    HIPCHK(c, hipMemcpyAsync(A.Utmp, c->Ut, 3 * N * c->es, hipMemcpyDeviceToDevice, c->s));
    CHK(tbl_begin_step(c));
    CHK(tbl_recycle_inflow(c));               // linearisation point of the convection
    const double *Ulin = (const double *)c->Ut;
    if (c->mixed) {
        launch_convert<double, float>((const float *)c->Ut, A.Ud, 3 * (long)N, c->s);
        Ulin = A.Ud;
    }
    launch_adj_bcs(g, A.Uend, A.lamU, A.lamP, A.grad, c->s);
    launch_adj_proj(g, A.lamU, A.lamP, A.lamPhi, c->s);
    CHK(solve_T(c, A, A.lamPhi, A.lamB, 0.0, 1.0));
    launch_adj_div(g, A.lamB, A.lamU, c->s);
    for (int p = 0; p < 3; p++) {
        launch_copy_interior(g, A.lamU + p * N, A.rhs, c->s);
        CHK(solve_T(c, A, A.rhs, A.lamQ + p * N, 1.0, -h));
    }
    launch_adj_mom(g, Ulin, A.lamQ, A.lamU, A.lamP, c->s);
    if (c->mixed) launch_convert<double, float>((const float *)A.Utmp, A.Uend, 3 * (long)N, c->s);
    else HIPCHK(c, hipMemcpyAsync(A.Uend, A.Utmp, 3 * N * 8, hipMemcpyDeviceToDevice, c->s));
    if (n - 1 >= 1) launch_adj_tau(g, A.lamU, A.i0, A.i1, A.cJ, c->s);
    HIPCHK(c, hipGetLastError());
    A.nadj++;
    return 0;
}
static long long binom_capped(int n, int k)
{
    long long r = 1;
    for (int q = 1; q <= k; q++) {
        r = r * (n - k + q) / q;
        if (r > (1LL << 40)) return 1LL << 40;
    }
    return r;
}
// reverse steps a+1..b; the state after step a is in snapshot slot_a and
// `free` further slots (slot_a+1 ...) may be used.  Checkpoint placement of
// the binomial schedule: with r the least repetition number such that
// C(free + r, free) >= b - a, the segment after the new checkpoint m is
// reversible with free-1 snapshots at repetition r.
static int revolve_split(int a, int b, int free)
{
    const int L = b - a;
    int r = 0;
    while (binom_capped(free + r, free) < L) r++;
    const long long bl = binom_capped(free - 1 + r, free - 1);
    int m = b - (int)(bl < (long long)(L - 1) ? bl : (long long)(L - 1));
    return m <= a ? a + 1 : m;
}
static int revolve(tbl_ctx *c, AdjWin &A, int a, int b, int slot_a, int free)
{
    if (b - a == 1) {
        CHK(snap_restore(c, A, slot_a));
        return adj_step(c, A, b);
    }
    if (free == 0) {
        for (int n = b; n > a; n--) {
            CHK(snap_restore(c, A, slot_a));
            for (int m = a + 1; m <= n - 1; m++) CHK(win_step(c, A, false));
            CHK(adj_step(c, A, n));
        }
        return 0;
    }
    const int m = revolve_split(a, b, free);
    CHK(snap_restore(c, A, slot_a));
    for (int k = a + 1; k <= m; k++) CHK(win_step(c, A, false));
    CHK(snap_store(c, A, slot_a + 1));
    CHK(revolve(c, A, m, b, slot_a + 1, free - 1));
    return revolve(c, A, a, m, slot_a, free);
}
// The schedule's first descent (restore a, step to m, store, recurse on
// (m, b)) runs the same steps as the window's J sweep: the sweep stores those
// checkpoints as it passes them (first_descent), and the reverse sweep starts
// here, skipping the forward steps of that descent.
static void first_descent(int n, int nsnap, std::vector<int> &at)
{
    at.assign(n + 1, -1);                      // at[state index] = slot, or -1
    int a = 0, slot = 0, free = nsnap;
    while (n - a > 1 && free > 0) {
        const int m = revolve_split(a, n, free);
        at[m] = slot + 1;
        a = m;
        slot++;
        free--;
    }
}
static int revolve_top(tbl_ctx *c, AdjWin &A, int a, int b, int slot_a, int free)
{
    if (b - a == 1 || free == 0) return revolve(c, A, a, b, slot_a, free);
    const int m = revolve_split(a, b, free);   // stored in slot_a + 1 by the J sweep
    CHK(revolve_top(c, A, m, b, slot_a + 1, free - 1));
    return revolve(c, A, a, m, slot_a, free);
}

extern "C" int tbl_adjoint_window(tbl_ctx *c, const tbl_adjoint_opts *o, const double *wall_w, double *J, double *dJdw,
                                  int *nfwd, int *nadj)
{
    if (!c || !o) return TBL_EARG;
    const tbl_params &p = c->p;
    if (p.scheme != TBL_SCHEME_SETBL || p.forc[0] != 0.0 || p.forc[1] != 0.0 || p.forc[2] != 0.0)
        return fail(c, TBL_EARG, "the adjoint window differentiates the dnssetbl convection scheme without forcing");
    if (!c->uploaded) return fail(c, TBL_ESTATE, "tbl_adjoint_window before tbl_upload");
    if (o->nsteps < 1 || o->nsnap < 1 || o->i0 < 2 || o->i1 < o->i0 || o->i1 > p.nX - 1)
        return fail(c, TBL_EARG, "invalid adjoint window options");
    for (int n = c->nstep + 1; n <= c->nstep + o->nsteps; n++)
        if (n % p.freq_tim_avg == 0) return fail(c, TBL_EARG, "the adjoint window crosses a statistics rescale");
    if (!c->fast) {
        char msg[256];
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (fastpc_create(c->g, c->Ab.data(), c->At.data(), c->Ap.data(), c->s, c->mixed, &c->fast, msg, sizeof(msg)))
            return fail(c, TBL_EARG, msg);
    }
    if (wall_w) {
        CHK(h2d(c, c->wall_w, wall_w, (size_t)p.nX * p.nY * 8));
        c->have_wall_w = true;
    }
    const size_t N = (size_t)c->g.N;
    AdjWin A;
    A.nstep0 = c->nstep;
    A.brmode = o->branch_mode;
    if (o->branch_mode == 1) {
        if (c->br_steps < o->nsteps) {
            if (c->br) hipFree(c->br);
            c->br = nullptr;
            c->br_steps = 0;
            CHK(dalloc(c, &c->br, (size_t)o->nsteps * N));
            c->br_steps = o->nsteps;
        }
        c->br_nstep0 = c->nstep;
    } else if (o->branch_mode == 2) {
        if (!c->br || c->br_nstep0 != c->nstep || c->br_steps < o->nsteps)
            return fail(c, TBL_ESTATE, "no recorded Peclet-switch pattern for this window");
        if (!o->forward_only) return fail(c, TBL_EARG, "a replayed switch pattern is for forward evaluations only");
    } else if (o->branch_mode != 0) {
        return fail(c, TBL_EARG, "invalid branch_mode");
    }
    A.i0 = o->i0 - 1;
    A.i1 = o->i1 - 1;
    A.cJ = 1.0 / ((double)o->nsteps * (double)(o->i1 - o->i0 + 1) * (double)(p.nY - 2));
    A.nslots = o->forward_only ? 1 : o->nsnap + 2;       // start, checkpoints, final state
    A.sn.assign(A.nslots, 0);
    A.sc.assign(A.nslots, 0);
    A.st.assign(A.nslots, 0.0);
    int rc = 0;
    auto run = [&]() -> int {
        CHK(dalloc(c, &A.Jd, 8));
        A.slot_bytes = 4 * N * c->es + N * 8;
        CHK(dalloc(c, &A.snapbuf, (size_t)A.nslots * A.slot_bytes));
        CHK(snap_store(c, A, 0));
        std::vector<int> at;
        if (!o->forward_only) first_descent(o->nsteps, o->nsnap, at);
        for (int n = 1; n <= o->nsteps; n++) {
            CHK(win_step(c, A, true));
            if (!o->forward_only && at[n] >= 0) CHK(snap_store(c, A, at[n]));
        }
        double Jh = 0.0;
        CHK(d2h(c, &Jh, A.Jd, 8));
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (J) *J = Jh;
        if (o->forward_only) return 0;
        const int fin = A.nslots - 1;
        CHK(snap_store(c, A, fin));
        CHK(dalloc(c, &A.lamU, 3 * N * 8));
        CHK(dalloc(c, &A.lamP, N * 8));
        CHK(dalloc(c, &A.lamQ, 3 * N * 8));
        CHK(dalloc(c, &A.lamPhi, N * 8));
        CHK(dalloc(c, &A.lamB, N * 8));
        CHK(dalloc(c, &A.rhs, N * 8));
        CHK(dalloc(c, &A.r, N * 8));
        CHK(dalloc(c, &A.d, N * 8));
        CHK(dalloc(c, &A.Uend, 3 * N * 8));
        CHK(dalloc(c, &A.Utmp, 3 * N * c->es));
        if (c->mixed) {
            CHK(dalloc(c, &A.Ud, 3 * N * 8));
            CHK(dalloc(c, &A.rf, N * 4));
            CHK(dalloc(c, &A.df, N * 4));
        }
        CHK(dalloc(c, &A.grad, (size_t)p.nX * p.nY * 8));
        if (c->mixed) launch_convert<double, float>((const float *)c->Ut, A.Uend, 3 * (long)N, c->s);
        else HIPCHK(c, hipMemcpyAsync(A.Uend, c->Ut, 3 * N * 8, hipMemcpyDeviceToDevice, c->s));
        launch_adj_tau(c->g, A.lamU, A.i0, A.i1, A.cJ, c->s);
        CHK(revolve_top(c, A, 0, o->nsteps, 0, o->nsnap));
        CHK(snap_restore(c, A, fin));
        if (dJdw) CHK(d2h(c, dJdw, A.grad, (size_t)p.nX * p.nY * 8));
        HIPCHK(c, hipStreamSynchronize(c->s));
        return 0;
    };
    rc = run();
    if (nfwd) *nfwd = A.nfwd;
    if (nadj) *nadj = A.nadj;
    hipStreamSynchronize(c->s);
    adj_free(A);
    return rc;
}

// ------------------------------------------------------------------ operator hooks
template <class T> static int op_impl(tbl_ctx *c, int op, const double *in, double *out)
{
    const size_t N = (size_t)c->g.N;
    T *x = F<T>(c->r), *y = F<T>(c->V);   // V slot 0 doubles as the output
    CHK(put_field(c, x, !c->mixed, in, N));
    // the Krylov work buffers keep zero faces (the separable preconditioner's
    // wall-normal transform carries face rows through): drop the input's faces;
    // the Poisson apply reads them as its Dirichlet values, so apply it first
    if (op == 0) {
        HIPCHK(c, hipMemsetAsync(y, 0, N * c->es, c->s));
        launch_poisson<T>(c->g, x, y, nullptr, nullptr, c->s);
        launch_zero_faces<T>(c->g, x, c->s);
        HIPCHK(c, hipGetLastError());
        CHK(get_field(c, y, !c->mixed, out, N));
        return 0;
    }
    launch_zero_faces<T>(c->g, x, c->s);
    HIPCHK(c, hipMemsetAsync(y, 0, N * c->es, c->s));
    launch_precond<T>(c->g, c->p.precond, c->fast, x, F<T>(c->scr), y, nullptr, c->s);
    HIPCHK(c, hipGetLastError());
    CHK(get_field(c, y, !c->mixed, out, N));
    return 0;
}
extern "C" int tbl_apply_operator(tbl_ctx *c, int op, const double *in, double *out)
{
    if (!c || !in || !out || op < 0 || op > 1) return TBL_EARG;
    return c->mixed ? op_impl<float>(c, op, in, out) : op_impl<double>(c, op, in, out);
}

// ------------------------------------------------------------------ timers
extern "C" int tbl_timer_enable(tbl_ctx *c, unsigned mask)
{
    if (!c) return TBL_EARG;
    timer_flush(c);
    c->tmask = mask;
    return 0;
}
extern "C" int tbl_timer_read(tbl_ctx *c, int *count, double *ms, double *bytes)
{
    if (!c) return TBL_EARG;
    timer_flush(c);
    for (int q = 0; q < TBL_K_NCLASS; q++) {
        if (count) count[q] = c->tcount[q];
        if (ms) ms[q] = c->tms[q];
        if (bytes) bytes[q] = c->tbytes[q];
        c->tcount[q] = 0;
        c->tms[q] = 0.0;
        c->tbytes[q] = 0.0;
    }
    return 0;
}
