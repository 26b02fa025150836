// tbl_diag.hip -- diagnostics, statistics and the recycling rescale
// (reference: dnssetbl/dnssetbl.f90:851-905, 936-1262; wall shear from
// dnstbl/dnstbl.f90:1620-1668).  All of it computes in double.
#include "tbl_dev.h"
#include "tbl_launch.h"

namespace tbl {

// ---------------------------------------------------------------- a12
// Divergence sum (dnssetbl.f90:851-905): one thread per (i,j) column marching
// a chunk of planes.  The loads run through a register prefetch ring of
// DV_R planes (3 in FP32; 1 in FP64, where deeper rings measured slower;
// buffer loads: uniform plane offsets, 32-bit in-plane offsets): step k uses
// the streamwise (U) and span (V) stencil values of plane k and enters W's
// plane k+2 into the wall-normal queue, all loaded
// DV_R steps earlier, then refills the slot.  Loads retire in issue order,
// so a step waits only for loads issued DV_R steps before it.  Every column
// of the block marches the same planes; a thread outside the field works on
// the clamped column and adds nothing.  Per point the arithmetic is the
// reference's, in double, summed in plane order.
#ifndef DV_R64
#define DV_R64 1
#endif
template <class T>
__global__ __launch_bounds__(RT_T) void k_div_check(Geo g, const T *__restrict__ Ut, double *__restrict__ part, int kchunk)
{
    constexpr int DV_R = sizeof(T) == 8 ? DV_R64 : 3;
    __shared__ double sh[RT_T / 64];
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int ir = blockIdx.x * 64 + (t & 63), jr = blockIdx.y * RT_J + (t >> 6) + 1;
    const bool live = ir >= 1 && ir <= nX - 2 && jr <= nY - 2;
    const int i = min(max(ir, 1), nX - 2), j = min(jr, nY - 2);
    const int k0 = 1 + blockIdx.z * kchunk, k1 = min(k0 + kchunk, nZ - 1);
    // planes read: k0-3 .. k1-1+DV_R+2 (clamped)
    const int kb = max(k0 - 3, 0), ke = min(k1 + DV_R + 1, nZ - 1);
    constexpr unsigned B = sizeof(T);
    const unsigned PS = (unsigned)g.sZ * B;
    const unsigned range = (unsigned)(ke - kb + 1) * PS;
    const long N = g.N, base = (long)kb * g.sZ;
    const __amdgpu_buffer_rsrc_t RU = buf_of(Ut + base, range), RV = buf_of(Ut + N + base, range),
                                 RW = buf_of(Ut + 2 * N + base, range);
    auto so = [&](int kk) { return (unsigned)(max(0, min(kk, nZ - 1)) - kb) * PS; };
    const unsigned ox = (unsigned)(i - 2 + nX * j) * B;           // i-2 (>= nX - 1 >= 0)
    const unsigned om2 = (unsigned)(jm2_of(j, nY) * nX + i) * B, om1 = (unsigned)(jm1_of(j, nY) * nX + i) * B;
    const unsigned op1 = (unsigned)(jp1_of(j, nY) * nX + i) * B, op2 = (unsigned)(jp2_of(j, nY) * nX + i) * B;
    const double c12 = g.c12;
    const bool xedge = (i == 1) || (i == nX - 2), yedge = (j == 1) || (j == nY - 2);
    // ring slot: U at i-2, i-1, i+1, i+2 and V at j-2, j-1, j+1, j+2 of
    // plane k; W of plane k+2
    T ru[DV_R][4], rv[DV_R][4], rw[DV_R];
    auto fill = [&](T (&u)[4], T (&v)[4], T &w, int k) {
        const unsigned s = so(k);
        u[0] = bldT<T>(RU, ox, s); u[1] = bldT<T>(RU, ox + B, s); u[2] = bldT<T>(RU, ox + 3 * B, s);
        u[3] = bldT<T>(RU, ox + 4 * B, s);
        v[0] = bldT<T>(RV, om2, s); v[1] = bldT<T>(RV, om1, s); v[2] = bldT<T>(RV, op1, s); v[3] = bldT<T>(RV, op2, s);
        w = bldT<T>(RW, ox + 2 * B, so(k + 2));
    };
    // W's wall-normal stencil (planes k-2..k+2 after step k's shift)
    double wq[5];
#pragma unroll
    for (int q = 0; q < 5; q++) wq[q] = (double)bldT<T>(RW, ox + 2 * B, so(k0 - 3 + q));
#pragma unroll
    for (int r = 0; r < DV_R; r++) fill(ru[r], rv[r], rw[r], k0 + r);
    double a = 0.0;
    auto step = [&](auto S, int k) {
        constexpr int r = decltype(S)::value;
        wq[0] = wq[1]; wq[1] = wq[2]; wq[2] = wq[3]; wq[3] = wq[4];
        wq[4] = (double)rw[r];
        double um2 = ru[r][0], um1 = ru[r][1], up1 = ru[r][2], up2 = ru[r][3];
        double vm2 = rv[r][0], vm1 = rv[r][1], vp1 = rv[r][2], vp2 = rv[r][3];
        // the slot's values are taken before it is refilled: the empty asm
        // keeps the conversions here (sunk to their uses, they would keep
        // the slot's registers live across the refill, and the loop would
        // copy values still in flight)
        asm("" : "+v"(um2), "+v"(um1), "+v"(up1), "+v"(up2), "+v"(vm2), "+v"(vm1), "+v"(vp1), "+v"(vp2),
            "+v"(wq[4]));
        __builtin_amdgcn_sched_barrier(0);
        // past the chunk's last plane the refill repeats that plane (the
        // values are never used, and the repeat is a cache hit, not HBM)
        fill(ru[r], rv[r], rw[r], min(k + DV_R, k1 - 1));
        __builtin_amdgcn_sched_barrier(0);
        const double e1 = (up1 - um1) * (0.5 * g.invdx);
        const double c1 = ((um2 - 8 * um1 + 8 * up1 - up2) * (c12 * g.invdx));
        const double e2 = (vp1 - vm1) * (0.5 * g.invdy);
        const double c2 = ((vm2 - 8 * vm1 + 8 * vp1 - vp2) * (c12 * g.invdy));
        const double rt1 = xedge ? e1 : c1, rt2 = yedge ? e2 : c2;
        const double e3 = ((wq[3] - wq[1]) * (0.5 * g.invdXi));
        const double c3 = ((wq[0] - 8 * wq[1] + 8 * wq[3] - wq[4]) * (c12 * g.invdXi));
        const double rt3 = (k <= 1 || k >= nZ - 2) ? e3 : c3;
        const double d = (rt1 + rt2 + rt3 * g.JacInv[k]);
        a += live ? d : 0.0;
        __builtin_amdgcn_sched_barrier(0);
    };
    int k = k0;
    for (; k + DV_R <= k1; k += DV_R) ring_steps<0, DV_R>(step, k);
    ring_tail<0, DV_R>(step, k, k1);
    double s = block_sum<RT_T>(a, sh, t);
    if (t == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
}

__global__ __launch_bounds__(1024) void k_fin_sum(const double *part, int nb, double *out)
{
    __shared__ double sh[16];
    double a = 0.0;
#pragma unroll 8
    for (int q = threadIdx.x; q < nb; q += 1024) a += part[q];
    double s = block_sum<1024>(a, sh, threadIdx.x);
    if (threadIdx.x == 0) *out = s;
}

// sum of nb partials; out[0] = |sum| / nint (the printed metric), out[1] = sum
__global__ __launch_bounds__(1024) void k_fin_div(const double *part, int nb, double nint, double *out)
{
    __shared__ double sh[16];
    double a = 0.0;
#pragma unroll 8
    for (int q = threadIdx.x; q < nb; q += 1024) a += part[q];
    double s = block_sum<1024>(a, sh, threadIdx.x);
    if (threadIdx.x == 0) {
        out[0] = fabs(s) / nint;
        out[1] = s;
    }
}

// ---------------------------------------------------------------- a13
struct Planes { int p[5]; double denom; };
template <class T>
__global__ __launch_bounds__(256) void k_mass_flux(Geo g, const T *__restrict__ Ut, Planes pl, double *__restrict__ part)
{
    __shared__ double sh[4];
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int q = blockIdx.y;
    const int i = pl.p[q] - 1;
    const long sk = g.sZ;
    const int nj = nY - 2;
    const long tot = (long)nj * (nZ - 2);
    double a = 0.0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += 256L * MF_BLK) {
        const int j = (int)(e % nj) + 1, k = (int)(e / nj) + 1;
        const double dAv = g.dy * (g.x3[k + 1] - g.x3[k]);
        const long c = (long)i + (long)nX * ((long)j + (long)nY * k);
        const long cp = c + (long)(jp1_of(j, nY) - j) * nX;
        a += (0.25 * ((double)Ut[c] + (double)Ut[cp] + (double)Ut[c + sk] + (double)Ut[cp + sk]) * (dAv));
    }
    double s = block_sum<256>(a, sh, threadIdx.x);
    if (threadIdx.x == 0) part[q * MF_BLK + blockIdx.x] = s;
}
__global__ void k_mass_fin(const double *__restrict__ part, double denom, double *__restrict__ out)
{
    const int q = threadIdx.x;
    if (q >= 5) return;
    double a = 0.0;
    for (int b = 0; b < MF_BLK; b++) a += part[q * MF_BLK + b];
    out[q] = a / denom;
}

// ---------------------------------------------------------------- a14
// thread per (i,k); the span sum runs in the reference's j order
template <class T>
__global__ __launch_bounds__(256) void k_stats(Geo g, const T *__restrict__ Ut, double *__restrict__ sums)
{
    // One thread per (i, k) walks the span rows in order (the sums keep the
    // reference's j order).  Row j's values are carried from the previous
    // row, so each row is loaded once; the loads run SR rows ahead through a
    // register ring of SR + 1 slots (buffer loads with 32-bit in-plane
    // offsets), unrolled by the slot count so that a slot keeps its
    // registers and no value in flight is copied.
    constexpr int SR = 6, NS = SR + 1;
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    if (k >= nZ) return;
    const bool live = i < nX;
    const int ic = min(i, nX - 1);
    const long N = g.N, M = (long)nX * nZ;
    constexpr unsigned B = sizeof(T);
    const long pk = (long)k * g.sZ;
    const unsigned PB = (unsigned)g.sZ * B;
    const __amdgpu_buffer_rsrc_t RU = buf_of(Ut + pk, PB), RV = buf_of(Ut + N + pk, PB),
                                 RW = buf_of(Ut + 2 * N + pk, PB);
    // q-th row of the sequence 1, 2, ..., nY-2, then the periodic wrap to 1
    // (jp1 of the last row); indices past the end repeat the wrap row
    auto off = [&](int q) { return (unsigned)(ic + nX * (q <= nY - 2 ? q : 1)) * B; };
    T su_[NS], sv_[NS], sw_[NS];
    auto fill = [&](int r, int q) {
        const unsigned o = off(q);
        su_[r] = bldT<T>(RU, o, 0); sv_[r] = bldT<T>(RV, o, 0); sw_[r] = bldT<T>(RW, o, 0);
    };
    double cu, cv, cw;
    {
        const unsigned o = off(1);
        cu = (double)bldT<T>(RU, o, 0); cv = (double)bldT<T>(RV, o, 0); cw = (double)bldT<T>(RW, o, 0);
    }
#pragma unroll
    for (int r = 0; r < SR; r++) fill(r, 2 + r);
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0, t7 = 0, t8 = 0;
    auto step = [&](auto S, int j) {
        constexpr int r = decltype(S)::value;
        // row j+1+SR into the slot the previous row read
        fill((r + SR) % NS, j + 1 + SR);
        __builtin_amdgcn_sched_barrier(0);
        const double nu = (double)su_[r], nv = (double)sv_[r], nw = (double)sw_[r];
        const double su = 0.50 * (cu + nu);
        const double sv = 0.50 * (cv + nv);
        const double sw = 0.50 * (cw + nw);
        t0 = t0 + su; t1 = t1 + sv; t2 = t2 + sw;
        t3 = t3 + su * su; t4 = t4 + sv * sv; t5 = t5 + sw * sw;
        t6 = t6 + su * sv; t7 = t7 + su * sw; t8 = t8 + sv * sw;
        cu = nu; cv = nv; cw = nw;
        __builtin_amdgcn_sched_barrier(0);
    };
    // rows j = 1 .. nY-2 (the ring slot of row j's successor is (j - 1) mod NS)
    int j = 1;
    for (; j + NS <= nY - 1; j += NS) ring_steps<0, NS>(step, j);
    ring_tail<0, NS>(step, j, nY - 1);
    if (!live) return;
    const double A = nY - 2;
    const long e = (long)i + (long)nX * k;
    sums[e] = sums[e] + (t0 / A);
    sums[e + M] = sums[e + M] + (t1 / A);
    sums[e + 2 * M] = sums[e + 2 * M] + (t2 / A);
    sums[e + 3 * M] = sums[e + 3 * M] + (t3 / A);
    sums[e + 4 * M] = sums[e + 4 * M] + (t4 / A);
    sums[e + 5 * M] = sums[e + 5 * M] + (t5 / A);
    sums[e + 6 * M] = sums[e + 6 * M] + (t6 / A);
    sums[e + 7 * M] = sums[e + 7 * M] + (t7 / A);
    sums[e + 8 * M] = sums[e + 8 * M] + (t8 / A);
}

// ---------------------------------------------------------------- a15
// averages and reset of the running sums (dnssetbl.f90:1018-1032, 1209-1219)
__global__ void k_rescale_avg(long M, int nsamples, double *sums, double *avg)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < 9 * M; e += (long)gridDim.x * 256) {
        avg[e] = (sums[e] / nsamples);
        sums[e] = 0.0;
    }
}
// fluctuations at the recycling plane (dnssetbl.f90:1034-1042)
template <class T>
__global__ void k_rescale_uf(Geo g, int i_rec0, const T *__restrict__ Ut, const double *__restrict__ avg,
                             double *__restrict__ uf)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (j > nY - 2 || k > nZ - 2) return;
    const long c = (long)i_rec0 + (long)nX * ((long)j + (long)nY * k), N = g.N, M = (long)nX * nZ;
    const long o = (long)j + (long)nY * k, P = (long)nY * nZ, a = (long)i_rec0 + (long)nX * k;
    uf[o] = (double)Ut[c] - avg[a];
    uf[o + P] = (double)Ut[c + N] - avg[a + M];
    uf[o + 2 * P] = (double)Ut[c + 2 * N] - avg[a + 2 * M];
}
// friction velocities, delta99, Lund weight, new(k), integral thicknesses
// (dnssetbl.f90:1055-1145); one block.
__global__ __launch_bounds__(256) void k_rescale_main(Geo g, int i_rec0, const double *__restrict__ avg,
                                                     RecDev *rec, double *Wt, int *newk, double *bltkns,
                                                     double *deltastar, double *theta, double *scratch)
{
    const int nX = g.nX, nZ = g.nZ;
    const double *usum2 = avg, *x3 = g.x3;
    const double Ue = g.Ue;
    const int t = threadIdx.x;
    // delta99: first node with U >= 0.99 Ue, else unchanged
    for (int i = t; i < nX; i += 256) {
        for (int k = 0; k < nZ; k++)
            if (usum2[i + (long)nX * k] >= (0.99 * Ue)) { bltkns[i] = x3[k]; break; }
    }
    __syncthreads();
    __shared__ double lam, lam03;
    if (t == 0) {
        const double utau1 = sqrt(((usum2[0 + (long)nX * 1] - usum2[0]) / (x3[1] - x3[0])) / g.Re);
        const double utau2 = sqrt(((usum2[i_rec0 + (long)nX * 1] - usum2[i_rec0]) / (x3[1] - x3[0])) / g.Re);
        lam = (utau1) / (utau2);
        lam03 = bltkns[i_rec0] / bltkns[0];
        rec->lamda = lam;
        rec->lamda03 = lam03;
    }
    __syncthreads();
    for (long e = t; e < (long)nX * nZ; e += 256) {
        const int i = (int)(e % nX), k = (int)(e / nX);
        const double eta = x3[k] / bltkns[i];
        Wt[e] = 0.5 * (1 + (tanh((4.0 * (eta - 0.2)) / ((1 - 2.0 * 0.2) * eta + 0.2)) / tanh(4.0)));
    }
    __syncthreads();
    if (t == 0) {   // recycling index map, the reference's two-phase scan (1-based values)
        double *eta = scratch, *zrecy = scratch + (nZ + 1);
        for (int k = 1; k <= nZ; k++) newk[k - 1] = nZ - 1;
        int ksep = 0;
        for (int k = 1; k <= nZ - 1; k++) {
            if (Wt[i_rec0 + (long)nX * (k - 1)] <= 0.99) { ksep = k; eta[k] = lam; }
            else eta[k] = lam03;
        }
        for (int k = 1; k <= nZ - 1; k++) zrecy[k] = eta[k] * x3[k - 1];
        int counter = 1;
        for (int k = 1; k <= nZ - 1; k++)
            if (counter <= ksep)
                if (x3[k - 1] >= zrecy[counter]) { newk[counter - 1] = k; counter = counter + 1; }
        counter = ksep + 1;
        for (int k = ksep + 1; k <= nZ - 1; k++)
            if (counter <= (nZ - 1))
                if (x3[k - 1] >= zrecy[counter]) { newk[counter - 1] = k; counter = counter + 1; }
        rec->ksep = ksep;
    }
    for (int i = t; i < nX; i += 256) {
        double ds = 0.0;
        for (int k = 0; k < nZ - 1; k++) {
            ds = ds + ((1.0 - (usum2[i + (long)nX * k] / Ue)) * (x3[k + 1] - x3[k]));
            if (x3[k] == bltkns[i]) break;
        }
        deltastar[i] = ds;
        double th = 0.0;
        for (int k = 0; k < nZ - 1; k++) {
            const double r = usum2[i + (long)nX * k] / Ue;
            th = th + ((r) * (1.0 - r) * (x3[k + 1] - x3[k]));
            if (x3[k] >= bltkns[i]) break;
        }
        theta[i] = th;
    }
}

// ---------------------------------------------------------------- a16
template <class T>
__global__ __launch_bounds__(RT_T) void k_ke(Geo g, int i_rec0, const T *__restrict__ Ut,
                                           const double *__restrict__ avg, double *__restrict__ part,
                                           int *guard, double *uf, int kchunk)
{
    __shared__ double sh[RT_T / 64];
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63);
    const int j = blockIdx.y * RT_J + (t >> 6) + 1;
    const int k0 = 1 + blockIdx.z * kchunk, k1 = min(k0 + kchunk, nZ - 1);
    double a = 0.0;
    int gd = 0;
    if (i >= 1 && i <= nX - 2 && j <= nY - 2) {
        const long N = g.N, M = (long)nX * nZ;
        for (int k = k0; k < k1; k++) {
            const long c = (long)i + (long)nX * ((long)j + (long)nY * k);
            const long e = (long)i + (long)nX * k;
            const double f1 = (double)Ut[c] - avg[e];
            const double f2 = (double)Ut[c + N] - avg[e + M];
            const double f3 = (double)Ut[c + 2 * N] - avg[e + 2 * M];
            const double q = f1 * f1 + f2 * f2 + f3 * f3;
            a += q;
            double mx = f1 > f2 ? f1 : f2;
            mx = mx > f3 ? mx : f3;
            if (mx > 25.0) gd |= 1;
            if (!(q == q)) gd |= 2;
            if (i == i_rec0) {
                const long o = (long)j + (long)nY * k, P = (long)nY * nZ;
                uf[o] = f1; uf[o + P] = f2; uf[o + 2 * P] = f3;
            }
        }
    }
    if (gd) atomicOr(guard, gd);
    double s = block_sum<RT_T>(a, sh, t);
    if (t == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
}

// ---------------------------------------------------------------- a17
// acc (optional): the running window sum, acc(i) += tau(i)
template <class T>
__global__ __launch_bounds__(1024) void k_wall_shear(Geo g, const T *__restrict__ Ut, double *tau, double *acc)
{
    // 64 streamwise stations per block (one per lane) x 16 span-row groups:
    // the 16 wavefronts form the terms of 16 rows at a time into an LDS tile,
    // and the first wavefront adds them in row order (the reference's
    // summation order, dnstbl/dnstbl.f90:1620-1668)
    __shared__ double tt[16][64];
    const int nX = g.nX, nY = g.nY;
    const int lane = threadIdx.x & 63, jg = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const bool station = i >= 1 && i <= nX - 2;
    const int ic = min(max(i, 1), nX - 2);
    const long N = g.N, sk = g.sZ;
    const T *U = Ut, *W = Ut + 2 * N;
    const double dx = g.dx, df1 = g.df1, df2 = g.df2;
    const bool xedge = (ic <= 2 || ic >= nX - 3);
    double t = 0.0;
    for (int j0 = 1; j0 <= nY - 2; j0 += 16) {
        const int j = j0 + jg;
        if (j <= nY - 2) {
            const long c = (long)ic + (long)nX * j;
            double mfb1;
            if (xedge)
                mfb1 = ((double)W[c + 1] - (double)W[c - 1]) / (2.0 * dx);
            else
                mfb1 = (-(double)W[c - 3] + (9.0 * (double)W[c - 2]) - (45.0 * (double)W[c - 1]) +
                        (45.0 * (double)W[c + 1]) - (9.0 * (double)W[c + 2]) + (double)W[c + 3]) / (60.0 * dx);
            const double fder = ((((double)U[c + 2 * sk] - (double)U[c]) * (df1 * df1)) +
                                 (((double)U[c] - (double)U[c + sk]) * (df2 * df2))) /
                                ((df2 * (df1 * df1)) - ((df2 * df2) * df1));
            tt[jg][lane] = (mfb1 + fder) / g.Re;
        }
        __syncthreads();
        if (jg == 0)
            for (int q = 0; q < 16 && j0 + q <= nY - 2; q++) t = t + tt[q][lane];
        __syncthreads();
    }
    if (jg != 0 || i >= nX) return;
    if (!station) { tau[i] = 0.0; return; }
    tau[i] = t;
    if (acc) acc[i] = t + acc[i];
}

// end of a wall-shear window (dnstbl/dnstbl.f90:1670-1682): the window mean
// sum/window, written as the span mean and the skin-friction coefficient;
// the sum restarts from zero
__global__ void k_tau_window(int nX, int nY, double Ue, int window, double *acc, double *mean, double *cf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nX) return;
    if (i == 0 || i == nX - 1) { mean[i] = 0.0; cf[i] = 0.0; acc[i] = 0.0; return; }
    const double s = acc[i] / (double)window;
    mean[i] = s / (double)(nY - 2);
    cf[i] = 2.0 * (s / (Ue * Ue * (double)(nY - 2)));
    acc[i] = 0.0;
}

// ---------------------------------------------------------------- launchers
static inline dim3 rgrid(const Geo &g, int *kchunk)
{
    RTiles r = rtiles_of(g.nX, g.nY, g.nZ);
    *kchunk = r.kchunk;
    return dim3(r.gx, r.gy, r.kb);
}

template <class T>
void launch_div_check(const Geo &g, const T *Ut, double *part, hipStream_t s)
{
    int kc;
    dim3 grid = rgrid(g, &kc);
    hipLaunchKernelGGL(k_div_check<T>, grid, dim3(RT_T), 0, s, g, Ut, part, kc);
}
void launch_fin_sum(const double *part, int nb, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(1024), 0, s, part, nb, out);
}
void launch_fin_div(const double *part, int nb, double nint, double *out2, hipStream_t s)
{
    hipLaunchKernelGGL(k_fin_div, dim3(1), dim3(1024), 0, s, part, nb, nint, out2);
}
template <class T>
void launch_mass_flux(const Geo &g, const T *Ut, const int *planes, double denom, double *out5, double *part,
                      hipStream_t s)
{
    // part holds 5 * MF_BLK doubles
    Planes pl;
    for (int q = 0; q < 5; q++) pl.p[q] = planes[q];
    pl.denom = denom;
    hipLaunchKernelGGL(k_mass_flux<T>, dim3(MF_BLK, 5), dim3(256), 0, s, g, Ut, pl, part);
    hipLaunchKernelGGL(k_mass_fin, dim3(1), dim3(64), 0, s, (const double *)part, denom, out5);
}
template <class T>
void launch_stats(const Geo &g, const T *Ut, double *sums, hipStream_t s)
{
    hipLaunchKernelGGL(k_stats<T>, dim3((g.nX + 255) / 256, g.nZ), dim3(256), 0, s, g, Ut, sums);
}
template <class T>
void launch_rescale(const Geo &g, int i_rec0, int nsamples, const T *Ut, double *sums, double *avg, RecDev *rec,
                    double *Wt, int *newk, double *bltkns, double *deltastar, double *theta, double *uf,
                    double *scan_scratch, hipStream_t s)
{
    const long M = (long)g.nX * g.nZ;
    long nb = (9 * M + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_rescale_avg, dim3((unsigned)nb), dim3(256), 0, s, M, nsamples, sums, avg);
    if (i_rec0 <= g.nX - 2)
        hipLaunchKernelGGL(k_rescale_uf<T>, dim3((g.nY - 2 + 255) / 256, g.nZ - 2), dim3(256), 0, s, g, i_rec0, Ut,
                           (const double *)avg, uf);
    hipLaunchKernelGGL(k_rescale_main, dim3(1), dim3(256), 0, s, g, i_rec0, (const double *)avg, rec, Wt, newk,
                       bltkns, deltastar, theta, scan_scratch);
}
template <class T>
void launch_ke(const Geo &g, int i_rec0, const T *Ut, const double *avg, double *part, int *guard, double *uf,
               hipStream_t s)
{
    int kc;
    dim3 grid = rgrid(g, &kc);
    hipLaunchKernelGGL(k_ke<T>, grid, dim3(RT_T), 0, s, g, i_rec0, Ut, avg, part, guard, uf, kc);
}
template <class T>
void launch_wall_shear(const Geo &g, const T *Ut, double *tau, hipStream_t s, double *acc)
{
    hipLaunchKernelGGL(k_wall_shear<T>, dim3((g.nX + 63) / 64), dim3(1024), 0, s, g, Ut, tau, acc);
}
void launch_tau_window(const Geo &g, int window, double *acc, double *mean, double *cf, hipStream_t s)
{
    hipLaunchKernelGGL(k_tau_window, dim3((g.nX + 255) / 256), dim3(256), 0, s, g.nX, g.nY, g.Ue, window, acc, mean,
                       cf);
}

#define TBL_INST(T)                                                                                           \
    template void launch_div_check<T>(const Geo &, const T *, double *, hipStream_t);                         \
    template void launch_mass_flux<T>(const Geo &, const T *, const int *, double, double *, double *, hipStream_t); \
    template void launch_stats<T>(const Geo &, const T *, double *, hipStream_t);                             \
    template void launch_rescale<T>(const Geo &, int, int, const T *, double *, double *, RecDev *, double *,  \
                                    int *, double *, double *, double *, double *, double *, hipStream_t);    \
    template void launch_ke<T>(const Geo &, int, const T *, const double *, double *, int *, double *, hipStream_t); \
    template void launch_wall_shear<T>(const Geo &, const T *, double *, hipStream_t, double *);
TBL_INST(double)
TBL_INST(float)
#undef TBL_INST

}  // namespace tbl
