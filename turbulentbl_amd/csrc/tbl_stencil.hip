// tbl_stencil.hip -- stencil kernels of the TBL step for gfx950.
//
// Every kernel keeps the arithmetic order of the reference expression it
// replaces (cited per kernel, dnssetbl/dnssetbl.f90), and the library is built
// with -ffp-contract=off, so in FP64 the stencil results are bitwise those of
// the reference's formulas.  Templates on the storage type T give the FP32
// (mixed-precision) variants; quantities the reference accumulates (residual
// norms) are always accumulated in double.
#include "tbl_dev.h"
#include "tbl_launch.h"

namespace tbl {

template <class T> __device__ __forceinline__ T tabs(T x) { return x < T(0) ? -x : x; }

// ----------------------------------------------------------------------------
// a3+a4: convection (Peclet-switched Kawamura-Kuwahara / central) + old
// pressure gradient + momentum RHS Q (dnssetbl.f90:306-369).
// One thread per (i,j) column marching a chunk of k planes; 1-D grid of 64 x RHS_TJ
// tiles remapped XCD-contiguously.  The wall-normal stencil values of the
// three components (planes k-2..k+2) and of the pressure (k-1..k+1) live in a
// register queue, so each plane is read once per column; every stencil value
// is loaded before the scheme is selected, which leaves the Peclet branches
// as arithmetic.  Each branch evaluates the reference's expression in its
// own operand order (bitwise in FP64).  Offsets are 32-bit inside one field
// (3N < 2^31 is checked at create time).
// SCHEME = TBL_SCHEME_DNSTBL evaluates the dnstbl generation of the same
// block (dnstbl/dnstbl.f90:620-766) instead: cell Peclet number signed by
// sign(1, ceiling(u)) and Kawamura-Kuwahara only inside 2 < Pe < 50
// (2 < Pe < 5 wall-normal), u (F(i+1) - F(i-1)) / (2 dx) at the streamwise
// edge nodes, first-order upwind at the span-edge rows, and
// Q = (U + dt (forc - conv)) - dt gradp.
// ----------------------------------------------------------------------------
#ifndef RHS_KCHUNK
#define RHS_KCHUNK 32   // planes per workgroup (register-queue kernel)
#endif
#define RHS_KCHUNK_F32 192   // most planes per workgroup of the FP32 tile kernel
#define RHS_TJ 8   // span rows per workgroup (64 x RHS_TJ columns): fewer span-halo re-reads
// waves per SIMD the register allocation must allow: 6 for the FP32 default
// scheme (80 VGPRs, no spill); the compiler's choice (5) otherwise, where 6
// would spill
#define RHS_WAVES(T, SCHEME) ((sizeof(T) == 4 && (SCHEME) == 0) ? 6 : 1)
// BR: the Peclet switches are recorded (brout) or replayed (brin) by an
// adjoint window; the stepping path instantiates BR = false, with no branch
// around a load in its plane loop (such a branch makes the compiler's
// in-order load waits drain every load in flight)
template <class T, int SCHEME, bool BR>
__global__ __launch_bounds__(64 * RHS_TJ) __attribute__((amdgpu_waves_per_eu(RHS_WAVES(T, SCHEME))))
void k_momentum_rhs(Geo g, const T *__restrict__ Ut,
                                                     const T *__restrict__ P, T *__restrict__ Q,
                                                     int kchunk, unsigned nbx, unsigned nby, unsigned nbz,
                                                     const unsigned char *__restrict__ brin,
                                                     unsigned char *__restrict__ brout, double *__restrict__ hpart)
{
    const unsigned L = xcd_remap(blockIdx.x, nbx * nby * nbz);
    const unsigned bx = L % nbx, by = (L / nbx) % nby, bz = L / (nbx * nby);
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int ir = bx * 64 + (threadIdx.x & 63), jr = by * RHS_TJ + (threadIdx.x >> 6) + 1;
    const bool live = ir >= 1 && ir <= nX - 2 && jr <= nY - 2;
    if (!live && !hpart) return;               // with hpart every thread reaches the block sums
    const int i = min(max(ir, 1), nX - 2), j = min(jr, nY - 2);
    const int k0 = 1 + bz * kchunk;
    const int k1 = live ? min(k0 + kchunk, nZ - 1) : k0;
    const long N = g.N;
    const int sk = (int)g.sZ;
    const T Re = T(g.Re), dx = T(g.dx), dy = T(g.dy), invdx = T(g.invdx), invdy = T(g.invdy),
            invdXi = T(g.invdXi), c12 = T(g.c12), dt = T(g.dt);
    const int dm1 = (jm1_of(j, nY) - j) * nX, dm2 = (jm2_of(j, nY) - j) * nX;
    const int dp1 = (jp1_of(j, nY) - j) * nX, dp2 = (jp2_of(j, nY) - j) * nX;
    const bool xedge = (i <= 1 || i >= nX - 2);
    const bool yedge = (j == 1 || j == nY - 2);
    const int c0 = i + nX * j;                                    // k = 0 plane
    const T *Fp[3] = {Ut, Ut + N, Ut + 2 * N};
    auto plane = [&](int kk) { return c0 + sk * max(0, min(kk, nZ - 1)); };
    // hpart: the first residual pass of the Helmholtz solve (dnssetbl.f90:378-386)
    // from the same values, as k_helm_resid forms it: per-block sums of |r|
    // and |Q| per component (a thread's k chunk summed in the storage type,
    // the block sum in double)
    T ha[3] = {T(0), T(0), T(0)}, hq[3] = {T(0), T(0), T(0)};
    T zq[3][5], pq[3];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const int o = plane(k0 - 2 + q);
#pragma unroll
        for (int p = 0; p < 3; p++) zq[p][q] = Fp[p][o];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) pq[q] = P[plane(k0 - 1 + q)];
    for (int k = k0; k < k1; k++) {
        const int c = c0 + sk * k;
        const int cm1 = c + dm1, cm2 = c + dm2, cp1 = c + dp1, cp2 = c + dp2;
        const T u = zq[0][2], v = zq[1][2], w = zq[2][2];
        const bool zedge = (k <= 1 || k >= nZ - 2);
        const T JacInv = T(g.JacInv[k]);
        const T dz3 = T(g.x3[k + 1] - g.x3[k]);
        // Peclet switches; an adjoint window may record them (brout) or replay
        // a recorded pattern (brin): bit 0 x, bit 1 y, bit 2 z
        bool kx, ky, kz;
        if (SCHEME == TBL_SCHEME_SETBL) {
            const T Pex = tabs(Re * u * dx), Pey = tabs(Re * v * dy), Pez = tabs(Re * w * dz3);
            kx = Pex > T(2.0); ky = Pey > T(2.0); kz = Pez > T(2.0);
        } else {
            // sgn = sign(1, ceiling(u)): -1 only for u <= -1
            auto sg = [](T a) { return ceil(a) < T(0) ? T(-1) : T(1); };
            const T Pex = sg(u) * Re * u * dx, Pey = sg(v) * Re * v * dy, Pez = (sg(w) * Re * w * dz3);
            kx = Pex > T(2.0) && Pex < T(50.0);
            ky = Pey > T(2.0) && Pey < T(50.0);
            kz = Pez > T(2.0) && Pez < T(5.0);
        }
        if (BR && brin) {
            const unsigned b = brin[c];
            kx = (b & 1u) != 0; ky = (b & 2u) != 0; kz = (b & 4u) != 0;
        }
        if (BR && brout) brout[c] = (unsigned char)((kx ? 1u : 0u) | (ky ? 2u : 0u) | (kz ? 4u : 0u));
        const T px1 = P[c + 1], px0 = P[c - 1], py1 = P[cp1], py0 = P[cm1];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const T *F = Fp[p];
            const T f0 = zq[p][2];
            // every x and y stencil value is in bounds for interior (i,j):
            // i +- 2 stays inside the field, j uses the periodic tables
            const T xm2 = F[c - 2], xm1 = F[c - 1], xp1 = F[c + 1], xp2 = F[c + 2];
            const T ym2 = F[cm2], ym1 = F[cm1], yp1 = F[cp1], yp2 = F[cp2];
            const T zm2 = zq[p][0], zm1 = zq[p][1], zp1 = zq[p][3], zp2 = zq[p][4];
            T rt1, rt2, rt3;
            if (xedge) {
                if (SCHEME == TBL_SCHEME_SETBL) rt1 = u * (xp1 - xm1) * T(0.5) * invdx;
                else rt1 = u * (xp1 - xm1) / (T(2.0) * dx);
            } else if (kx) {
                T rt11 = tabs(u) * (xp2 - (T(4) * (xp1)) + (T(6) * (f0)) - (T(4) * (xm1)) + xm2) * (T(0.25) * invdx);
                T rt12 = (u) * (-xp2 + (T(8) * (xp1)) - (T(8) * (xm1)) + xm2) * (c12 * invdx);
                rt1 = rt11 + rt12;
            } else {
                rt1 = u * ((xm2 - T(8) * xm1 + T(8) * xp1 - xp2) * (c12 * invdx));
            }
            if (yedge) {
                if (SCHEME == TBL_SCHEME_SETBL) {
                    rt2 = v * (yp1 - ym1) * T(0.5) * invdy;
                } else {
                    const T a1 = v > T(0) ? v : T(0), b1 = v < T(0) ? v : T(0);
                    rt2 = a1 * ((f0 - ym1) * (invdy)) + b1 * ((yp1 - f0) * (invdy));
                }
            } else if (ky) {
                T rt11 = (tabs(v)) * (yp2 - (T(4) * (yp1)) + (T(6) * (f0)) - (T(4) * (ym1)) + ym2) * (T(0.25) * invdy);
                T rt12 = (v) * (-yp2 + (T(8) * (yp1)) - (T(8) * (ym1)) + ym2) * (c12 * invdy);
                rt2 = rt11 + rt12;
            } else {
                rt2 = v * ((ym2 - T(8) * ym1 + T(8) * yp1 - yp2) * (c12 * invdy));
            }
            if (zedge) {
                rt3 = (w * (zp1 - zm1) * (T(0.5) * invdXi));
            } else if (kz) {
                T rt11 = (tabs(w)) * (zp2 - (T(4) * (zp1)) + (T(6) * (f0)) - (T(4) * (zm1)) + zm2) * (T(0.25) * invdXi);
                T rt12 = (w) * (-zp2 + (T(8) * (zp1)) - (T(8) * (zm1)) + zm2) * (c12 * invdXi);
                rt3 = (rt11 + rt12);
            } else {
                rt3 = w * ((zm2 - T(8) * zm1 + T(8) * zp1 - zp2) * (c12 * invdXi));
            }
            const T cv = rt1 + rt2 + (rt3 * JacInv);
            T gp;
            if (p == 0) gp = T(0.5) * (px1 - px0) * (invdx);
            else if (p == 1) gp = T(0.5) * (py1 - py0) * (invdy);
            else gp = T(0.5) * (pq[2] - pq[0]) * (invdXi * JacInv);
            T qv;
            if (SCHEME == TBL_SCHEME_SETBL) qv = f0 - dt * (cv + gp);
            else qv = (f0 + (dt * (T(g.forc[p]) - cv))) - (dt * gp);
            Q[c + (long)p * N] = qv;
            if (hpart) {
                const T rt1h = T(g.Ae) * xp1 + T(g.Aw) * xm1 + T(g.An) * yp1 + T(g.As) * ym1;
                const T rt2h = T(g.At[k]) * zp1 + T(g.Ab[k]) * zm1;
                const T resLoc = qv - (T(1.0) - dt / Re * T(g.Ap[k])) * f0 + dt * (rt1h + rt2h) / Re;
                ha[p] += tabs(resLoc);
                hq[p] += tabs(qv);
            }
        }
        // advance the queues by one plane
        const int on = plane(k + 3), op = plane(k + 2);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            zq[p][0] = zq[p][1]; zq[p][1] = zq[p][2]; zq[p][2] = zq[p][3]; zq[p][3] = zq[p][4];
            zq[p][4] = Fp[p][on];
        }
        pq[0] = pq[1]; pq[1] = pq[2]; pq[2] = P[op];
    }
    if (hpart) {
        __shared__ double hs[RHS_TJ];
        const unsigned nb = gridDim.x;
        for (int p = 0; p < 3; p++) {
            const double s1 = block_sum<64 * RHS_TJ>((double)ha[p], hs, threadIdx.x);
            if (threadIdx.x == 0) hpart[(long)p * nb + blockIdx.x] = s1;
            const double s2 = block_sum<64 * RHS_TJ>((double)hq[p], hs, threadIdx.x);
            if (threadIdx.x == 0) hpart[(long)(3 + p) * nb + blockIdx.x] = s2;
        }
    }
}

// ----------------------------------------------------------------------------
// a3+a4 in FP32 with the default scheme (dnssetbl.f90:306-369): the stencils,
// Peclet switches and fused first Helmholtz residual of k_momentum_rhs,
// evaluated with fused multiply-adds, in a software-pipelined plane loop.
//  * Every load is a buffer load: one resource per field and block, based at
//    the block's first plane, a uniform plane offset (scalar) and the
//    thread's 32-bit in-plane offsets, so the loop has no vector address
//    arithmetic.
//  * Plane k+1's x and y neighbours are loaded while plane k is formed, into
//    the other of two register sets (the sets swap roles every plane); the
//    wall-normal queue plane k+4 (pressure k+3) is loaded after them and
//    enters the queue one plane later.  A scheduling barrier keeps these
//    loads ahead of plane k's arithmetic.
// The hardware retires loads in issue order, and a register copy of a
// loaded value waits for its load: with this order, waiting for plane k's
// neighbours never waits for a wall-normal (HBM) load of the same or the
// previous plane, and the queue shift copies values loaded a plane earlier.
// ----------------------------------------------------------------------------
// barrier on LDS data only: the thread's LDS writes are complete, global loads
// may stay in flight (the "memory" clobber keeps the compiler from moving
// memory accesses across it)
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// one convection derivative with fused multiply-adds: uc = u c12/h,
// ua = |u|/(4h) where the Kawamura-Kuwahara form applies (else 0),
// uh = u/(2h) for the edge form
//   central  uc (m2 - 8 m1 + 8 p1 - p2)
//   KK       central + ua (p2 - 4 p1 + 6 f0 - 4 m1 + m2)
//   edge     uh (p1 - m1)
__device__ __forceinline__ float conv_fma(float m2, float m1, float f0, float p1, float p2, float uc, float ua,
                                          float uh, bool edge, float &s1)
{
    const float e = p1 - m1;
    s1 = p1 + m1;
    const float d3 = __builtin_fmaf(8.0f, e, m2 - p2);
    const float d4 = __builtin_fmaf(-4.0f, s1, __builtin_fmaf(6.0f, f0, p2 + m2));
    const float r = __builtin_fmaf(ua, d4, uc * d3);
    return edge ? uh * e : r;
}

#ifndef ZB_REV
#define ZB_REV 1
#endif
// prefetch ring depth (planes), plane tile of one field: 64 + 4 columns by
// RHS_TJ + 4 rows and a sink row for the slots of threads outside the field
#ifndef MR_R
#define MR_R 3
#endif
#define MT_W 68
#define MT_F (MT_W * (RHS_TJ + 5))
#define MT_SINK (MT_W * (RHS_TJ + 4))
template <bool BR>
__global__ __launch_bounds__(64 * RHS_TJ) __attribute__((amdgpu_waves_per_eu(4)))
void k_momentum_rhs_f32(Geo g, const float *__restrict__ Ut, const float *__restrict__ P, float *__restrict__ Q,
                        int kchunk, unsigned nbx, unsigned nby, unsigned nbz,
                        const unsigned char *__restrict__ brin, unsigned char *__restrict__ brout,
                        double *__restrict__ hpart)
{
    using T = float;
    // plane tiles of U, V, W and P (double-buffered by plane parity) and the
    // per-plane coefficients of the block's planes
    __shared__ float tile[2][4][MT_F];
    __shared__ float cf[5][RHS_KCHUNK_F32];
    const unsigned L = xcd_remap(blockIdx.x, nbx * nby * nbz);
    const unsigned bx = L % nbx, by = (L / nbx) % nby, bz = L / (nbx * nby);
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x, lx = t & 63, ly = t >> 6;
    // tile columns start at i0 = 64 bx (cache-line aligned row segments); the
    // first tile's column 0 is a wall column, loaded into the tile as the x
    // neighbour of i = 1 but not updated
    const int i0 = bx * 64, j0 = by * RHS_TJ + 1;
    const int ti = min(64, nX - 1 - i0), tj = min(RHS_TJ, nY - 1 - j0);   // tile columns and live rows
    // every thread marches the same planes (uniform loop bounds and plane
    // offsets keep the buffer operands scalar); a thread outside the field
    // works on the clamped column, stores to an offset past the buffer's
    // range (discarded by the hardware), writes its tile slot to a sink and
    // has its residual terms masked
    const bool intile = lx < ti && ly < tj;
    const bool live = intile && i0 + lx >= 1;
    const int i = i0 + min(lx, ti - 1), j = j0 + min(ly, tj - 1);
    const int k0 = 1 + bz * kchunk;
    const int k1 = min(k0 + kchunk, nZ - 1);
    // planes the block reads: k0-3 .. k0+kchunk+MR_R+1 (clamped)
    const int kb = max(k0 - 3, 0), ke = min(k0 + kchunk + MR_R + 1, nZ - 1);
    const unsigned PS = (unsigned)g.sZ * 4u;
    const unsigned range = (unsigned)(ke - kb + 1) * PS;
    const long N = g.N, base = (long)kb * g.sZ;
    __amdgpu_buffer_rsrc_t R[4], RQ[3];
#pragma unroll
    for (int p = 0; p < 3; p++) {
        R[p] = buf_of(Ut + p * N + base, range);
        RQ[p] = buf_of(Q + p * N + base, range);
    }
    R[3] = buf_of(P + base, range);
    auto so = [&](int kk) { return (unsigned)(max(0, min(kk, nZ - 1)) - kb) * PS; };
    const T Re = T(g.Re), dx = T(g.dx), dy = T(g.dy), invdx = T(g.invdx), invdy = T(g.invdy),
            invdXi = T(g.invdXi), c12 = T(g.c12), dt = T(g.dt);
    const T Ae = T(g.Ae), An = T(g.An), dtRe = dt / Re;
    // JacInv, dz3, Ab, At, 1 - dt/Re Ap of the block's planes
    if (t < kchunk) {
        const int kk = min(k0 + t, nZ - 2);
        cf[0][t] = T(g.JacInv[kk]);
        cf[1][t] = T(g.x3[kk + 1] - g.x3[kk]);
        cf[2][t] = T(g.Ab[kk]);
        cf[3][t] = T(g.At[kk]);
        cf[4][t] = T(1.0) - dt / Re * T(g.Ap[kk]);
    }
    const unsigned oc = (unsigned)(i + nX * j) * 4u;             // centre, in-plane bytes
    const unsigned ost = live ? oc : 0x80000000u;                 // store offset (range < 2^31 bytes)
    // residual mask: a thread outside the field reads tile slots no thread
    // wrote (any bit pattern, NaN included), so its terms are cleared with a
    // bitwise AND, not scaled by zero
    const unsigned wl = live ? 0xFFFFFFFFu : 0u;
    const int cslot = intile ? (2 + ly) * MT_W + 2 + lx : MT_SINK;  // tile slot of the centre
    // the thread's halo element (at most one per field and plane): span
    // halo rows j0-2, j0-1 and the two rows after the last live row
    // (periodic: virtual row v is row ((v - 1) mod (nY-2)) + 1), streamwise
    // halo columns i0-2, i0-1 and the two after the last tile column
    // (clamped to the field; the clamped values are never used)
    int hslot = MT_SINK;
    unsigned hoff = 0x80000000u;                                  // no element: load returns 0
    {
        const int n = nY - 2;
        auto prow = [&](int v) { return ((v - 1) % n + n) % n + 1; };
        if (t < 256) {
            const int hr = t >> 6, hc = t & 63;
            const int row = hr < 2 ? hr : tj + hr;
            const int v = hr < 2 ? j0 - 2 + hr : j0 + tj + hr - 2;
            if (hc < ti) {
                hslot = row * MT_W + 2 + hc;
                hoff = (unsigned)(prow(v) * nX + i0 + hc) * 4u;
            }
        } else if (t < 256 + 4 * RHS_TJ) {
            const int u = t - 256, hr = u >> 2, hcl = u & 3;
            const int col = hcl < 2 ? hcl : ti + hcl;
            const int ii = hcl < 2 ? i0 - 2 + hcl : i0 + ti + hcl - 2;
            if (hr < tj) {
                hslot = (2 + hr) * MT_W + col;
                hoff = (unsigned)((j0 + hr) * nX + min(max(ii, 0), nX - 1)) * 4u;
            }
        }
    }
    const bool xedge = (i <= 1 || i >= nX - 2);
    const bool yedge = (j == 1 || j == nY - 2);
    T ha[3] = {T(0), T(0), T(0)}, hq[3] = {T(0), T(0), T(0)};
    // queues (before step k's shift): planes k-3..k+1 of the components,
    // k-2..k of the pressure
    T zq[3][5], pq[3];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const unsigned s = so(k0 - 3 + q);
#pragma unroll
        for (int p = 0; p < 3; p++) zq[p][q] = bld(R[p], oc, s);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) pq[q] = bld(R[3], oc, so(k0 - 2 + q));
    // prefetch ring, slot s for the steps k = k0 + s (mod MR_R): centres of
    // plane k+2 (components) and k+1 (pressure), halo elements of plane k
    T rc[MR_R][3], rp[MR_R], rh[MR_R][4];
#pragma unroll
    for (int s = 0; s < MR_R; s++) {
#pragma unroll
        for (int p = 0; p < 3; p++) rc[s][p] = bld(R[p], oc, so(k0 + s + 2));
        rp[s] = bld(R[3], oc, so(k0 + s + 1));
#pragma unroll
        for (int f = 0; f < 4; f++) rh[s][f] = bld(R[f], hoff, so(k0 + s));
    }
    // clear the tiles once (slots outside a partial tile are never written
    // but are read by threads outside the field); the barrier orders the
    // clearing before the first plane's writes.  cf is complete at that
    // barrier too.
    for (int e = t; e < 2 * 4 * MT_F; e += 64 * RHS_TJ) (&tile[0][0][0])[e] = 0.0f;
    lds_barrier();
    const int a = ly * MT_W + lx;   // neighbour (i + di, j + dj) at a + (2 + dj) MT_W + 2 + di
    auto step = [&](auto S, int k) {
        constexpr int s = decltype(S)::value;
        // queues: planes k-2..k+2 and k-1..k+1 (values loaded MR_R planes
        // ago).  A value enters the queue by an explicit register move: the
        // queue keeps it for several planes, and the slot's own registers
        // must be free for the refill below (otherwise the loop copies
        // values still in flight)
#pragma unroll
        for (int p = 0; p < 3; p++) {
#pragma unroll
            for (int q = 0; q < 4; q++) zq[p][q] = zq[p][q + 1];
            asm("v_mov_b32 %0, %1" : "=v"(zq[p][4]) : "v"(rc[s][p]));
        }
        pq[0] = pq[1]; pq[1] = pq[2];
        asm("v_mov_b32 %0, %1" : "=v"(pq[2]) : "v"(rp[s]));
        // plane k into the tile
        float *tb = &tile[k & 1][0][0];
#pragma unroll
        for (int p = 0; p < 3; p++) tb[p * MT_F + cslot] = zq[p][2];
        tb[3 * MT_F + cslot] = pq[1];
#pragma unroll
        for (int f = 0; f < 4; f++) tb[f * MT_F + hslot] = rh[s][f];
        __builtin_amdgcn_sched_barrier(0);
        // refill the slot for plane k + MR_R; past the last plane the chunk
        // needs (centres k1+1, pressure k1, halos k1-1) a refill repeats that
        // plane, a cache hit, since those values are never used
#pragma unroll
        for (int p = 0; p < 3; p++) rc[s][p] = bld(R[p], oc, so(min(k + MR_R + 2, k1 + 1)));
        rp[s] = bld(R[3], oc, so(min(k + MR_R + 1, k1)));
#pragma unroll
        for (int f = 0; f < 4; f++) rh[s][f] = bld(R[f], hoff, so(min(k + MR_R, k1 - 1)));
        lds_barrier();
        const T u = zq[0][2], v = zq[1][2], w = zq[2][2];
        const bool zedge = (k <= 1 || k >= nZ - 2);
        const int kl = k - k0;
        const T JacInv = cf[0][kl], dz3 = cf[1][kl], Abk = cf[2][kl], Atk = cf[3][kl], aP = cf[4][kl];
        bool kx = tabs(Re * u * dx) > T(2.0), ky = tabs(Re * v * dy) > T(2.0), kz = tabs(Re * w * dz3) > T(2.0);
        const long c = (long)k * g.sZ + (long)(i + nX * j);
        if (BR && brin) {
            const unsigned b = brin[c];
            kx = (b & 1u) != 0; ky = (b & 2u) != 0; kz = (b & 4u) != 0;
        }
        if (BR && brout && live) brout[c] = (unsigned char)((kx ? 1u : 0u) | (ky ? 2u : 0u) | (kz ? 4u : 0u));
        const T ucx = u * (c12 * invdx), uax = kx ? tabs(u) * (T(0.25) * invdx) : T(0), uhx = u * (T(0.5) * invdx);
        const T ucy = v * (c12 * invdy), uay = ky ? tabs(v) * (T(0.25) * invdy) : T(0), uhy = v * (T(0.5) * invdy);
        const T ucz = w * (c12 * invdXi), uaz = kz ? tabs(w) * (T(0.25) * invdXi) : T(0), uhz = w * (T(0.5) * invdXi);
        const T *tp = tb + 3 * MT_F + a;
        const T px1 = tp[2 * MT_W + 3], px0 = tp[2 * MT_W + 1], py1 = tp[3 * MT_W + 2], py0 = tp[MT_W + 2];
        const unsigned sk = so(k);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const T *tf = tb + p * MT_F + a;
            const T f0 = zq[p][2];
            const T xm2 = tf[2 * MT_W], xm1 = tf[2 * MT_W + 1], xp1 = tf[2 * MT_W + 3], xp2 = tf[2 * MT_W + 4];
            const T ym2 = tf[2], ym1 = tf[MT_W + 2], yp1 = tf[3 * MT_W + 2], yp2 = tf[4 * MT_W + 2];
            const T zm2 = zq[p][0], zm1 = zq[p][1], zp1 = zq[p][3], zp2 = zq[p][4];
            T sx, sy, sz;
            const T rt1 = conv_fma(xm2, xm1, f0, xp1, xp2, ucx, uax, uhx, xedge, sx);
            const T rt2 = conv_fma(ym2, ym1, f0, yp1, yp2, ucy, uay, uhy, yedge, sy);
            const T rt3 = conv_fma(zm2, zm1, f0, zp1, zp2, ucz, uaz, uhz, zedge, sz);
            const T cv = __builtin_fmaf(rt3, JacInv, rt1 + rt2);
            T gp;
            if (p == 0) gp = (px1 - px0) * (T(0.5) * invdx);
            else if (p == 1) gp = (py1 - py0) * (T(0.5) * invdy);
            else gp = (pq[2] - pq[0]) * (T(0.5) * invdXi * JacInv);
            const T qv = __builtin_fmaf(-dt, cv + gp, f0);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, qv), RQ[p], ost, sk, 0);
            // first Helmholtz residual (summed only when hpart is given)
            // Ae = Aw = 1/dx^2 and An = As = 1/dy^2 (set so at create time)
            const T rt1h = __builtin_fmaf(An, sy, Ae * sx);
            const T rt2h = __builtin_fmaf(Abk, zm1, Atk * zp1);
            const T res = __builtin_fmaf(dtRe, rt1h + rt2h, __builtin_fmaf(-aP, f0, qv));
            ha[p] += __builtin_bit_cast(T, __builtin_bit_cast(unsigned, tabs(res)) & wl);
            hq[p] += __builtin_bit_cast(T, __builtin_bit_cast(unsigned, tabs(qv)) & wl);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    int k = k0;
    for (; k + MR_R <= k1; k += MR_R) ring_steps<0, MR_R>(step, k);
    ring_tail<0, MR_R>(step, k, k1);
    if (hpart) {
        __shared__ double hs[RHS_TJ];
        const unsigned nb = gridDim.x;
        for (int p = 0; p < 3; p++) {
            const double s1 = block_sum<64 * RHS_TJ>((double)ha[p], hs, threadIdx.x);
            if (threadIdx.x == 0) hpart[(long)p * nb + blockIdx.x] = s1;
            const double s2 = block_sum<64 * RHS_TJ>((double)hq[p], hs, threadIdx.x);
            if (threadIdx.x == 0) hpart[(long)(3 + p) * nb + blockIdx.x] = s2;
        }
    }
}

// ----------------------------------------------------------------------------
// a5: implicit diffusion (dnssetbl.f90:371-433), red-black Gauss-Seidel.
// Colour of a node = parity of its 1-based i+j+k; red (even) first.
// ----------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void helm_terms(const Geo &g, const T *u, long c, int j, int k,
                                           T &rt1, T &rt2)
{
    const int nX = g.nX, nY = g.nY;
    const long cp = c + (long)(jp1_of(j, nY) - j) * nX, cm = c + (long)(jm1_of(j, nY) - j) * nX;
    rt1 = T(g.Ae) * u[c + 1] + T(g.Aw) * u[c - 1] + T(g.An) * u[cp] + T(g.As) * u[cm];
    rt2 = T(g.At[k]) * u[c + g.sZ] + T(g.Ab[k]) * u[c - g.sZ];
}

// In-plane neighbours of a (64 x RT_J)-column tile through LDS: plane k of
// the tile and its one-node halo ((64+2) x (RT_J+2) values, the span halo rows
// from the periodic tables) is staged cooperatively, so that each value is
// loaded from memory once per tile instead of by up to five threads.  The
// staging is pipelined: a thread's tile values of plane k+2 are loaded into
// registers while plane k is computed and written to LDS one plane later, so
// no barrier waits on memory; three plane buffers rotate (plane k+1 is
// written while plane k is read, and a buffer is overwritten only after the
// barrier that ends its last read).
#define TL_X (64 + 2)
#define TL_Y (RT_J + 2)
#define TL_N (TL_X * TL_Y)
#define TL_PER ((TL_N + RT_T - 1) / RT_T)                        // tile values per thread
// global (0-based) span row of tile row ty (0 = the halo below row j0 = by RT_J + 1)
__device__ __forceinline__ int tile_row(int ty, int j0, int nY)
{
    if (ty == 0) return jm1_of(j0, nY);
    const int j = j0 + ty - 1;
    if (j <= nY - 2) return j;
    return j == nY - 1 ? 1 : -1;                // the halo above the last row, or unused
}
// in-plane offsets of this thread's tile values (-1: outside the field or
// unused); tile column 0 is field column i0 (i0 = -1 for the first tile)
__device__ __forceinline__ void tile_offsets(int *off, int t, int i0, int j0, const Geo &g)
{
#pragma unroll
    for (int q = 0; q < TL_PER; q++) {
        const int e = t + q * RT_T, ty = e / TL_X, tx = e - ty * TL_X;
        const int i = i0 + tx, j = e < TL_N ? tile_row(ty, j0, g.nY) : -1;
        off[q] = (i >= 0 && i <= g.nX - 1 && j >= 0) ? i + g.nX * j : -1;
    }
}
// loads are unconditional (unused slots read the plane's first value), so
// that the compiler's in-order wait counts stay exact across the loop
template <class V, class F>
__device__ __forceinline__ void tile_load(V *reg, const int *off, int k, const Geo &g, F val)
{
#pragma unroll
    for (int q = 0; q < TL_PER; q++) reg[q] = val(max(off[q], 0) + (int)g.sZ * k);
}
template <class V>
__device__ __forceinline__ void tile_store(V *buf, const V *reg, const int *off, int t)
{
#pragma unroll
    for (int q = 0; q < TL_PER; q++)
        if (off[q] >= 0) buf[t + q * RT_T] = reg[q];
}

struct GsCtrl {
    double sumnor[3], res[3];
    double qnorm[3];         // L1 norm of the right-hand side (first residual pass)
    int active[3], sweeps[3];
};

// L1 residual of every active component (dnssetbl.f90:378-386): block
// partials (and, on the first pass, the L1 norm of Q).  Each block covers a
// 64 x RT_J (i,j) tile and a chunk of k planes.
template <class T>
__global__ __launch_bounds__(RT_T) void k_helm_resid(Geo g, const T *__restrict__ Q, const T *__restrict__ Ut,
                                                   const GsCtrl *__restrict__ ctrl, double *__restrict__ part, int first,
                                                   int kchunk)
{
    __shared__ double sh[RT_T / 64];
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63);
    const int j = blockIdx.y * RT_J + (t >> 6) + 1;
    const int k0 = 1 + blockIdx.z * kchunk, k1 = min(k0 + kchunk, nZ - 1);
    const bool in = (i >= 1 && i <= nX - 2 && j <= nY - 2);
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const T dt = T(g.dt), Re = T(g.Re);
    int act[3];
    for (int p = 0; p < 3; p++) act[p] = ctrl->active[p];
    double a[3] = {0.0, 0.0, 0.0}, aq[3] = {0.0, 0.0, 0.0};
    if (in) {
        // one component at a time (a third of the L2 working set for the
        // halo reuse between neighbouring tiles); u(k-1), u(k), u(k+1) rotate
        // through registers
        const int sk = (int)g.sZ;                                // 3N < 2^31 (checked at create)
        const int dn = (jp1_of(j, nY) - j) * nX, ds = (jm1_of(j, nY) - j) * nX;
        const int c0 = i + nX * j;
        const T Ae = T(g.Ae), Aw = T(g.Aw), An = T(g.An), As = T(g.As);
#pragma unroll
        for (int p = 0; p < 3; p++) {
            if (!act[p]) continue;                               // uniform across the grid
            const T *u = Ut + (long)p * g.N;
            const T *qp = Q + (long)p * g.N;
            T zm = u[c0 + sk * (k0 - 1)], z0 = u[c0 + sk * k0];
            double ap = 0.0, aqp = 0.0;
            for (int k = k0; k < k1; k++) {
                const int c = c0 + sk * k;
                const T zp = u[c + sk];
                const T rt1 = Ae * u[c + 1] + Aw * u[c - 1] + An * u[c + dn] + As * u[c + ds];
                const T rt2 = T(g.At[k]) * zp + T(g.Ab[k]) * zm;
                const T q = qp[c];
                T resLoc = q - (T(1.0) - dt / Re * T(g.Ap[k])) * z0 + dt * (rt1 + rt2) / Re;
                ap += (double)tabs(resLoc);
                aqp += (double)tabs(q);
                zm = z0;
                z0 = zp;
            }
            a[p] = ap;
            aq[p] = aqp;
        }
    }
    for (int p = 0; p < 3; p++) {
        if (!act[p]) continue;                   // uniform across the grid
        double s = block_sum<RT_T>(a[p], sh, t);
        if (t == 0) part[(long)p * nb + b] = s;
        if (first) {
            double sq = block_sum<RT_T>(aq[p], sh, t);
            if (t == 0) part[(long)(3 + p) * nb + b] = sq;
        }
    }
}

// Stopping rule of dnssetbl.f90:389-392.  floor_coef > 0 (FP32 storage only)
// adds: stop once the L1 residual is at the roundoff level of the stored
// field, floor_coef * ||Q||_1, which the reference's 1e-6 relative target
// cannot go below in single precision.
__global__ __launch_bounds__(1024) void k_helm_fin(GsCtrl *ctrl, const double *__restrict__ part, int nb, int it,
                                                  double errorstop, double floor_coef)
{
    __shared__ double sh[16];
    const double tiny = (double)1e-12f;    // REAL(4) literal, dnssetbl.f90:392
    for (int p = 0; p < 3; p++) {
        if (!ctrl->active[p]) continue;
        double a = 0.0;
        // fixed per-thread strided order, then fixed block reduction: deterministic
#pragma unroll 8
        for (int q = threadIdx.x; q < nb; q += 1024) a += part[(long)p * nb + q];
        double res = block_sum<1024>(a, sh, threadIdx.x);
        double qn = 0.0;
        if (it == 1) {
            double aq = 0.0;
            for (int q = threadIdx.x; q < nb; q += 1024) aq += part[(long)(3 + p) * nb + q];
            qn = block_sum<1024>(aq, sh, threadIdx.x);
        }
        if (threadIdx.x == 0) {
            if (it == 1) { ctrl->sumnor[p] = res; ctrl->qnorm[p] = qn; }
            ctrl->res[p] = res;
            double sumav = res / ctrl->sumnor[p];
            if (sumav < errorstop || res < tiny || (floor_coef > 0.0 && res <= floor_coef * ctrl->qnorm[p])) {
                ctrl->active[p] = 0;
                ctrl->sweeps[p] = it - 1;
            }
        }
        __syncthreads();
    }
}

// one colour of a red-black sweep; thread per same-colour node along i
template <class T>
__global__ __launch_bounds__(256) void k_helm_sweep(Geo g, const T *__restrict__ Q, T *__restrict__ Ut,
                                                   const GsCtrl *__restrict__ ctrl, int colour)
{
    const int nX = g.nX, nY = g.nY;
    const int t = threadIdx.x;
    const int j = blockIdx.y * 4 + (t >> 6) + 1;
    const int k = blockIdx.z + 1;
    if (j > nY - 2) return;
    const int ifirst = (((j + k) & 1) == colour) ? 1 : 2;
    const int i = ifirst + 2 * (blockIdx.x * 64 + (t & 63));
    if (i > nX - 2) return;
    const long c = (long)i + (long)nX * ((long)j + (long)nY * k);
    const T dt = T(g.dt), Re = T(g.Re);
    for (int p = 0; p < 3; p++) {
        if (!ctrl->active[p]) continue;
        T *u = Ut + (long)p * g.N;
        T rt1, rt2;
        helm_terms(g, u, c, j, k, rt1, rt2);
        u[c] = (Q[c + (long)p * g.N] + dt / Re * (rt1 + rt2)) / (T(1.0) - dt / Re * T(g.Ap[k]));
    }
}

// one colour of a z-line red-black Gauss-Seidel sweep: every (i,j) column of
// the colour (parity of 1-based i+j) solves its tridiagonal wall-normal system
// exactly, with the x/y neighbours at their current values.  The strong
// wall-normal coupling of the stretched grid is then solved directly, which
// is what makes the point sweep slow.  The line matrix
//   a(k) = -h Ab(k), b(k) = 1 - h Ap(k), c(k) = -h At(k),  h = dt/Re,
// is the same for every column.  The interior planes are split into C
// balanced chunks of at most LC planes; a workgroup is 64 columns (one per
// lane) x C wavefronts (one per chunk).  Each wavefront eliminates its chunk
// in registers so that every row depends only on the chunk's first and last
// unknowns (partitioned Thomas); the 2C first/last unknowns of a column are
// solved through LDS by wavefront 0 (Thomas); every wavefront then recovers
// and stores its rows.  Each value is read and written once per launch: no
// elimination intermediates go through HBM.  All coefficient parts are
// precomputed on the host (zf layout: a, R, c', a'', c'' at [0,5nZ) in
// blocks of nZ, then c0 [C], then the reduced factors rA, rg, re [2C each]).
#define ZL_MAXC 16
template <class T, int LC>
__global__ __launch_bounds__(64 * ZL_MAXC) void k_helm_zline(Geo g, const T *__restrict__ Q, T *__restrict__ Ut,
                                                             const T *__restrict__ zf, const GsCtrl *__restrict__ ctrl,
                                                             int colour, int C, unsigned nbx, unsigned nblk,
                                                             unsigned pw)
{
    __shared__ T sD[2 * ZL_MAXC * 64];
    __shared__ T sF[3 * 2 * ZL_MAXC];                            // reduced-system factors rA, rg, re
    const int nX = g.nX, nY = g.nY, nZ = g.nZ, nz = nZ - 2;
    // blocks run span-row-first inside panels of pw column tiles, so that the
    // rows j-1, j, j+1 (the span neighbours of row j) are swept at nearly the
    // same time and their second and third reads hit L2
    const unsigned L = xcd_remap(blockIdx.x, nblk);
    const unsigned nyr = nblk / nbx;
    const unsigned P = L / (pw * nyr), w_ = min(pw, nbx - P * pw), rem = L - P * pw * nyr;
    const int bx = (int)(P * pw + rem % w_), j = (int)(rem / w_) + 1;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // chunk of this wavefront
    const int ifirst = (((colour + j) & 1) == 1) ? 1 : 2;
    const int i = ifirst + 2 * (bx * 64 + lane);
    const bool live = i <= nX - 2;
    const int kb = 1 + (int)(((long)w * nz) / C), M = 1 + (int)(((long)(w + 1) * nz) / C) - kb;
    // 32-bit offsets inside one component (N < 2^31 is checked at create time)
    const int sk = (int)g.sZ;
    const long N = g.N;
    const int c0 = i + nX * j;                                    // k = 0 plane
    const int dn = (jp1_of(j, nY) - j) * nX, ds = (jm1_of(j, nY) - j) * nX;
    const T h = T(g.dt) / T(g.Re);
    const T Ae = T(g.Ae), Aw = T(g.Aw), An = T(g.An), As = T(g.As);
    const T *zA = zf, *zR = zf + nZ, *zC = zf + 2 * nZ, *zA2 = zf + 3 * nZ, *zC2 = zf + 4 * nZ;
    const T *zc0 = zf + 5 * nZ, *rA = zc0 + C, *rg = rA + 2 * C, *re = rg + 2 * C;
    // the reduced system padded to 2 ZL_MAXC unknowns with identity rows
    // (rA = 0, rg = 1, re = 0, right-hand side 0), so that its two sweeps are
    // straight-line code whose LDS reads can be issued ahead (made visible
    // by the barrier after the first forward elimination)
    for (int e = threadIdx.x; e < 2 * ZL_MAXC; e += blockDim.x) {
        const bool in = e < 2 * C;
        sF[e] = (in && e > 0) ? rA[e] : T(0);                     // the first row has no sub-diagonal
        sF[2 * ZL_MAXC + e] = in ? rg[e] : T(1);
        sF[4 * ZL_MAXC + e] = in ? re[e] : T(0);
    }
    struct Op { T q, e, w, n, s; };
    for (int p = 0; p < 3; p++) {
        if (!ctrl->active[p]) continue;                          // uniform over the grid
        T *u = Ut + (long)p * N;
        const T *q = Q + (long)p * N;
        // per-lane bases at the chunk's first plane; every other offset
        // (planes, span neighbours) is wave-uniform and stays in scalar registers
        const T *qb = q + (c0 + sk * kb);
        T *ub = u + (c0 + sk * kb);
        auto ld = [&](int r, Op &o) {
            const int o_ = sk * r;
            o.q = qb[o_]; o.e = ub[o_ + 1]; o.w = ub[o_ - 1]; o.n = ub[o_ + dn]; o.s = ub[o_ + ds];
        };
        T dp[LC];
        if (live) {
            // forward elimination; the operands of the next ZPF planes are in
            // flight (the scheduling barrier keeps the compiler from hoisting
            // every load of the chunk, which would exceed the register budget)
            constexpr int ZPF = 2;
            Op ring[ZPF];
#pragma unroll
            for (int r = 0; r < ZPF; r++)
                if (r < M) ld(r, ring[r]);
#pragma unroll
            for (int r = 0; r < LC; r++) {
                if (r < M) {
                    const int k = kb + r;
                    const Op cur = ring[r % ZPF];
                    if (r + ZPF < M) ld(r + ZPF, ring[r % ZPF]);
                    T rhs = cur.q + h * (Ae * cur.e + Aw * cur.w + An * cur.n + As * cur.s);
                    if (k == 1) rhs = rhs + h * (T(g.Ab[k]) * ub[sk * r - sk]);
                    if (k == nZ - 2) rhs = rhs + h * (T(g.At[k]) * ub[sk * r + sk]);
                    if (r == 0) dp[0] = rhs;
                    else dp[r] = (rhs - zA[k] * dp[r - 1]) * zR[k];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int r = LC - 3; r >= 1; r--)
                if (r <= M - 3) dp[r] = dp[r] - zC[kb + r] * dp[r + 1];
            T DL = dp[1];
#pragma unroll
            for (int r = 2; r < LC; r++)
                if (r == M - 1) DL = dp[r];
            sD[(2 * w) * 64 + lane] = dp[0] - zc0[w] * dp[1];
            sD[(2 * w + 1) * 64 + lane] = DL;
        }
        __syncthreads();
        if (w == 0 && live) {
            T y = sD[lane] * sF[2 * ZL_MAXC];
            sD[lane] = y;
#pragma unroll
            for (int s2 = 1; s2 < 2 * ZL_MAXC; s2++) {
                if (s2 < 2 * C) {
                    y = (sD[s2 * 64 + lane] - sF[s2] * y) * sF[2 * ZL_MAXC + s2];
                    sD[s2 * 64 + lane] = y;
                }
            }
            T x = y;
#pragma unroll
            for (int s2 = 2 * ZL_MAXC - 2; s2 >= 0; s2--) {
                if (s2 <= 2 * C - 2) {
                    x = sD[s2 * 64 + lane] - sF[4 * ZL_MAXC + s2] * x;
                    sD[s2 * 64 + lane] = x;
                }
            }
        }
        __syncthreads();
        if (live) {
            // the other colour's value in this lane's aligned column pair
            // (2m, 2m+1) is stored back unchanged with the new value, so that
            // rows are written in whole lines (no thread changes it during
            // this colour pass)
            const int po = (i & 1) ? -1 : 1;
            const T xf = sD[(2 * w) * 64 + lane], xl = sD[(2 * w + 1) * 64 + lane];
#pragma unroll
            for (int r = 0; r < LC; r++) {
                if (r < M) {
                    const int k = kb + r;
                    T x;
                    if (r == 0) x = xf;
                    else if (r == M - 1) x = xl;
                    else x = dp[r] - zA2[k] * xf - zC2[k] * xl;
                    const T o = ub[sk * r + po];
                    ub[sk * r] = x;
                    ub[sk * r + po] = o;
                }
            }
        }
        __syncthreads();                                         // sD is reused by the next component
    }
}

// One pass of a z-line zebra Gauss-Seidel sweep over span rows: pass
// `colour` solves every wall-normal column of the rows with parity(1-based j)
// == colour exactly (the same partitioned Thomas solve and factors as
// k_helm_zline), with the x neighbours (i-1, i+1: columns of the same rows)
// at their values from the start of the pass and the span neighbours (the
// other rows) at their current values.  Reads come from `cur` (x neighbours,
// the column's wall and top values) and `ysrc` (span rows: `cur` in the first
// pass, `dst` in the second), writes go to `dst`, so no value is read after
// another block has overwritten it in the same pass.  A workgroup is 64
// consecutive columns of one row (one per lane) x C wavefronts (one per
// chunk), so every load and store is a whole, coalesced row segment.  A
// component whose stopping test has ended is copied through unchanged, so
// that both buffers stay consistent and the host swaps them after each
// sweep.
template <class T, int LC, int ZPF>
// FP32 with 16-plane chunks: registers allocated for six waves per SIMD (80
// VGPRs, no spill), so that two 12-wave workgroups share a CU
__global__ __launch_bounds__(64 * ZL_MAXC)
__attribute__((amdgpu_waves_per_eu((sizeof(T) == 4 && LC == 16) ? 6 : 1))) void k_helm_zebra(Geo g, const T *__restrict__ Q, const T *__restrict__ cur,
                                                             const T *ysrc, T *dst, const T *__restrict__ zf,
                                                             const GsCtrl *__restrict__ ctrl, int colour, int C,
                                                             unsigned nbx, unsigned nblk, unsigned pw)
{
    __shared__ T sD[2 * ZL_MAXC * 64];
    __shared__ T sF[3 * 2 * ZL_MAXC];                            // reduced-system factors rA, rg, re
    const int nX = g.nX, nY = g.nY, nZ = g.nZ, nz = nZ - 2;
    // blocks run row-first inside panels of pw column tiles, so that the rows
    // j and j+2 of this colour, which share the span neighbour row j+1, are
    // swept at nearly the same time and its second read hits L2
    const unsigned L = xcd_remap(blockIdx.x, nblk);
    const unsigned nyr = nblk / nbx;
    const unsigned P = L / (pw * nyr), w_ = min(pw, nbx - P * pw), rem = L - P * pw * nyr;
    // ZB_REV: the second colour's rows run from the last to the first, so
    // that each pass starts where the previous one ended (the rows of one
    // colour are independent within a pass: the order changes no result)
    const unsigned rr = rem / w_, ro = (ZB_REV && colour) ? nyr - 1 - rr : rr;
    const int bx = (int)(P * pw + rem % w_), j = 1 + 2 * (int)ro + colour;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // chunk of this wavefront
    // tiles start at i = 64 bx, so that a tile's row segments are whole
    // cache lines (Q, the span rows and the new values: two 128-byte lines
    // per FP32 segment instead of three); the boundary columns i = 0 and
    // nX-1 are lanes that solve nothing, and each interior column is still
    // solved by one lane
    const int i = bx * 64 + lane;
    const bool live = i >= 1 && i <= nX - 2;
    const int kb = 1 + (int)(((long)w * nz) / C), M = 1 + (int)(((long)(w + 1) * nz) / C) - kb;
    // 32-bit offsets inside one component (N < 2^31 is checked at create time)
    const int sk = (int)g.sZ;
    const long N = g.N;
    const int c0 = i + nX * j;                                    // k = 0 plane
    const int dn = (jp1_of(j, nY) - j) * nX, ds = (jm1_of(j, nY) - j) * nX;
    const T h = T(g.dt) / T(g.Re);
    const T Ae = T(g.Ae), Aw = T(g.Aw), An = T(g.An), As = T(g.As);
    const T *zA = zf, *zR = zf + nZ, *zC = zf + 2 * nZ, *zA2 = zf + 3 * nZ, *zC2 = zf + 4 * nZ;
    const T *zc0 = zf + 5 * nZ, *rA = zc0 + C, *rg = rA + 2 * C, *re = rg + 2 * C;
    for (int e = threadIdx.x; e < 2 * ZL_MAXC; e += blockDim.x) {
        const bool in = e < 2 * C;
        sF[e] = (in && e > 0) ? rA[e] : T(0);                     // the first row has no sub-diagonal
        sF[2 * ZL_MAXC + e] = in ? rg[e] : T(1);
        sF[4 * ZL_MAXC + e] = in ? re[e] : T(0);
    }
    struct Op { T q, e, w, n, s; };
    for (int p = 0; p < 3; p++) {
        // per-lane bases at the chunk's first plane; every other offset
        // (planes, neighbours, the other arrays) is wave-uniform
        const T *cb = cur + (long)p * N + (c0 + sk * kb);
        if (!ctrl->active[p]) {                                  // uniform over the grid
            T *db = dst + (long)p * N + (c0 + sk * kb);
            if (live)
                for (int r = 0; r < M; r++) db[sk * r] = cb[sk * r];
            continue;
        }
        const T *qb = Q + (long)p * N + (c0 + sk * kb);
        const T *yb = ysrc + (long)p * N + (c0 + sk * kb);
        auto ld = [&](int r, Op &o) {
            const int o_ = sk * r;
            o.q = qb[o_]; o.e = cb[o_ + 1]; o.w = cb[o_ - 1]; o.n = yb[o_ + dn]; o.s = yb[o_ + ds];
        };
        T dp[LC];
        if (live) {
            // the operands of the next ZPF planes are in flight.  Every load
            // is unconditional (plane indices clamped to the chunk; the wall
            // and top values, used by the first and last chunk only, are
            // loaded once up front): a branch around a load makes the
            // compiler's in-order wait counts drain the whole ring
            const T wall = cb[-sk], top = cb[sk * M];
            const T Abw = T(g.Ab[kb]), Att = T(g.At[kb + M - 1]);
            // ZPF + 1 slots: the slot refilled in step r is never the one
            // step r reads, so the new loads need not wait for its compute
            Op ring[ZPF + 1];
#pragma unroll
            for (int r = 0; r < ZPF; r++) ld(min(r, M - 1), ring[r]);
#pragma unroll
            for (int r = 0; r < LC; r++) {
                // outside the r < M test (M is below LC by at most one plane)
                ld(min(r + ZPF, M - 1), ring[(r + ZPF) % (ZPF + 1)]);
                if (r < M) {
                    const int k = kb + r;
                    const Op cur_ = ring[r % (ZPF + 1)];
                    T rhs = cur_.q + h * (Ae * cur_.e + Aw * cur_.w + An * cur_.n + As * cur_.s);
                    if (k == 1) rhs = rhs + h * (Abw * wall);
                    if (k == nZ - 2) rhs = rhs + h * (Att * top);
                    if (r == 0) dp[0] = rhs;
                    else dp[r] = (rhs - zA[k] * dp[r - 1]) * zR[k];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int r = LC - 3; r >= 1; r--)
                if (r <= M - 3) dp[r] = dp[r] - zC[kb + r] * dp[r + 1];
            T DL = dp[1];
#pragma unroll
            for (int r = 2; r < LC; r++)
                if (r == M - 1) DL = dp[r];
            sD[(2 * w) * 64 + lane] = dp[0] - zc0[w] * dp[1];
            sD[(2 * w + 1) * 64 + lane] = DL;
        }
        __syncthreads();
        if (w == 0 && live) {
            T y = sD[lane] * sF[2 * ZL_MAXC];
            sD[lane] = y;
#pragma unroll
            for (int s2 = 1; s2 < 2 * ZL_MAXC; s2++) {
                if (s2 < 2 * C) {
                    y = (sD[s2 * 64 + lane] - sF[s2] * y) * sF[2 * ZL_MAXC + s2];
                    sD[s2 * 64 + lane] = y;
                }
            }
            T x = y;
#pragma unroll
            for (int s2 = 2 * ZL_MAXC - 2; s2 >= 0; s2--) {
                if (s2 <= 2 * C - 2) {
                    x = sD[s2 * 64 + lane] - sF[4 * ZL_MAXC + s2] * x;
                    sD[s2 * 64 + lane] = x;
                }
            }
        }
        __syncthreads();
        if (live) {
            T *db = dst + (long)p * N + (c0 + sk * kb);
            const T xf = sD[(2 * w) * 64 + lane], xl = sD[(2 * w + 1) * 64 + lane];
#pragma unroll
            for (int r = 0; r < LC; r++) {
                if (r < M) {
                    const int k = kb + r;
                    T x;
                    if (r == 0) x = xf;
                    else if (r == M - 1) x = xl;
                    else x = dp[r] - zA2[k] * xf - zC2[k] * xl;
                    db[sk * r] = x;
                }
            }
        }
        __syncthreads();                                         // sD is reused by the next component
    }
}

// the boundary and ghost values (i = 0, nX-1 columns, j = 0, nY-1 rows,
// k = 0, nZ-1 planes) of the three velocity components, src -> dst
template <class T>
__global__ void k_copy_faces3(Geo g, const T *__restrict__ src, T *__restrict__ dst)
{
    const long nX = g.nX, nY = g.nY, nZ = g.nZ, N = g.N;
    const long fz = nX * nY, fy = nX * nZ, fx = nY * nZ, per = 2 * (fz + fy + fx);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * per) return;
    const long p = t / per;
    t -= p * per;
    long c;
    if (t < 2 * fz) { const long k = (t < fz) ? 0 : nZ - 1; c = t % fz + fz * k; }
    else if ((t -= 2 * fz) < 2 * fy) { const long j = (t < fy) ? 0 : nY - 1, q = t % fy; c = q % nX + nX * (j + nY * (q / nX)); }
    else { t -= 2 * fy; const long i = (t < fx) ? 0 : nX - 1, q = t % fx; c = i + nX * (q % nY + nY * (q / nY)); }
    dst[c + p * N] = src[c + p * N];
}

// ----------------------------------------------------------------------------
// a6: PPE right-hand side, face-average divergence / dt (dnssetbl.f90:437-447)
// computed in double from T inputs; output double.
// ----------------------------------------------------------------------------
// gate (optional): returns at once while any Gauss-Seidel component is still
// active, so the launch can be queued before the host has read the stopping test
template <class T>
__global__ __launch_bounds__(256) void k_ppe_rhs(Geo g, const T *__restrict__ Ut, double *__restrict__ pRhs,
                                                const GsCtrl *__restrict__ gate)
{
    if (gate && (gate->active[0] | gate->active[1] | gate->active[2])) return;
    const int nX = g.nX, nY = g.nY;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63);
    const int j = blockIdx.y * 4 + (t >> 6) + 1;
    const int k = blockIdx.z + 1;
    if (i < 1 || i > nX - 2 || j > nY - 2) return;
    const long c = (long)i + (long)nX * ((long)j + (long)nY * k), N = g.N, sk = g.sZ;
    const long cp = c + (long)(jp1_of(j, nY) - j) * nX, cm = c + (long)(jm1_of(j, nY) - j) * nX;
    const T *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    double rt1 = ((((double)U[c + 1] + (double)U[c]) * 0.5) - ((((double)U[c] + (double)U[c - 1]) * 0.5))) * g.invdx;
    double rt2 = ((((double)V[cp] + (double)V[c]) * 0.5) - ((((double)V[c] + (double)V[cm]) * 0.5))) * g.invdy;
    double rt3 = 2.0 * ((((double)W[c + sk] + (double)W[c]) * 0.5) - ((((double)W[c] + (double)W[c - sk]) * 0.5))) /
                 (g.x3[k + 1] - g.x3[k - 1]);
    pRhs[c] = (rt1 + rt2 + rt3) / g.dt;
}

// ----------------------------------------------------------------------------
// a7: Poisson operator  out = A x  on the interior (dnssetbl.f90:574-583).
// `active` (may be null) lets GMRES skip the launch once converged.
// ----------------------------------------------------------------------------
// `scale` (may be null): out = (A x) * (*scale) -- GMRES passes the scale
// sigma_j of the stored Krylov vector so that A sees the normalised vector.
template <class T>
__global__ __launch_bounds__(256) void k_poisson(Geo g, const T *__restrict__ x, T *__restrict__ out,
                                                const int *__restrict__ active, const double *__restrict__ scale,
                                                const T *__restrict__ v0, double *__restrict__ part,
                                                int kchunk, unsigned nbx, unsigned nby, unsigned nbz)
{
    if (active && !*active) return;
    __shared__ double sh[4];
    const unsigned L = xcd_remap(blockIdx.x, nbx * nby * nbz);
    const unsigned bx = L % nbx, by = (L / nbx) % nby, bz = L / (nbx * nby);
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int i = bx * 64 + (t & 63);
    const int j = by * 4 + (t >> 6) + 1;
    double acc = 0.0;                                            // <out, v0> (first Arnoldi step)
    if (i >= 1 && i <= nX - 2 && j <= nY - 2) {
        // each thread marches a chunk of k planes; the wall-normal neighbours
        // x(k-1), x(k), x(k+1) rotate through registers (each plane read once)
        const int k0 = 1 + (int)bz * kchunk, k1 = min(k0 + kchunk, nZ - 1);
        const int sk = (int)g.sZ;                                // N < 2^31 (checked at create)
        const int c0 = i + nX * j;
        const int dn = (jp1_of(j, nY) - j) * nX, ds = (jm1_of(j, nY) - j) * nX;
        const T Aw = T(g.Aw), Ae = T(g.Ae), As = T(g.As), An = T(g.An);
        T zm = x[c0 + sk * (k0 - 1)], z0 = x[c0 + sk * k0];
        for (int k = k0; k < k1; k++) {
            const int c = c0 + sk * k;
            const T zp = x[c + sk];
            const T ax = Aw * x[c - 1] + Ae * x[c + 1] + As * x[c + ds] + An * x[c + dn] +
                         T(g.Ab[k]) * zm + T(g.At[k]) * zp + T(g.Ap[k]) * z0;
            const T w = scale ? ax * T(*scale) : ax;
            out[c] = w;
            if (part) acc += (double)v0[c] * (double)w;
            zm = z0;
            z0 = zp;
        }
    }
    if (part) {
        const double s = block_sum<256>(acc, sh, t);
        if (t == 0) part[blockIdx.x] = s;
    }
}

// r = pRhs - A PCorr in double (dnssetbl.f90:498-507); stores r as Tr and
// block partials of ||r||^2 (double).  Boundary entries of r are left alone.
// With bpart, also block partials of ||pRhs||^2 (the first residual of a solve
// reads every interior pRhs value once, so the norm of b costs no extra pass).
template <class Tr>
__global__ __launch_bounds__(RT_T) void k_resid(Geo g, const double *__restrict__ pRhs,
                                              const double *__restrict__ PC, Tr *__restrict__ r,
                                              double *__restrict__ part, double *__restrict__ bpart, int kchunk)
{
    __shared__ double sh[RT_T / 64];
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63);
    const int j = blockIdx.y * RT_J + (t >> 6) + 1;
    const int k0 = 1 + blockIdx.z * kchunk, k1 = min(k0 + kchunk, nZ - 1);
    double a = 0.0, bb = 0.0;
    if (i >= 1 && i <= nX - 2 && j <= nY - 2) {
        const int sk = (int)g.sZ;                                // N < 2^31 (checked at create)
        const int dn = (jp1_of(j, nY) - j) * nX, ds = (jm1_of(j, nY) - j) * nX;
        const int c0 = i + nX * j;
        double zm = PC[c0 + sk * (k0 - 1)], z0 = PC[c0 + sk * k0];   // PCorr(k-1), PCorr(k) in registers
        for (int k = k0; k < k1; k++) {
            const int c = c0 + sk * k;
            const double zp = PC[c + sk];
            const double b = pRhs[c];
            double rr = b - (g.Aw * PC[c - 1] + g.Ae * PC[c + 1] + g.As * PC[c + ds] + g.An * PC[c + dn] +
                             g.Ab[k] * zm + g.At[k] * zp + g.Ap[k] * z0);
            if (r) r[c] = (Tr)rr;
            a += rr * rr;
            bb += b * b;
            zm = z0;
            z0 = zp;
        }
    }
    const unsigned blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    double s = block_sum<RT_T>(a, sh, t);
    if (t == 0) part[blk] = s;
    if (bpart) {                                                 // uniform over the grid
        s = block_sum<RT_T>(bb, sh, t);
        if (t == 0) bpart[blk] = s;
    }
}

// The restart residual of the next refinement cycle with the previous cycle's
// one-vector correction folded in: PCn = PC + y z0 (y = the cycle's solution
// coefficient, the same arithmetic as k_xupd) is formed at the seven stencil
// points from PC and z0, written to a second buffer (neighbours must see the
// old PC), and r = pRhs - A PCn.  Faces of PCn are not written (both buffers
// keep zero faces during a solve).  Bitwise the same as k_xupd then k_resid.
// The updated in-plane neighbours come from an LDS tile (see tile_load).
template <class Tr>
__global__ __launch_bounds__(RT_T) void k_resid_upd(Geo g, const double *__restrict__ pRhs,
                                                  const double *__restrict__ PC, const Tr *__restrict__ Z,
                                                  const GmCtrl *__restrict__ gc, double *__restrict__ PCn,
                                                  Tr *__restrict__ r, double *__restrict__ part, int kchunk,
                                                  int gated)
{
    // gated: queued before the host knows that the Krylov cycle has ended;
    // returns at once while the cycle is still active
    if (gated && gc->st.active) return;
    __shared__ double sh[RT_T / 64];
    __shared__ double tl[3][TL_N];
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    // the centre columns start at i = 64 bx (cache-line aligned), the tile one column before
    const int i0 = blockIdx.x * 64 - 1, j0 = blockIdx.y * RT_J + 1;
    const int i = i0 + (t & 63) + 1;
    const int j = j0 + (t >> 6);
    const int k0 = 1 + blockIdx.z * kchunk, k1 = min(k0 + kchunk, nZ - 1);
    const double y = gc->st.kit >= 1 ? gc->coef[0] : 0.0;
    const bool live = i >= 1 && i <= nX - 2 && j <= nY - 2;
    const int sk = (int)g.sZ;                                    // N < 2^31 (checked at create)
    auto upd = [&](int cc) { return PC[cc] + (double)Z[cc] * y; };
    const int lc = ((t >> 6) + 1) * TL_X + (t & 63) + 1;         // this thread's node in the tile
    const int c0 = i + nX * j;
    double a = 0.0;
    double zm = live ? upd(c0 + sk * (k0 - 1)) : 0.0;
    int off[TL_PER];
    tile_offsets(off, t, i0, j0, g);
    double pre[TL_PER];
    tile_load(pre, off, k0, g, upd);
    tile_store(tl[0], pre, off, t);
    tile_load(pre, off, k0 + 1, g, upd);
    for (int k = k0; k < k1; k++) {
        const int q = (k - k0) % 3;
        double *cur = tl[q], *nxt = tl[(q + 1) % 3];
        const int c = c0 + sk * k, cl = live ? c : 0;
        // before the tile loads of plane k+2; unconditional (in-field addresses)
        const double b = pRhs[cl], Ab = g.Ab[k], At = g.At[k], Ap = g.Ap[k];
        tile_store(nxt, pre, off, t);                            // plane k+1
        tile_load(pre, off, min(k + 2, k1), g, upd);
        lds_barrier();
        if (live) {
            const double z0 = cur[lc], zp = nxt[lc];
            double rr = b - (g.Aw * cur[lc - 1] + g.Ae * cur[lc + 1] + g.As * cur[lc - TL_X] +
                                   g.An * cur[lc + TL_X] + Ab * zm + At * zp + Ap * z0);
            PCn[c] = z0;
            if (r) r[c] = (Tr)rr;
            a += rr * rr;
            zm = z0;
        }
    }
    double s = block_sum<RT_T>(a, sh, t);
    if (t == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
}

// ----------------------------------------------------------------------------
// a8: red-black SSOR preconditioner out = M^{-1} r, in two passes.
// The reference applies a lexicographic per-rank SSOR (dnssetbl.f90:509-532);
// here the same SSOR product (D/w+U)^{-1}((2-w)/w)D(D/w+L)^{-1} is taken in
// red-black order, whose forward red values and backward black values are
// pointwise and are recomputed where needed instead of stored:
//   pass 1 (black):  xb = -w (r_b + sum_nb A_nb xr_nb) / Ap,  xr = -w r_r / Ap
//   pass 2 (all):    black: yb = -w (((2-w)/w) Ap xb) / Ap
//                    red:   yr = -w (((2-w)/w) Ap xr + sum_nb A_nb yb_nb) / Ap
// with every product written in the order of the sweep formulas, so the
// result equals a literal red-black SSOR sweep bit for bit.
// ----------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T ssor_xr(const Geo &g, const T *r, long c, int k, T w)
{
    return -w * (r[c] + T(0.0)) * T(g.rAp[k]);
}
template <class T>
__device__ __forceinline__ T ssor_yb(const Geo &g, const T *xb, long c, int k, T w, T c2w)
{
    T s = c2w * T(g.Ap[k]) * xb[c];
    return -w * (s + T(0.0)) * T(g.rAp[k]);
}

template <class T>
__global__ __launch_bounds__(256) void k_ssor_fwd(Geo g, const T *__restrict__ r, T *__restrict__ xb,
                                                 const int *__restrict__ active)
{
    if (active && !*active) return;
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int j = blockIdx.y * 4 + (t >> 6) + 1;
    const int k = blockIdx.z + 1;
    if (j > nY - 2) return;
    const int ifirst = (((j + k) & 1) == 1) ? 1 : 2;
    const int i = ifirst + 2 * (blockIdx.x * 64 + (t & 63));
    if (i > nX - 2) return;
    const long c = (long)i + (long)nX * ((long)j + (long)nY * k), sk = g.sZ;
    const long cp = c + (long)(jp1_of(j, nY) - j) * nX, cm = c + (long)(jm1_of(j, nY) - j) * nX;
    const T w = T(g.omega);
    const T xw = (i > 1) ? ssor_xr(g, r, c - 1, k, w) : T(0);
    const T xe = (i < nX - 2) ? ssor_xr(g, r, c + 1, k, w) : T(0);
    const T xs = ssor_xr(g, r, cm, k, w);
    const T xn = ssor_xr(g, r, cp, k, w);
    const T xbo = (k > 1) ? ssor_xr(g, r, c - sk, k - 1, w) : T(0);
    const T xto = (k < nZ - 2) ? ssor_xr(g, r, c + sk, k + 1, w) : T(0);
    const T low = T(g.Aw) * xw + T(g.Ae) * xe + T(g.As) * xs + T(g.An) * xn + T(g.Ab[k]) * xbo + T(g.At[k]) * xto;
    xb[c] = -w * (r[c] + low) * T(g.rAp[k]);
}

template <class T>
__global__ __launch_bounds__(256) void k_ssor_bwd(Geo g, const T *__restrict__ r, const T *__restrict__ xb,
                                                 T *__restrict__ out, const int *__restrict__ active,
                                                 unsigned nbx, unsigned nby, unsigned nbz)
{
    if (active && !*active) return;
    const unsigned L = xcd_remap(blockIdx.x, nbx * nby * nbz);
    const unsigned bx = L % nbx, by = (L / nbx) % nby, bz = L / (nbx * nby);
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int i = bx * 64 + (t & 63);
    const int j = by * 4 + (t >> 6) + 1;
    const int k = bz + 1;
    if (i < 1 || i > nX - 2 || j > nY - 2) return;
    const long c = (long)i + (long)nX * ((long)j + (long)nY * k), sk = g.sZ;
    const T w = T(g.omega);
    const T c2w = (T(2) - w) / w;
    if (((i + j + k + 1) & 1) == 1) {              // black
        out[c] = ssor_yb(g, xb, c, k, w, c2w);
        return;
    }
    const long cp = c + (long)(jp1_of(j, nY) - j) * nX, cm = c + (long)(jm1_of(j, nY) - j) * nX;
    const T yw = (i > 1) ? ssor_yb(g, xb, c - 1, k, w, c2w) : T(0);
    const T ye = (i < nX - 2) ? ssor_yb(g, xb, c + 1, k, w, c2w) : T(0);
    const T ys = ssor_yb(g, xb, cm, k, w, c2w);
    const T yn = ssor_yb(g, xb, cp, k, w, c2w);
    const T ybo = (k > 1) ? ssor_yb(g, xb, c - sk, k - 1, w, c2w) : T(0);
    const T yto = (k < nZ - 2) ? ssor_yb(g, xb, c + sk, k + 1, w, c2w) : T(0);
    const T up = T(g.Aw) * yw + T(g.Ae) * ye + T(g.As) * ys + T(g.An) * yn + T(g.Ab[k]) * ybo + T(g.At[k]) * yto;
    const T xr = ssor_xr(g, r, c, k, w);
    const T s = c2w * T(g.Ap[k]) * xr;
    out[c] = -w * (s + up) * T(g.rAp[k]);
}

// Jacobi / identity preconditioners (out = r/Ap or r on the interior)
template <class T>
__global__ __launch_bounds__(256) void k_diag_pc(Geo g, const T *__restrict__ r, T *__restrict__ out,
                                                const int *__restrict__ active, int jacobi)
{
    if (active && !*active) return;
    const int nX = g.nX, nY = g.nY;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63);
    const int j = blockIdx.y * 4 + (t >> 6) + 1;
    const int k = blockIdx.z + 1;
    if (i < 1 || i > nX - 2 || j > nY - 2) return;
    const long c = (long)i + (long)nX * ((long)j + (long)nY * k);
    out[c] = jacobi ? r[c] * T(g.rAp[k]) : r[c];
}

// ----------------------------------------------------------------------------
// Boundary helpers
// ----------------------------------------------------------------------------
// zero the six faces of a scalar field (PCorr before the solve, dnssetbl.f90:451-460)
template <class T>
__global__ void k_zero_faces(Geo g, T *x)
{
    const long nX = g.nX, nY = g.nY, nZ = g.nZ;
    const long fz = nX * nY, fy = nX * nZ, fx = nY * nZ;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * fz) { long k = (t < fz) ? 0 : nZ - 1; long q = t % fz; x[q + fz * k] = T(0); return; }
    t -= 2 * fz;
    if (t < 2 * fy) { long j = (t < fy) ? 0 : nY - 1; long q = t % fy; long i = q % nX, k = q / nX; x[i + nX * (j + nY * k)] = T(0); return; }
    t -= 2 * fy;
    if (t < 2 * fx) { long i = (t < fx) ? 0 : nX - 1; long q = t % fx; long j = q % nY, k = q / nY; x[i + nX * (j + nY * k)] = T(0); }
}

// ----------------------------------------------------------------------------
// a10: PCorr extrapolation (dnssetbl.f90:697-708), then pressure update and
// projection (:711-726).  Three face kernels keep the reference's order:
// wall-normal faces, span copies, inflow/outflow faces.
// ----------------------------------------------------------------------------
__global__ void k_pc_zfaces(Geo g, double *PC)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
    if (i >= nX || j > nY - 2) return;
    const long s = g.sZ, b = (long)i + (long)nX * j;
    PC[b] = ((4.0 * PC[b + s]) - PC[b + 2 * s]) / 3.0;
    const long t = b + s * (nZ - 1);
    PC[t] = (5.0 * PC[t - s] + (-4.0 * PC[t - 2 * s] + PC[t - 3 * s])) / 2.0;
}
__global__ void k_pc_span(Geo g, double *PC)
{
    const int nX = g.nX, nY = g.nY;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= nX) return;
    const long b = (long)i + (long)g.sZ * k;
    PC[b] = PC[b + (long)nX * (nY - 2)];
    PC[b + (long)nX * (nY - 1)] = PC[b + (long)nX];
}
__global__ void k_pc_xfaces(Geo g, double *PC)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (j > nY - 2 || k >= nZ) return;
    const long b = (long)nX * ((long)j + (long)nY * k);
    PC[b] = (4.0 * PC[b + 1] - PC[b + 2]) / 3.0;
    const long e = b + nX - 1;
    PC[e] = (5.0 * PC[e - 1] + (-4.0 * PC[e - 2] + PC[e - 3])) / 2.0;
}

template <class T>
__global__ __launch_bounds__(256) void k_project(Geo g, const double *__restrict__ PC, T *__restrict__ Press,
                                                T *__restrict__ Ut)
{
    const int nX = g.nX, nY = g.nY;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63);
    const int j = blockIdx.y * 4 + (t >> 6) + 1;
    const int k = blockIdx.z + 1;
    if (i < 1 || i > nX - 2 || j > nY - 2) return;
    const long c = (long)i + (long)nX * ((long)j + (long)nY * k), sk = g.sZ, N = g.N;
    const long cp = c + (long)(jp1_of(j, nY) - j) * nX, cm = c + (long)(jm1_of(j, nY) - j) * nX;
    Press[c] = (T)((double)Press[c] + PC[c]);
    const double g1 = 0.5 * (PC[c + 1] - PC[c - 1]) * (g.invdx);
    const double g2 = 0.5 * (PC[cp] - PC[cm]) * (g.invdy);
    const double g3 = 0.5 * (PC[c + sk] - PC[c - sk]) * (g.invdXi * g.JacInv[k]);
    Ut[c] = (T)((double)Ut[c] - (g.dt * g1));
    Ut[c + N] = (T)((double)Ut[c + N] - (g.dt * g2));
    Ut[c + 2 * N] = (T)((double)Ut[c + 2 * N] - (g.dt * g3));
}

// ----------------------------------------------------------------------------
// a11: velocity and pressure boundary conditions (dnssetbl.f90:728-797)
// ----------------------------------------------------------------------------
// wall (k=1: no slip, w = wall_w control hook) and top (k=nZ: 2nd-order zero gradient)
template <class T>
__global__ void k_bc_vel_z(Geo g, T *Ut, const double *wall_w)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
    if (i >= nX || j > nY - 2) return;
    const long s = g.sZ, b = (long)i + (long)nX * j, top = b + s * (nZ - 1);
    for (int p = 0; p < 3; p++) {
        T *u = Ut + (long)p * g.N;
        u[b] = (p == 2 && wall_w) ? (T)wall_w[b] : T(0.0);
    }
    for (int p = 0; p < 3; p++) {
        T *u = Ut + (long)p * g.N;
        u[top] = (T(4.0) * u[top - s] - u[top - 2 * s]) / T(3.0);
    }
}
// outflow (i=nX) for every k, after the wall/top values are set
template <class T>
__global__ void k_bc_vel_out(Geo g, T *Ut)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y;
    if (j > nY - 2 || k >= nZ) return;
    const long e = (long)nX - 1 + (long)nX * ((long)j + (long)nY * k);
    for (int p = 0; p < 3; p++) {
        T *u = Ut + (long)p * g.N;
        u[e] = (T(4.0) * u[e - 1] - u[e - 2]) / T(3.0);
    }
}
// span periodic ghosts j=1 <- nY-1, j=nY <- 2 for ncomp components
template <class T>
__global__ void k_span_copy(Geo g, T *f, int ncomp)
{
    const int nX = g.nX, nY = g.nY;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= nX) return;
    for (int p = 0; p < ncomp; p++) {
        T *u = f + (long)p * g.N;
        const long b = (long)i + (long)g.sZ * k;
        u[b] = u[b + (long)nX * (nY - 2)];
        u[b + (long)nX * (nY - 1)] = u[b + (long)nX];
    }
}
// bottom (from wall-normal momentum) and top pressure, interior i and j
template <class T>
__global__ void k_bc_press_z(Geo g, const T *__restrict__ Ut, T *__restrict__ Press)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + 1;
    if (i > nX - 2 || j > nY - 2) return;
    const long s = g.sZ, b = (long)i + (long)nX * j;
    const T *W = Ut + 2 * g.N;
    const double dx = g.dx, dy = g.dy, df1 = g.df1, df2 = g.df2, Re = g.Re;
    const double w1 = W[b], w2 = W[b + s], w3 = W[b + 2 * s];
    const double rt11 = w1 * (w2 - w1) / df1;
    const double rt1 = ((double)W[b - 1] - 2.0 * w1 + (double)W[b + 1]) / (dx * dx);
    const double rt2 = ((double)W[b - nX] - 2.0 * w1 + (double)W[b + nX]) / (dy * dy);
    const double rt3 = (2.0 * df1 * w3 - 2.0 * df2 * w2 + 2.0 * (df2 - df1) * w1) / (df1 * (df2 * df2) - (df2 * (df1 * df1)));
    const double rt12 = -rt11 + (rt1 + rt2 + rt3) / Re;
    Press[b] = (T)((double)Press[b + s] - rt12 * df1);
    const long t = b + s * (nZ - 1);
    const double r1 = (11.0 * (double)W[t] - 18.0 * (double)W[t - s] + 9.0 * (double)W[t - 2 * s] - 2.0 * (double)W[t - 3 * s]) / (6.0 * g.dXi);
    const double r2 = (r1 * g.JacInv[nZ - 1]);
    Press[t] = (T)(r2 / Re);
}
// outflow and inflow pressure, k = 2..nZ-1 (inflow from streamwise momentum)
template <class T>
__global__ void k_bc_press_x(Geo g, const T *__restrict__ Ut, const T *__restrict__ Utold, T *__restrict__ Press)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (j > nY - 2 || k > nZ - 2) return;
    const long N = g.N, sk = g.sZ;
    const long c = (long)nX * ((long)j + (long)nY * k);
    const T *U = Ut, *V = Ut + N, *W = Ut + 2 * N;
    const double dx = g.dx, dy = g.dy, dXi = g.dXi, Re = g.Re, dt = g.dt;
    const long e = c + nX - 1;
    Press[e] = (T)((3.0 * (double)U[e] - 4.0 * (double)U[e - 1] + (double)U[e - 2]) / (2.0 * dx * Re));
    const double u1 = U[c];
    const double rt1 = u1 * (((double)U[c + 1] - u1) / (dx));
    const double rt2 = (double)V[c] * (((double)U[c + nX] - (double)U[c - nX]) / (2.0 * dy));
    const double rt3 = (double)W[c] * (((double)U[c + sk] - (double)U[c - sk]) / (2.0 * dXi));
    const double rt4 = (u1 + (double)U[c + 2] - 2 * (double)U[c + 1]) / (dx * dx);
    const double rt5 = ((double)U[c + nX] + (double)U[c - nX] - 2 * u1) / (dy * dy);
    double rt11, rt6;
    if (k <= 1) {
        rt11 = ((double)U[c + sk] - u1) / dXi;
        rt6 = ((double)U[c + 2 * sk] - 2.0 * (double)U[c + sk] + u1) / (dXi * dXi);
    } else if (k >= nZ - 2) {
        rt11 = (u1 - (double)U[c - sk]) / dXi;
        rt6 = ((double)U[c - 2 * sk] - 2.0 * (double)U[c - sk] + u1) / (dXi * dXi);
    } else {
        rt11 = ((double)U[c + sk] - (double)U[c - sk]) / (2.0 * dXi);
        rt6 = (-2 * u1 + (double)U[c + sk] + (double)U[c - sk]) / (dXi * dXi);
    }
    const double Ji = g.JacInv[k];
    rt6 = rt6 * (Ji * Ji) - g.Jac2[k] * rt11 * (Ji * Ji * Ji);
    const double rt12 = (u1 - (double)Utold[c]) / dt;
    Press[c] = (T)(((2 * dx * ((rt12 + rt1 + rt2 + rt3 * Ji) - ((rt4 + rt5 + rt6) / Re))) - (double)Press[c + 2] + (4 * (double)Press[c + 1])) / 3.0);
}

// Utold(1,:,:,1) = Ut(1,:,:,1): the part of Utold = Ut (dnssetbl.f90:257)
// that the step reads
template <class T>
__global__ void k_save_inflow(Geo g, const T *__restrict__ Ut, T *__restrict__ Utold)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)g.nY * g.nZ) return;
    const long c = e * g.nX;                                   // i = 0 of row (j, k)
    Utold[c] = Ut[c];
}
template <class T>
void launch_save_inflow(const Geo &g, const T *Ut, T *Utold, hipStream_t s)
{
    const long n = (long)g.nY * g.nZ;
    hipLaunchKernelGGL(k_save_inflow<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, Ut, Utold);
}
template void launch_save_inflow<float>(const Geo &, const float *, float *, hipStream_t);
template void launch_save_inflow<double>(const Geo &, const double *, double *, hipStream_t);

// ----------------------------------------------------------------------------
// a2: recycling/rescaling inflow plane (dnssetbl.f90:264-303), k = 2..nZ-1
// ----------------------------------------------------------------------------
template <class T>
__global__ void k_recycle(Geo g, T *Ut, int i_rec0, double eps, const double *__restrict__ lamda_p,
                          const double *__restrict__ avg, const double *__restrict__ Wt,
                          const int *__restrict__ newk, const double *__restrict__ Ume,
                          const double *__restrict__ uf)
{
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (j > nY - 2 || k > nZ - 2) return;
    const long M = (long)nX * nZ, N = g.N;
    const double lamda = *lamda_p;
    const double *usum2 = avg, *vsum2 = avg + M, *wsum2 = avg + 2 * M;
    const long c = (long)nX * ((long)j + (long)nY * k);
    if (usum2[i_rec0 + (long)nX * k] < (0.99 * g.Ue)) {
        const int nk = newk[k] - 1;
        const long ak = i_rec0 + (long)nX * nk;
        const double Um = eps * (usum2[ak]) + (1 - eps) * Ume[k];
        const double Vm = lamda * vsum2[ak];
        const double Wm = wsum2[ak];
        const double W1 = Wt[(long)nX * k];
        const long o = (long)j + (long)nY * nk, P = (long)nY * nZ;
        const double f1 = lamda * uf[o], f2 = lamda * uf[o + P], f3 = lamda * uf[o + 2 * P];
        Ut[c] = (T)((Um + f1) * (1.0 - W1) + (Um + f1) * W1);
        Ut[c + N] = (T)((Vm + f2) * (1.0 - W1) + (Vm + f2) * W1);
        Ut[c + 2 * N] = (T)((Wm + f3) * (1.0 - W1) + (Wm + f3) * W1);
    } else {
        const long a = i_rec0 + (long)nX * k;
        Ut[c] = (T)usum2[a];
        Ut[c + N] = (T)vsum2[a];
        Ut[c + 2 * N] = (T)wsum2[a];
    }
}

// ============================================================================
// host launchers
// ============================================================================
// column tiles start at i = 0 (cache-line aligned row segments) and cover i <= nX-2
static inline dim3 tiles3(const Geo &g) { return dim3((g.nX - 1 + 63) / 64, (g.nY - 2 + 3) / 4, g.nZ - 2); }
static inline dim3 pairs3(const Geo &g) { return dim3(((g.nX - 1) / 2 + 63) / 64, (g.nY - 2 + 3) / 4, g.nZ - 2); }
static inline dim3 rgrid(const Geo &g, int *kchunk)
{
    RTiles r = rtiles_of(g.nX, g.nY, g.nZ);
    *kchunk = r.kchunk;
    return dim3(r.gx, r.gy, r.kb);
}
unsigned resid_blocks(const Geo &g) { RTiles r = rtiles_of(g.nX, g.nY, g.nZ); return (unsigned)(r.gx * r.gy * r.kb); }

// planes per block of the FP32 tile kernel: the column is cut into the
// number of chunks that minimises (waves of blocks over the resident slots) x
// (planes per chunk + 5, the queue prologue and halo planes); the slots are
// MI355X's 256 CUs x 2 resident blocks of 512 threads at 128 VGPRs.  A
// chunk's buffer range stays below 2^31 bytes (stores of lanes outside the
// field go to offset 2^31).  0: no chunk fits, the register-queue kernel runs.
static int rhs_chunk_f32(const Geo &g)
{
    const long tiles = (long)((g.nX - 1 + 63) / 64) * ((g.nY - 2 + RHS_TJ - 1) / RHS_TJ), slots = 512;
    const int nz = g.nZ - 2;
    int best = 0;
    long bestc = -1;
    for (int nbz = 1; nbz <= nz; nbz++) {
        const int kc = (nz + nbz - 1) / nbz, nb = (nz + kc - 1) / kc;
        if (kc > RHS_KCHUNK_F32 || (long)(kc + MR_R + 5) * g.sZ * 4 >= (1L << 31)) continue;
        const long cost = (tiles * nb + slots - 1) / slots * (kc + 5);
        if (bestc < 0 || cost < bestc) { bestc = cost; best = kc; }
        if (kc <= 8) break;
    }
    return best;
}
static bool rhs_tile_kernel(const Geo &g, bool single)
{
    return single && g.scheme == TBL_SCHEME_SETBL && rhs_chunk_f32(g) > 0;
}
static int rhs_chunk(const Geo &g, bool single)
{
    return rhs_tile_kernel(g, single) ? rhs_chunk_f32(g) : RHS_KCHUNK;
}
// column tiles start at i = 0 (cache-line aligned) and cover i <= nX-2
static unsigned rhs_nbx(const Geo &g, bool)
{
    return (unsigned)((g.nX - 1 + 63) / 64);
}
unsigned rhs_blocks(const Geo &g, bool single)
{
    const int kc = rhs_chunk(g, single);
    return rhs_nbx(g, single) * (unsigned)((g.nY - 2 + RHS_TJ - 1) / RHS_TJ) *
           (unsigned)((g.nZ - 2 + kc - 1) / kc);
}
template <class T>
void launch_momentum_rhs(const Geo &g, const T *Ut, const T *P, T *Q, hipStream_t s, const unsigned char *brin,
                         unsigned char *brout, double *hpart)
{
    const int kchunk = rhs_chunk(g, sizeof(T) == 4);
    unsigned nbx = rhs_nbx(g, sizeof(T) == 4), nby = (g.nY - 2 + RHS_TJ - 1) / RHS_TJ, nbz = (g.nZ - 2 + kchunk - 1) / kchunk;
    const dim3 grid(nbx * nby * nbz), blk(64 * RHS_TJ);
    if constexpr (sizeof(T) == 4) {
        if (rhs_tile_kernel(g, true)) {
            const float *Uf = (const float *)Ut, *Pf = (const float *)P;
            float *Qf = (float *)Q;
            if (brin || brout)
                hipLaunchKernelGGL(k_momentum_rhs_f32<true>, grid, blk, 0, s, g, Uf, Pf, Qf, kchunk, nbx, nby, nbz,
                                   brin, brout, hpart);
            else
                hipLaunchKernelGGL(k_momentum_rhs_f32<false>, grid, blk, 0, s, g, Uf, Pf, Qf, kchunk, nbx, nby, nbz,
                                   brin, brout, hpart);
            return;
        }
    }
    if (g.scheme == TBL_SCHEME_DNSTBL)
        hipLaunchKernelGGL((k_momentum_rhs<T, TBL_SCHEME_DNSTBL, false>), grid, blk, 0, s, g, Ut, P, Q, kchunk, nbx, nby,
                           nbz, brin, brout, hpart);
    else if (brin || brout)
        hipLaunchKernelGGL((k_momentum_rhs<T, TBL_SCHEME_SETBL, true>), grid, blk, 0, s, g, Ut, P, Q, kchunk, nbx, nby,
                           nbz, brin, brout, hpart);
    else
        hipLaunchKernelGGL((k_momentum_rhs<T, TBL_SCHEME_SETBL, false>), grid, blk, 0, s, g, Ut, P, Q, kchunk, nbx, nby,
                           nbz, brin, brout, hpart);
}

template <class T>
void launch_helm_resid(const Geo &g, const T *Q, const T *Ut, const void *ctrl, double *part, int first, hipStream_t s)
{
    int kc;
    dim3 grid = rgrid(g, &kc);
    hipLaunchKernelGGL(k_helm_resid<T>, grid, dim3(RT_T), 0, s, g, Q, Ut, (const GsCtrl *)ctrl, part, first, kc);
}
void launch_helm_fin(const Geo &g, void *ctrl, const double *part, int it, double errorstop, double floor_coef,
                     hipStream_t s, int nb)
{
    hipLaunchKernelGGL(k_helm_fin, dim3(1), dim3(1024), 0, s, (GsCtrl *)ctrl, part, nb > 0 ? nb : (int)resid_blocks(g),
                       it, errorstop, floor_coef);
}
template <class T>
void launch_helm_sweep(const Geo &g, const T *Q, T *Ut, const void *ctrl, int colour, hipStream_t s)
{
    hipLaunchKernelGGL(k_helm_sweep<T>, pairs3(g), dim3(256), 0, s, g, Q, Ut, (const GsCtrl *)ctrl, colour);
}
template <class T>
void launch_helm_zline(const Geo &g, const T *Q, T *Ut, const T *zf, const void *ctrl, int colour, hipStream_t s)
{
    // partition of the interior planes (the zline_chunks rule): LC = 16 for
    // nZ-2 <= 256, else 32; nZ-2 <= 512 is checked at create time
    const int nz = g.nZ - 2, LC = nz <= 256 ? 16 : 32, C = (nz + LC - 1) / LC;
    const unsigned nbx = (unsigned)(((g.nX - 1) / 2 + 63) / 64), nblk = nbx * (unsigned)(g.nY - 2);
    const unsigned pw = nbx < 2 ? nbx : 2;
    if (LC == 16)
        hipLaunchKernelGGL((k_helm_zline<T, 16>), dim3(nblk), dim3(64 * C), 0, s, g, Q, Ut, zf, (const GsCtrl *)ctrl,
                           colour, C, nbx, nblk, pw);
    else
        hipLaunchKernelGGL((k_helm_zline<T, 32>), dim3(nblk), dim3(64 * C), 0, s, g, Q, Ut, zf, (const GsCtrl *)ctrl,
                           colour, C, nbx, nblk, pw);
}
template <class T>
void launch_helm_zebra(const Geo &g, const T *Q, const T *cur, const T *ysrc, T *dst, const T *zf, const void *ctrl,
                       int colour, hipStream_t s)
{
    const int nz = g.nZ - 2, LC = nz <= 256 ? 16 : 32, C = (nz + LC - 1) / LC;
    // column tiles start at i = 0 (cache-line aligned) and cover i <= nX-2
    const unsigned nbx = (unsigned)((g.nX - 1 + 63) / 64), nblk = nbx * (unsigned)((g.nY - 2) / 2);
#ifndef ZB_PW
#define ZB_PW 2
#endif
    const unsigned pw = nbx < ZB_PW ? nbx : ZB_PW;
    // prefetch depth (planes of operands in flight beyond the one being
    // solved): one, which keeps FP32 16-plane chunks at 80 VGPRs (two
    // workgroups per CU; depths 3 and 5 at one workgroup per CU were slower)
    // and FP64 free of spills
    constexpr int ZPF16 = 1;
    if (LC == 16)
        hipLaunchKernelGGL((k_helm_zebra<T, 16, ZPF16>), dim3(nblk), dim3(64 * C), 0, s, g, Q, cur, ysrc, dst, zf,
                           (const GsCtrl *)ctrl, colour, C, nbx, nblk, pw);
    else
        hipLaunchKernelGGL((k_helm_zebra<T, 32, 1>), dim3(nblk), dim3(64 * C), 0, s, g, Q, cur, ysrc, dst, zf,
                           (const GsCtrl *)ctrl, colour, C, nbx, nblk, pw);
}
template <class T>
void launch_copy_faces3(const Geo &g, const T *src, T *dst, hipStream_t s)
{
    const long per = 2 * ((long)g.nX * g.nY + (long)g.nX * g.nZ + (long)g.nY * g.nZ);
    hipLaunchKernelGGL(k_copy_faces3<T>, dim3((unsigned)((3 * per + 255) / 256)), dim3(256), 0, s, g, src, dst);
}
size_t gs_ctrl_bytes() { return sizeof(GsCtrl); }
void gs_ctrl_init(void *host, int nactive_mask)
{
    GsCtrl *c = (GsCtrl *)host;
    for (int p = 0; p < 3; p++) {
        c->sumnor[p] = 0; c->res[p] = 0; c->qnorm[p] = 0;
        c->active[p] = (nactive_mask >> p) & 1; c->sweeps[p] = 0;
    }
}
void gs_ctrl_read(const void *host, int *active, int *sweeps, double *res)
{
    const GsCtrl *c = (const GsCtrl *)host;
    for (int p = 0; p < 3; p++) { active[p] = c->active[p]; sweeps[p] = c->sweeps[p]; res[p] = c->res[p]; }
}

template <class T>
void launch_ppe_rhs(const Geo &g, const T *Ut, double *pRhs, hipStream_t s, const void *gate)
{
    hipLaunchKernelGGL(k_ppe_rhs<T>, tiles3(g), dim3(256), 0, s, g, Ut, pRhs, (const GsCtrl *)gate);
}

#ifndef POISSON_KC
#define POISSON_KC 16   // wall-normal planes per thread of the Poisson apply
#endif
unsigned poisson_blocks(const Geo &g)
{
    return (unsigned)(((g.nX - 1 + 63) / 64) * ((g.nY - 2 + 3) / 4) * ((g.nZ - 2 + POISSON_KC - 1) / POISSON_KC));
}
template <class T>
unsigned launch_poisson(const Geo &g, const T *x, T *out, const int *active, const double *scale, hipStream_t s,
                        const T *v0, double *part)
{
    const int kchunk = POISSON_KC;
    unsigned nbx = (g.nX - 1 + 63) / 64, nby = (g.nY - 2 + 3) / 4, nbz = (g.nZ - 2 + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(k_poisson<T>, dim3(nbx * nby * nbz), dim3(256), 0, s, g, x, out, active, scale, v0, part, kchunk,
                       nbx, nby, nbz);
    return nbx * nby * nbz;
}

template <class Tr>
void launch_resid(const Geo &g, const double *pRhs, const double *PC, Tr *r, double *part, hipStream_t s,
                  double *bpart)
{
    int kc;
    dim3 grid = rgrid(g, &kc);
    hipLaunchKernelGGL(k_resid<Tr>, grid, dim3(RT_T), 0, s, g, pRhs, PC, r, part, bpart, kc);
}

template <class Tr>
void launch_resid_upd(const Geo &g, const double *pRhs, const double *PC, const Tr *Z, const GmCtrl *gc, double *PCn,
                      Tr *r, double *part, hipStream_t s, int gated)
{
    int kc;
    dim3 grid = rgrid(g, &kc);
    hipLaunchKernelGGL(k_resid_upd<Tr>, grid, dim3(RT_T), 0, s, g, pRhs, PC, Z, gc, PCn, r, part, kc, gated);
}

// ----------------------------------------------------------------------------
// The reference's own orderings of the two iterative solves, inside each
// wall-normal slab of its MPI decomposition (slab bounds dnssetbl.f90:76-92,
// in g.slab as 0-based first/last interior planes):
//  * Gauss-Seidel of the implicit diffusion (:395-430): lexicographic (i
//    fastest, then j, then k) inside a slab; the ghost planes below and above
//    the slab hold the neighbouring slabs' values from before the sweep (the
//    reference exchanges them after every sweep), read here from a snapshot
//    taken before each sweep;
//  * block SSOR of the pressure GMRES (:509-532, 586-610): lexicographic
//    forward and backward sweeps inside a slab, with zero values across the
//    slab faces and the span faces (no periodic coupling).
// A lexicographic sweep is a chain: point (i,j,k) reads the new values of
// (i-1,j,k), (i,j-1,k) and (i,j,k-1) and the old values of the other
// neighbours; with the periodic span wrap, row j = 1 reads the old value of
// row nY-2 and row nY-2 the new value of row 1.  Both sweeps run as a
// wavefront over the lines: in step t line (j, k) updates its block
// t - j - k.  Every new neighbour of a point then lies in the same block
// before it or in a block updated in an earlier step, and every old
// neighbour in a block of a later step (the wrap's new row 1 ny - 1 steps
// earlier), so the values read are exactly the lexicographic loop's and the
// results are the reference's, bit for bit.
// One workgroup per slab (and velocity component).  Thread (jj, kc) owns span
// row jj and chunk kc of KL consecutive wall-normal lines of the slab (rows
// padded to whole wavefronts); one barrier per step.
//  * Gauss-Seidel: blocks of SL_B points (SL_B/2 in FP64), neighbour values
//    from memory (a value another thread wrote in an earlier step is visible
//    after that step's barrier); ceil((nX-2)/SL_B) + (nY-3) + planes steps
//    per sweep.
//  * block SSOR: one point per block, the i-1 and k-1 neighbours in
//    registers (LDS across chunks), the j-1 neighbour from the previous lane
//    by wave shuffle (LDS across wavefronts); (nX-2) + (nY-3) + planes steps
//    per sweep.
// ----------------------------------------------------------------------------
#define SL_MAXT 1024
#define SL_MAXW (SL_MAXT / 64)
#define SL_B 8
// threads of a slab workgroup with KL lines per thread: up to 1024 for
// KL <= 8, 512 above
#define SL_BOUND(KL) ((KL) <= 8 ? SL_MAXT : SL_MAXT / 2)
// launch shape of a slab kernel: KL lines per thread, nkc chunks, rows
// padded to nyp; 0 if the slabs need more threads than one workgroup has
static int slab_shape(const Geo &g, int nkmax, int &KL, int &nkc, int &nyp)
{
    nyp = (g.nY - 2 + 63) / 64 * 64;
    for (KL = 4; KL <= 32; KL *= 2) {
        nkc = (nkmax + KL - 1) / KL;
        if (nyp * nkc <= SL_BOUND(KL)) return nyp * nkc;
    }
    return 0;
}

// snapshot of the ghost planes of every slab before a sweep; layout
// [component][slab][below, above] x (nX nY)
template <class T>
__global__ void k_slab_halo(Geo g, const T *__restrict__ Ut, T *__restrict__ halo, const GsCtrl *__restrict__ ctrl)
{
    const int p = blockIdx.z, hs = blockIdx.y, r = hs >> 1;
    if (!ctrl->active[p]) return;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.sZ) return;
    const int k = (hs & 1) ? g.slab[2 * r + 1] + 1 : g.slab[2 * r] - 1;
    halo[((long)p * 2 * g.nslab + hs) * g.sZ + e] = Ut[(long)p * g.N + (long)k * g.sZ + e];
}

template <class T, int KL>
__global__ __launch_bounds__(SL_BOUND(KL)) void k_helm_lex(Geo g, const T *__restrict__ Qall, T *Uall,
                                                           const T *__restrict__ halo, const GsCtrl *__restrict__ ctrl,
                                                           int nkc, int nyp)
{
    const int r = blockIdx.x, p = blockIdx.y;
    if (!ctrl->active[p]) return;
    const int tid = threadIdx.x;
    const int jj = tid % nyp, kc = tid / nyp;
    const int nX = g.nX, nY = g.nY, nx = nX - 2, ny = nY - 2;
    const int kb = g.slab[2 * r], nk = g.slab[2 * r + 1] - kb + 1;
    const long sZ = g.sZ;
    const bool live = jj < ny;
    T *u = Uall + (long)p * g.N;
    const T *q = Qall + (long)p * g.N;
    const T *hb = halo + (long)(p * 2 * g.nslab + 2 * r) * sZ, *ht = hb + sZ;
    const int j = jj + 1;
    // span neighbours: row j+1 (old; for the last row the wrap's row 1, new)
    // and row j-1 (new, from the previous step; for the first row the
    // wrap's row nY-2, old)
    const long dn = (long)(jp1_of(live ? j : 1, nY) - j) * nX, ds = (long)(jm1_of(live ? j : 1, nY) - j) * nX;
    const T dt = T(g.dt), Re = T(g.Re), Ae = T(g.Ae), Aw = T(g.Aw), An = T(g.An), As = T(g.As);
    constexpr int LB = sizeof(T) == 8 ? SL_B / 2 : SL_B;     // points per block (FP64: registers)
    const int nbk = (nx + LB - 1) / LB;
    const int steps = nbk + ny + nk - 2;
    for (int t = 0; t < steps; t++) {
#pragma unroll 1
        for (int l = 0; l < KL; l++) {
            const int kk = kc * KL + l, bi = t - jj - kk;
            if (live && kk < nk && bi >= 0 && bi < nbk) {
                const int k = kb + kk, i0 = bi * LB;
                const long c0 = (long)(i0 + 1) + (long)nX * ((long)j + (long)nY * k), pl0 = c0 - (long)k * sZ;
                const T Abk = T(g.Ab[k]), Atk = T(g.At[k]), Apk = T(g.Ap[k]);
                // i-1 of the block's first point: the face (first block) or the
                // line's previous block, written by this thread one step earlier
                T uw = u[c0 - 1];
                // every operand of the block is read before its first store:
                // none is written inside the block except i+1, which must be
                // read old anyway (points past the line end read in-field
                // values that are not used)
                T qv[LB], ue[LB], un[LB], us[LB], ub[LB], ut[LB];
#pragma unroll
                for (int b = 0; b < LB; b++) {
                    const long c = c0 + b, pl = pl0 + b;
                    qv[b] = q[c];
                    ue[b] = u[c + 1];
                    un[b] = u[c + dn];
                    us[b] = u[c + ds];
                    ub[b] = kk == 0 ? hb[pl] : u[c - sZ];
                    ut[b] = kk == nk - 1 ? ht[pl] : u[c + sZ];
                }
#pragma unroll
                for (int b = 0; b < LB; b++) {
                    if (i0 + b < nx) {
                        const T rt1 = Ae * ue[b] + Aw * uw + An * un[b] + As * us[b];
                        const T rt2 = Atk * ut[b] + Abk * ub[b];
                        const T v = (qv[b] + dt / Re * (rt1 + rt2)) / (T(1.0) - dt / Re * Apk);
                        u[c0 + b] = v;
                        uw = v;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// block SSOR, one workgroup per slab: forward sweep (the stored values are
// already scaled by ((2 - w)/w) Ap), then the backward sweep in the mirrored
// order; the forward values a neighbour reads are the unscaled ones
template <class T, int KL>
__global__ __launch_bounds__(SL_BOUND(KL)) void k_ssor_block(Geo g, const T *__restrict__ rin, T *out,
                                                        const int *__restrict__ active, int nkc, int nyp)
{
    if (active && !*active) return;
    const int r = blockIdx.x;
    __shared__ T sW[2][SL_MAXW][KL];
    __shared__ T sK[2][SL_MAXT];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int jj = tid % nyp, kc = tid / nyp;
    const int nX = g.nX, nY = g.nY, nx = nX - 2, ny = nY - 2;
    const int kb = g.slab[2 * r], nk = g.slab[2 * r + 1] - kb + 1;
    const bool live = jj < ny;
    const int j = jj + 1, jc = live ? j : 1;
    const T om = T(g.omega);
    const int steps = nx + ny + nk - 2;
    // per line: its plane's coefficients, and the address of its first point
    // (the point of a step follows from the step index)
    T Abk[KL], Atk[KL], Apk[KL], rApk[KL];
    long c0l[KL];
#pragma unroll
    for (int l = 0; l < KL; l++) {
        const int k = kb + min(kc * KL + l, nk - 1);
        Abk[l] = T(g.Ab[k]); Atk[l] = T(g.At[k]); Apk[l] = T(g.Ap[k]); rApk[l] = T(g.rAp[k]);
        c0l[l] = 1 + (long)nX * ((long)jc + (long)nY * k);
    }
    // A line's operand of step t (rin forward, the forward result backward)
    // is loaded SD steps ahead into an SD + 1 slot register ring; the step
    // loop is unrolled by the slot count, so a slot keeps its registers.
    // Loads are unconditional at clamped points (the values of steps a line
    // does not update are never used).  Backward, a line's future points are
    // written only by the line itself, so the values loaded ahead are the
    // forward results.
    constexpr int SD = 2, NS = SD + 1;
    T rv[NS][KL];
    T xprev[KL];
    auto ld_f = [&](int sl, int t) {
#pragma unroll
        for (int l = 0; l < KL; l++) {
            const int ii = min(max(t - jj - (kc * KL + l), 0), nx - 1);
            rv[sl][l] = rin[c0l[l] + ii];
        }
    };
#pragma unroll
    for (int l = 0; l < KL; l++) xprev[l] = T(0);
#pragma unroll
    for (int q = 0; q < SD; q++) ld_f(q, q);
    auto fwd = [&](auto S, int t) {
        constexpr int sl = decltype(S)::value;
        ld_f((sl + SD) % NS, t + SD);
        const int b = t & 1;
        const T kin = kc > 0 ? sK[b ^ 1][tid - nyp] : T(0);
#pragma unroll
        for (int l = KL - 1; l >= 0; l--) {
            T sj = __shfl_up(xprev[l], 1, 64);
            if (lane == 0 && jj > 0) sj = sW[b ^ 1][w - 1][l];
            const int kk = kc * KL + l, ii = t - jj - kk;
            if (live && kk < nk && ii >= 0 && ii < nx) {
                const long c = c0l[l] + ii;
                const T rw = ii == 0 ? T(0) : xprev[l];
                const T rs = jj == 0 ? T(0) : sj;
                const T rb = kk == 0 ? T(0) : (l == 0 ? kin : xprev[l - 1]);
                const T v = -om * (rv[sl][l] + T(g.Aw) * rw + T(g.As) * rs + Abk[l] * rb) * rApk[l];
                out[c] = (T(2) - om) / om * Apk[l] * v;
                xprev[l] = v;
            }
        }
        if (lane == 63) {
#pragma unroll
            for (int l = 0; l < KL; l++) sW[b][w][l] = xprev[l];
        }
        sK[b][tid] = xprev[KL - 1];
        __syncthreads();
    };
    int t = 0;
    for (; t + NS <= steps; t += NS) ring_steps<0, NS>(fwd, t);
    ring_tail<0, NS>(fwd, t, steps);
    // backward: mirrored coordinates (nx-1-ii, ny-1-jj, nk-1-kk); the new
    // neighbours are i+1 (itself), j+1 (next lane) and k+1 (next line)
#pragma unroll
    for (int l = 0; l < KL; l++) xprev[l] = T(0);
    const int jr = ny - 1 - jj;
    auto ld_b = [&](int sl, int t) {
#pragma unroll
        for (int l = 0; l < KL; l++) {
            const int ir = min(max(t - jr - (nk - 1 - (kc * KL + l)), 0), nx - 1);
            rv[sl][l] = out[c0l[l] + (nx - 1 - ir)];
        }
    };
#pragma unroll
    for (int q = 0; q < SD; q++) ld_b(q, q);
    auto bwd = [&](auto S, int t) {
        constexpr int sl = decltype(S)::value;
        ld_b((sl + SD) % NS, t + SD);
        const int b = t & 1;
        const T kin = kc < nkc - 1 ? sK[b ^ 1][tid + nyp] : T(0);
        // lowest line first: line l reads line l+1's value of step t-1
#pragma unroll
        for (int l = 0; l < KL; l++) {
            T sj = __shfl_down(xprev[l], 1, 64);              // row jj+1, step t-1
            if (lane == 63 && jj < ny - 1) sj = sW[b ^ 1][w + 1][l];
            const int kk = kc * KL + l, ir = t - jr - (nk - 1 - kk), ii = nx - 1 - ir;
            if (live && kk < nk && ir >= 0 && ir < nx) {
                const long c = c0l[l] + ii;
                const T re = ii == nx - 1 ? T(0) : xprev[l];
                const T rn = jj == ny - 1 ? T(0) : sj;
                const T rt = kk == nk - 1 ? T(0) : (l == KL - 1 ? kin : xprev[l + 1]);
                const T v = -om * (rv[sl][l] + T(g.Ae) * re + T(g.An) * rn + Atk[l] * rt) * rApk[l];
                out[c] = v;
                xprev[l] = v;
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int l = 0; l < KL; l++) sW[b][w][l] = xprev[l];
        }
        sK[b][tid] = xprev[0];
        __syncthreads();
    };
    t = 0;
    for (; t + NS <= steps; t += NS) ring_steps<0, NS>(bwd, t);
    ring_tail<0, NS>(bwd, t, steps);
}

int slab_threads(const Geo &g, int nkmax)
{
    int KL, nkc, nyp;
    return slab_shape(g, nkmax, KL, nkc, nyp);
}
template <class T>
void launch_helm_lex(const Geo &g, const T *Q, T *Ut, T *halo, const void *ctrl, int nkmax, hipStream_t s)
{
    int KL, nkc, nyp;
    const int nt = slab_shape(g, nkmax, KL, nkc, nyp);
    const GsCtrl *gc = (const GsCtrl *)ctrl;
    hipLaunchKernelGGL(k_slab_halo<T>, dim3((unsigned)((g.sZ + 255) / 256), 2 * g.nslab, 3), dim3(256), 0, s, g,
                       (const T *)Ut, halo, gc);
    const dim3 grid(g.nslab, 3), blk(nt);
    switch (KL) {
    case 4: hipLaunchKernelGGL((k_helm_lex<T, 4>), grid, blk, 0, s, g, Q, Ut, (const T *)halo, gc, nkc, nyp); break;
    case 8: hipLaunchKernelGGL((k_helm_lex<T, 8>), grid, blk, 0, s, g, Q, Ut, (const T *)halo, gc, nkc, nyp); break;
    case 16: hipLaunchKernelGGL((k_helm_lex<T, 16>), grid, blk, 0, s, g, Q, Ut, (const T *)halo, gc, nkc, nyp); break;
    default: hipLaunchKernelGGL((k_helm_lex<T, 32>), grid, blk, 0, s, g, Q, Ut, (const T *)halo, gc, nkc, nyp); break;
    }
}
template <class T>
static void launch_ssor_block(const Geo &g, const T *r, T *out, const int *active, hipStream_t s)
{
    int KL, nkc, nyp;
    const int nt = slab_shape(g, g.nkmax, KL, nkc, nyp);
    const dim3 grid(g.nslab), blk(nt);
    switch (KL) {
    case 4: hipLaunchKernelGGL((k_ssor_block<T, 4>), grid, blk, 0, s, g, r, out, active, nkc, nyp); break;
    case 8: hipLaunchKernelGGL((k_ssor_block<T, 8>), grid, blk, 0, s, g, r, out, active, nkc, nyp); break;
    case 16: hipLaunchKernelGGL((k_ssor_block<T, 16>), grid, blk, 0, s, g, r, out, active, nkc, nyp); break;
    default: hipLaunchKernelGGL((k_ssor_block<T, 32>), grid, blk, 0, s, g, r, out, active, nkc, nyp); break;
    }
}

template <class T>
void launch_precond(const Geo &g, int pc, const FastPC *fast, const T *r, T *scratch, T *out, const int *active,
                    hipStream_t s, int part)
{
    if (pc == PC_FAST) {
        run_fastpc<T>(fast, g, r, out, active, s, part);
    } else if (pc == PC_SSOR_BLOCK) {
        launch_ssor_block<T>(g, r, out, active, s);
    } else if (pc == PC_SSOR_RB) {
        hipLaunchKernelGGL(k_ssor_fwd<T>, pairs3(g), dim3(256), 0, s, g, r, scratch, active);
        unsigned nbx = (g.nX - 1 + 63) / 64, nby = (g.nY - 2 + 3) / 4, nbz = g.nZ - 2;
        hipLaunchKernelGGL(k_ssor_bwd<T>, dim3(nbx * nby * nbz), dim3(256), 0, s, g, r, (const T *)scratch, out, active, nbx, nby, nbz);
    } else {
        hipLaunchKernelGGL(k_diag_pc<T>, tiles3(g), dim3(256), 0, s, g, r, out, active, pc == PC_JACOBI ? 1 : 0);
    }
}

template <class T>
void launch_zero_faces(const Geo &g, T *x, hipStream_t s)
{
    long tot = 2L * ((long)g.nX * g.nY + (long)g.nX * g.nZ + (long)g.nY * g.nZ);
    hipLaunchKernelGGL(k_zero_faces<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, x);
}
template void launch_zero_faces<double>(const Geo &, double *, hipStream_t);
template void launch_zero_faces<float>(const Geo &, float *, hipStream_t);

template <class T>
void launch_project(const Geo &g, double *PC, T *Press, T *Ut, hipStream_t s)
{
    hipLaunchKernelGGL(k_pc_zfaces, dim3((g.nX + 255) / 256, g.nY - 2), dim3(256), 0, s, g, PC);
    hipLaunchKernelGGL(k_pc_span, dim3((g.nX + 255) / 256, g.nZ), dim3(256), 0, s, g, PC);
    hipLaunchKernelGGL(k_pc_xfaces, dim3((g.nY - 2 + 255) / 256, g.nZ), dim3(256), 0, s, g, PC);
    hipLaunchKernelGGL(k_project<T>, tiles3(g), dim3(256), 0, s, g, (const double *)PC, Press, Ut);
}

template <class T>
void launch_bcs(const Geo &g, T *Ut, const T *Utold, T *Press, const double *wall_w, hipStream_t s)
{
    hipLaunchKernelGGL(k_bc_vel_z<T>, dim3((g.nX + 255) / 256, g.nY - 2), dim3(256), 0, s, g, Ut, wall_w);
    hipLaunchKernelGGL(k_bc_vel_out<T>, dim3((g.nY - 2 + 255) / 256, g.nZ), dim3(256), 0, s, g, Ut);
    hipLaunchKernelGGL(k_span_copy<T>, dim3((g.nX + 255) / 256, g.nZ), dim3(256), 0, s, g, Ut, 3);
    hipLaunchKernelGGL(k_bc_press_z<T>, dim3((g.nX - 2 + 255) / 256, g.nY - 2), dim3(256), 0, s, g, (const T *)Ut, Press);
    hipLaunchKernelGGL(k_bc_press_x<T>, dim3((g.nY - 2 + 255) / 256, g.nZ - 2), dim3(256), 0, s, g, (const T *)Ut, Utold, Press);
    hipLaunchKernelGGL(k_span_copy<T>, dim3((g.nX + 255) / 256, g.nZ), dim3(256), 0, s, g, Press, 1);
}

template <class T>
void launch_recycle(const Geo &g, T *Ut, int i_rec0, double eps, const double *lamda, const double *avg,
                    const double *Wt, const int *newk, const double *Ume, const double *uf, hipStream_t s)
{
    hipLaunchKernelGGL(k_recycle<T>, dim3((g.nY - 2 + 255) / 256, g.nZ - 2), dim3(256), 0, s, g, Ut, i_rec0,
                       eps, lamda, avg, Wt, newk, Ume, uf);
}

#define TBL_INST(T)                                                                                        \
    template void launch_momentum_rhs<T>(const Geo &, const T *, const T *, T *, hipStream_t, const unsigned char *, \
                                         unsigned char *, double *);                                        \
    template void launch_helm_resid<T>(const Geo &, const T *, const T *, const void *, double *, int, hipStream_t); \
    template void launch_helm_sweep<T>(const Geo &, const T *, T *, const void *, int, hipStream_t);      \
    template void launch_helm_zline<T>(const Geo &, const T *, T *, const T *, const void *, int, hipStream_t); \
    template void launch_helm_zebra<T>(const Geo &, const T *, const T *, const T *, T *, const T *, const void *, int, \
                                       hipStream_t);                                                                    \
    template void launch_copy_faces3<T>(const Geo &, const T *, T *, hipStream_t);                                     \
    template void launch_helm_lex<T>(const Geo &, const T *, T *, T *, const void *, int, hipStream_t);                \
    template void launch_ppe_rhs<T>(const Geo &, const T *, double *, hipStream_t, const void *);          \
    template unsigned launch_poisson<T>(const Geo &, const T *, T *, const int *, const double *, hipStream_t, \
                                        const T *, double *);                                              \
    template void launch_resid<T>(const Geo &, const double *, const double *, T *, double *, hipStream_t,   \
                                  double *);                                                             \
    template void launch_resid_upd<T>(const Geo &, const double *, const double *, const T *, const GmCtrl *,     \
                                      double *, T *, double *, hipStream_t, int);                              \
    template void launch_precond<T>(const Geo &, int, const FastPC *, const T *, T *, T *, const int *, hipStream_t, int); \
    template void launch_project<T>(const Geo &, double *, T *, T *, hipStream_t);                        \
    template void launch_bcs<T>(const Geo &, T *, const T *, T *, const double *, hipStream_t);           \
    template void launch_recycle<T>(const Geo &, T *, int, double, const double *, const double *, const double *, \
                                    const int *, const double *, const double *, hipStream_t);
TBL_INST(double)
TBL_INST(float)
#undef TBL_INST

}  // namespace tbl
