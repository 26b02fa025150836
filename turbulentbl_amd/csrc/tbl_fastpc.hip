// tbl_fastpc.hip -- fast-diagonalisation preconditioner for the pressure
// Poisson operator of dnssetbl/dnssetbl.f90:574-583.
//
// The reference's operator is separable:  A = Ax + Ay + Az with
//   Ax = tridiag(Aw, -2Aw, Ae)          in i, homogeneous Dirichlet at i = 1, nX,
//   Ay = tridiag(As, -2An, An)          in j, periodic with period nY-2 (jm1/jp1),
//   Az = tridiag(Ab(k), Ap(k)+2Aw+2An, At(k)) in k, Dirichlet at k = 1, nZ.
// Ay is diagonalised by the real orthonormal Fourier basis Fy.  Az is
// similar to a symmetric tridiagonal matrix through a diagonal scaling S
// (s_{k+1}/s_k = sqrt(Ab(k+1)/At(k)), well conditioned on the tanh grid), so
// Az = (S Q) diag(lz) (S Q)^{-1} with Q orthogonal (Jacobi eigensolver on the
// host).  After the two transforms every streamwise line (q, n) is a
// constant-coefficient tridiagonal system
//   Aw p(i-1) + (-2Aw + ly(n) + lz(q)) p(i) + Ae p(i+1) = r^(i),
// solved by one wavefront per line (partitioned Thomas in registers, cyclic
// reduction of the lane-boundary unknowns across the wave).
// The four transforms are products with small dense matrices on the matrix
// cores (hand-written kernels: split-bf16 operands for FP32 storage, FP64
// products for FP64 storage).
// As M^{-1} inside GMRES this keeps the reference's solver while needing only
// a few Krylov iterations.  The same transforms solve the shifted operator
// alpha I + beta A (the implicit-diffusion Helmholtz operator is I - (dt/Re) A)
// and, with the wall-normal transforms exchanged and transposed, the
// transposed systems the discrete adjoint needs (Ay and the line matrices are
// symmetric, Fy is orthonormal).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include <type_traits>

#include "tbl_dev.h"
#include "tbl_launch.h"

namespace tbl {

#define LS_MAXN 2048  // line-solver capacity: nX-2 <= 64 lanes x 32 unknowns

struct FastPC {
    int nx = 0, ny = 0, nz = 0;                          // interior sizes
    double *ly = nullptr, *lz = nullptr;                 // eigenvalues (full-index arrays)
    void *t1 = nullptr, *t2 = nullptr;                   // work fields (storage type)
    bool single = false;
    // FP32 storage: the transform matrices split into three bf16 parts in
    // matrix-core fragment order (see k_xform): z forward, y forward, y
    // inverse, z inverse, and the transposed z pair; null in FP64 contexts
    unsigned short *bz_f = nullptr, *by_f = nullptr, *by_i = nullptr, *bz_i = nullptr;
    unsigned short *bz_ft = nullptr, *bz_it = nullptr;
    // FP64 storage: the transforms in FP64 fragment order (see k_xform64):
    // z forward / inverse, y forward / inverse, and the transposed z pair of
    // the adjoint solves; null in FP32 contexts
    double *dz_f = nullptr, *dz_i = nullptr, *dy_f = nullptr, *dy_i = nullptr, *dz_ft = nullptr, *dz_it = nullptr;
};

static int setfail(char *err, int len, const char *m)
{
    if (err && len > 0) snprintf(err, len, "%s", m);
    return 1;
}

// cyclic Jacobi eigensolver for a dense symmetric n x n matrix (row-major a);
// on return a holds the eigenvalues on its diagonal and v the eigenvectors
// (columns), v row-major n x n.
static void jacobi_eig(int n, std::vector<double> &a, std::vector<double> &v)
{
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = 0; q < n; q++) {
                if (p == q) diag += a[(size_t)p * n + p] * a[(size_t)p * n + p];
                else off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
            }
        if (off <= 1e-30 * diag) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a[(size_t)p * n + q];
                if (fabs(apq) <= 1e-300) continue;
                const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {          // columns p, q
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - s * akq;
                    a[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {          // rows p, q
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - s * aqk;
                    a[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
}

// ----------------------------------------------------------------------------
// FP32 transforms on the matrix cores with split bf16 operands:
//   C[m, n] = sum_k A[m, k] B[k, n],  A[m, k] at A + row(m) + k lda,
//   row(m) = (m % Mp) + (m / Mp) bstride, C likewise with ldc.
// Every FP32 operand is written exactly as x = x_h + x_m + x_l with three
// bf16 parts (8 significant bits each: x_h keeps the top 8 bits of the
// significand, x_m the next 8, x_l the rest).  The six products with
// |part ratio| >= 2^-16 (hh, hm, mh, hl, lh, mm) are accumulated in FP32 by
// v_mfma_f32_32x32x16_bf16; the dropped terms are below 2^-23 relative, so
// the result has FP32 accuracy, as an SGEMM's would, at the bf16 matrix-core
// rate.  B (the transform, K x N) is split once on the host from its FP64
// values.
// A workgroup of 4 wavefronts owns 128 rows and NT 32-column tiles of C
// (accumulators for every tile in registers); wider transforms (more than
// 8 tiles) split the tiles over ncb column blocks, which run next to each
// other on one XCD so that their A rows are read from HBM once.  The k loop
// runs over groups of 16: the group's A values (8 per lane) are loaded two
// groups ahead and split in registers; the group's B fragments (3 parts x NT
// tiles) are staged through LDS, double-buffered, so that B is read from L2
// once per workgroup.
// The product is formed as C^T = B^T A^T, so the accumulator's lane index
// runs over rows m and the stores are coalesced along the contiguous row
// dimension.  Fragments: the k index of lane l, element e is (l / 32) * 8 + e
// of its 16-wide group, identical for both operands.
// ----------------------------------------------------------------------------
typedef __bf16 xf_v8bf __attribute__((ext_vector_type(8)));
typedef float xf_v16f __attribute__((ext_vector_type(16)));
typedef unsigned xf_v4u __attribute__((ext_vector_type(4)));

static inline unsigned short bf16_rne(float x)
{
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
static inline double bf16_val(unsigned short h)
{
    const unsigned u = (unsigned)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return (double)x;
}

// exact three-part split of two FP32 values, packed as bf16 pairs (element
// 2p in the low half, 2p + 1 in the high half)
__device__ __forceinline__ void xf_split2(float x0, float x1, unsigned &h, unsigned &m, unsigned &l)
{
    const unsigned h0 = __float_as_uint(x0) & 0xFFFF0000u, h1 = __float_as_uint(x1) & 0xFFFF0000u;
    const float r0 = x0 - __uint_as_float(h0), r1 = x1 - __uint_as_float(h1);
    const unsigned m0 = __float_as_uint(r0) & 0xFFFF0000u, m1 = __float_as_uint(r1) & 0xFFFF0000u;
    const unsigned l0 = __float_as_uint(r0 - __uint_as_float(m0)), l1 = __float_as_uint(r1 - __uint_as_float(m1));
    h = (h0 >> 16) | h1;
    m = (m0 >> 16) | m1;
    l = (l0 >> 16) | (l1 & 0xFFFF0000u);
}

#define XF_ROWS 128
#ifndef XF_MAXT
#define XF_MAXT 8
#endif
template <int NT>
__global__ __launch_bounds__(256) void k_xform(const float *__restrict__ A, float *__restrict__ C,
                                              const xf_v4u *__restrict__ Bf, int K, int KG, int N, long Mtot, long Mp,
                                              long bstride, long lda, long ldc, int NTt, int ncb,
                                              const int *__restrict__ active)
{
    if (active && !*active) return;                          // gated launch of an ended solve
    constexpr int SL = 3 * NT * 64;                          // fragments of one k group
    constexpr int PT = (SL + 255) / 256;                     // staged per thread
    __shared__ xf_v4u sB[2][PT * 256];                        // padded: staging needs no bounds test
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // 1-D grid: row block rb, column block cb (fastest), XCD-contiguous
    const unsigned L = xcd_remap(blockIdx.x, gridDim.x);
    const int cb = (int)(L % (unsigned)ncb);
    const long rb = (long)(L / (unsigned)ncb);
    const int nt0 = cb * NT;                                 // first tile of this column block
    const long m = rb * XF_ROWS + wv * 32 + (lane & 31);
    const bool mok = m < Mtot;
    const float *Ar = A + (mok ? (m % Mp) + (m / Mp) * bstride : 0);
    const int kh = (lane >> 5) * 8;
    xf_v16f acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] = xf_v16f{};
    // loads are unconditional at in-range addresses, masked after (the
    // padding k >= K and rows m >= Mtot are zeroed at the split, so that
    // nothing consumes a load before it is needed).  Offsets are 32-bit (the
    // field holds fewer than 2^30 values, checked at create time):
    // row + min(k lda, (K-1) lda), with the uniform part k0 lda scalar.
    const int row32 = (int)(Ar - A), lda32 = (int)lda, khl = kh * lda32, kcap = (K - 1) * lda32;
    auto ldA = [&](int kgl, float *dst) {
        const int s0 = kgl * 16 * lda32;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const unsigned off = (unsigned)(row32 + min(s0 + e * lda32 + khl, kcap));
            dst[e] = __builtin_nontemporal_load((const float *)((const char *)A + off * 4u));
        }
    };
    // staged fragment e of a group: part e / (NT 64), tile nt0 + (e / 64) % NT
    // (clamped to the last tile: columns past N are computed and not stored)
    int boff[PT];
#pragma unroll
    for (int q = 0; q < PT; q++) {
        const int e = min(t + 256 * q, SL - 1), part = e / (NT * 64), nt = (e / 64) % NT;
        boff[q] = (part * NTt + min(nt0 + nt, NTt - 1)) * 64 + (e & 63);
    }
    auto ldB = [&](int kgl, xf_v4u *dst) {
        const long g0 = (long)min(kgl, KG - 1) * 3 * NTt * 64;
#pragma unroll
        for (int q = 0; q < PT; q++) dst[q] = Bf[g0 + boff[q]];
    };
    // Two-deep pipeline over k groups, unrolled by two so that each parity
    // has its own staging registers: the step of group g splits A(g) (loaded
    // two steps earlier), issues the loads of group g + 2 into the registers
    // it has just freed, forms the 6 NT products from the B slice in LDS
    // buffer g % 2, then stores the B slice of group g + 1 (loaded one step
    // earlier) into the other buffer.  The group count is padded to even; a
    // padding group is masked to zero at the split.
    float an[2][8];
    xf_v4u st[2][PT];
    ldA(0, an[0]);
    ldB(0, st[0]);
    ldA(1, an[1]);
    ldB(1, st[1]);
#pragma unroll
    for (int q = 0; q < PT; q++) sB[0][t + 256 * q] = st[0][q];
    __syncthreads();
    auto step = [&](int kg, auto bc) {
        constexpr int b = decltype(bc)::value;
        xf_v4u ah, am, al;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            unsigned h, mm, l;
            const int k0 = kg * 16 + kh + 2 * p;
            xf_split2((mok && k0 < K) ? an[b][2 * p] : 0.0f, (mok && k0 + 1 < K) ? an[b][2 * p + 1] : 0.0f, h, mm,
                      l);
            ah[p] = h; am[p] = mm; al[p] = l;
        }
        __builtin_amdgcn_sched_barrier(0);
        ldA(kg + 2, an[b]);
        ldB(kg + 2, st[b]);
        __builtin_amdgcn_sched_barrier(0);
        const xf_v8bf Ah = __builtin_bit_cast(xf_v8bf, ah), Am = __builtin_bit_cast(xf_v8bf, am),
                      Al = __builtin_bit_cast(xf_v8bf, al);
        // B fragments are read one n tile ahead of their products
        xf_v4u bh = sB[b][lane], bm = sB[b][NT * 64 + lane], bl = sB[b][2 * NT * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const xf_v8bf Bh = __builtin_bit_cast(xf_v8bf, bh);
            const xf_v8bf Bm = __builtin_bit_cast(xf_v8bf, bm);
            const xf_v8bf Bl = __builtin_bit_cast(xf_v8bf, bl);
            if (nt + 1 < NT) {
                bh = sB[b][(nt + 1) * 64 + lane];
                bm = sB[b][(NT + nt + 1) * 64 + lane];
                bl = sB[b][(2 * NT + nt + 1) * 64 + lane];
            }
            xf_v16f c = acc[nt];
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bm, Am, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bl, Ah, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bh, Al, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bm, Ah, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bh, Am, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bh, Ah, c, 0, 0, 0);
            acc[nt] = c;
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int q = 0; q < PT; q++) sB[b ^ 1][t + 256 * q] = st[b ^ 1][q];
        __syncthreads();
    };
    for (int kg = 0; kg < KG; kg += 2) {
        step(kg, std::integral_constant<int, 0>{});
        step(kg + 1, std::integral_constant<int, 1>{});
    }
    if (mok) {
        float *Cr = C + (Ar - A);
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int n = (nt0 + nt) * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (n < N) Cr[(long)n * ldc] = acc[nt][r];
            }
    }
}

// B (K x N, b(k, n)) split into bf16 parts and laid out as the first MFMA
// operand, slice by k group: [kg][part][n tile][lane] x 8 (B^T: row n =
// lane % 32 of the tile, k = (lane / 32) * 8 + e of the group)
template <class Fn>
static std::vector<unsigned short> xform_split(int K, int N, Fn b)
{
    const int KG = (K + 15) / 16, NT = (N + 31) / 32;
    std::vector<unsigned short> out((size_t)KG * 3 * NT * 64 * 8, 0);
    for (int kg = 0; kg < KG; kg++)
        for (int nt = 0; nt < NT; nt++)
            for (int l = 0; l < 64; l++)
                for (int e = 0; e < 8; e++) {
                    const int k = kg * 16 + (l >> 5) * 8 + e, n = nt * 32 + (l & 31);
                    double x = (k < K && n < N) ? b(k, n) : 0.0;
                    for (int part = 0; part < 3; part++) {
                        const unsigned short h = bf16_rne((float)x);
                        out[((((size_t)kg * 3 + part) * NT + nt) * 64 + l) * 8 + e] = h;
                        x -= bf16_val(h);
                    }
                }
    return out;
}

static void launch_xform(int K, int N, const unsigned short *Bf, const float *A, float *C, long Mtot, long Mp,
                         long bstride, long lda, long ldc, hipStream_t s, const int *active = nullptr)
{
    const int KG = (K + 15) / 16, NTt = (N + 31) / 32;
    const int ncb = (NTt + XF_MAXT - 1) / XF_MAXT, NT = (NTt + ncb - 1) / ncb;   // at most XF_MAXT tiles per block
    const dim3 grid((unsigned)(((Mtot + XF_ROWS - 1) / XF_ROWS) * ncb)), blk(256);
    const xf_v4u *B = (const xf_v4u *)Bf;
#define XF_CASE(q) case q: hipLaunchKernelGGL(k_xform<q>, grid, blk, 0, s, A, C, B, K, KG, N, Mtot, Mp, bstride, lda, ldc, NTt, ncb, active); break;
    switch (NT) {
        XF_CASE(1) XF_CASE(2) XF_CASE(3) XF_CASE(4) XF_CASE(5) XF_CASE(6) XF_CASE(7) XF_CASE(8)
        default: break;
    }
#undef XF_CASE
}

// ----------------------------------------------------------------------------
// FP64 transforms on the matrix cores: the same product and row map as the
// FP32 kernel, C[m, n] = sum_k A[m, k] B[k, n], formed in FP64 throughout by
// v_mfma_f64_16x16x4_f64 (a DGEMM's arithmetic, no library).
// Fragments of the 16 x 16 x 4 product (64 lanes, one double each):
//   first operand (16 x 4):  lane l holds row l % 16, k = l / 16;
//   second operand (4 x 16): lane l holds k = l / 16, column l % 16;
//   result (16 x 16):        lane l, element r holds row 4 r + l / 16,
//                            column l % 16 (the FP64 product's own order,
//                            unlike the 4 (l / 16) + r of the FP32 ones).
// The product is formed as C^T = B^T A^T: the result's rows are columns n of
// C and its columns rows m, so a lane's results lie along the contiguous row
// index m and each store writes four 128-byte row segments.
// A workgroup of 4 wavefronts owns 64 MT rows (MT row tiles of 16 per wave)
// and NT 16-column tiles.  The k loop runs over groups of 16 (four products
// per tile pair): the group's A values (4 MT per lane) are loaded two groups
// ahead into registers; its B slice, packed on the host in fragment order
// (NT x 256 doubles, contiguous per column block), is staged through LDS,
// double-buffered, and read by all four waves.
// ----------------------------------------------------------------------------
typedef double xd_v4d __attribute__((ext_vector_type(4)));

template <int MT, int NT>
__global__ __launch_bounds__(256) void k_xform64(const double *__restrict__ A, double *__restrict__ C,
                                                const double *__restrict__ Bf, int K, int KG, int N, long Mtot,
                                                long Mp, long bstride, long lda, long ldc, int NTt, int ncb,
                                                const int *__restrict__ active)
{
    if (active && !*active) return;                          // gated launch of an ended solve
    __shared__ double sB[2][NT * 256];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, kl = lane >> 4;
    // 1-D grid: row block rb, column block cb (fastest), XCD-contiguous
    const unsigned L = xcd_remap(blockIdx.x, gridDim.x);
    const int cb = (int)(L % (unsigned)ncb);
    const long rb = (long)(L / (unsigned)ncb);
    const int nt0 = cb * NT;
    // row offsets (element units; fields hold fewer than 2^31 values), rows
    // past Mtot clamped to the last row: loaded, never stored
    unsigned rowo[MT];
    bool mok[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        const long m = rb * (64 * MT) + wv * (16 * MT) + mt * 16 + (lane & 15);
        mok[mt] = m < Mtot;
        const long mm = mok[mt] ? m : Mtot - 1;
        rowo[mt] = (unsigned)((mm % Mp) + (mm / Mp) * bstride);
    }
    const unsigned lda32 = (unsigned)lda;
    // A values of group kgl: element mt * 4 + ks is A[m(mt), kgl 16 + 4 ks + kl]
    // (k clamped in range; the value is masked to zero at use when k >= K)
    auto ldA = [&](int kgl, double *dst) {
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const unsigned k = (unsigned)min(kgl * 16 + ks * 4 + kl, K - 1);
#pragma unroll
            for (int mt = 0; mt < MT; mt++)
                dst[mt * 4 + ks] = __builtin_nontemporal_load(A + (rowo[mt] + k * lda32));
        }
    };
    // B slice of group kgl for this column block: tile q of the block (clamped
    // to the last tile: columns past N are computed and not stored)
    auto ldB = [&](int kgl, double *dst) {
        const long g0 = (long)min(kgl, KG - 1) * NTt * 256;
#pragma unroll
        for (int q = 0; q < NT; q++) dst[q] = Bf[g0 + (long)min(nt0 + q, NTt - 1) * 256 + t];
    };
    xd_v4d acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[mt][nt] = xd_v4d{};
    // Two-deep pipeline over k groups, unrolled by two so that each parity
    // has its own staging registers (as in the FP32 kernel): the step of group
    // g masks A(g), issues the loads of group g + 2 into the registers it has
    // just freed, forms the 4 MT NT products from LDS buffer g % 2, then
    // stores the B slice of group g + 1 into the other buffer.  The group
    // count is padded to even; a padding group's A values are masked to zero.
    double an[2][4 * MT];
    double st[2][NT];
    ldA(0, an[0]);
    ldB(0, st[0]);
    ldA(1, an[1]);
    ldB(1, st[1]);
#pragma unroll
    for (int q = 0; q < NT; q++) sB[0][q * 256 + t] = st[0][q];
    __syncthreads();
    auto step = [&](int kg, auto bc) {
        constexpr int b = decltype(bc)::value;
        double a[4 * MT];
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const bool kin = kg * 16 + ks * 4 + kl < K;
#pragma unroll
            for (int mt = 0; mt < MT; mt++) a[mt * 4 + ks] = kin ? an[b][mt * 4 + ks] : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);
        ldA(kg + 2, an[b]);
        ldB(kg + 2, st[b]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const double bf = sB[b][nt * 256 + ks * 64 + lane];
#pragma unroll
                for (int mt = 0; mt < MT; mt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf, a[mt * 4 + ks], acc[mt][nt], 0, 0, 0);
            }
#pragma unroll
        for (int q = 0; q < NT; q++) sB[b ^ 1][q * 256 + t] = st[b ^ 1][q];
        __syncthreads();
    };
    for (int kg = 0; kg < KG; kg += 2) {
        step(kg, std::integral_constant<int, 0>{});
        step(kg + 1, std::integral_constant<int, 1>{});
    }
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        if (!mok[mt]) continue;
        double *Cr = C + rowo[mt];
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int n = (nt0 + nt) * 16 + 4 * r + kl;
                if (n < N) Cr[(long)n * ldc] = acc[mt][nt][r];
            }
    }
}

// B (K x N, b(k, n)) in FP64 fragment order, slice by k group:
// [kg][n tile][k step][lane], element = b(16 kg + 4 step + lane / 16,
// 16 tile + lane % 16), zero past K and N
template <class Fn>
static std::vector<double> xform64_pack(int K, int N, Fn b)
{
    const int KG = (K + 15) / 16, NT = (N + 15) / 16;
    std::vector<double> out((size_t)KG * NT * 256, 0.0);
    for (int kg = 0; kg < KG; kg++)
        for (int nt = 0; nt < NT; nt++)
            for (int ks = 0; ks < 4; ks++)
                for (int l = 0; l < 64; l++) {
                    const int k = kg * 16 + ks * 4 + (l >> 4), n = nt * 16 + (l & 15);
                    if (k < K && n < N) out[(((size_t)kg * NT + nt) * 4 + ks) * 64 + l] = b(k, n);
                }
    return out;
}

// Launch shape: at most 8 column tiles per block; two row tiles per wave
// unless that leaves fewer than 2048 workgroups, then one; narrower column
// blocks while the grid is still below 2048 workgroups and a block keeps at
// least two tiles.
static void launch_xform64(int K, int N, const double *Bf, const double *A, double *C, long Mtot, long Mp,
                           long bstride, long lda, long ldc, hipStream_t s, const int *active = nullptr)
{
    const int KG = (K + 15) / 16, NTt = (N + 15) / 16;
    int ncb = (NTt + 7) / 8;
    auto blocks = [&](int mt, int nc) { return ((Mtot + 64 * mt - 1) / (64 * mt)) * nc; };
    const int MT = blocks(2, ncb) >= 2048 ? 2 : 1;
    while (blocks(MT, ncb) < 2048 && (NTt + ncb) / (ncb + 1) >= 2) ncb++;
    const int NT = (NTt + ncb - 1) / ncb;
    ncb = (NTt + NT - 1) / NT;
    const dim3 grid((unsigned)blocks(MT, ncb)), blk(256);
#define XD_CASE(m, q) case q: hipLaunchKernelGGL((k_xform64<m, q>), grid, blk, 0, s, A, C, Bf, K, KG, N, Mtot, Mp, bstride, lda, ldc, NTt, ncb, active); break;
    if (MT == 2) {
        switch (NT) {
            XD_CASE(2, 1) XD_CASE(2, 2) XD_CASE(2, 3) XD_CASE(2, 4) XD_CASE(2, 5) XD_CASE(2, 6) XD_CASE(2, 7) XD_CASE(2, 8)
            default: break;
        }
    } else {
        switch (NT) {
            XD_CASE(1, 1) XD_CASE(1, 2) XD_CASE(1, 3) XD_CASE(1, 4) XD_CASE(1, 5) XD_CASE(1, 6) XD_CASE(1, 7) XD_CASE(1, 8)
            default: break;
        }
    }
#undef XD_CASE
}

int fastpc_create(const Geo &g, const double *Ab, const double *At, const double *Ap, hipStream_t s, bool single,
                  FastPC **out, char *err, int errlen)
{
    const int nx = g.nX - 2, ny = g.nY - 2, nz = g.nZ - 2;
    if (nx > LS_MAXN) return setfail(err, errlen, "fastpc: nX-2 exceeds the line-solver capacity (2048)");
    FastPC *f = new FastPC();
    f->single = single;
    f->nx = nx; f->ny = ny; f->nz = nz;
    // --- z: symmetrise Az (interior k = 1..nZ-2) and diagonalise
    std::vector<double> sc(nz), T((size_t)nz * nz, 0.0), Q;
    sc[0] = 1.0;
    for (int q = 0; q < nz - 1; q++) {
        const double lo = Ab[q + 2], up = At[q + 1];          // Ab(k+1), At(k) for k = q+1
        if (!(lo > 0.0) || !(up > 0.0)) { delete f; return setfail(err, errlen, "fastpc: Ab*At <= 0, operator not symmetrisable"); }
        sc[q + 1] = sc[q] * sqrt(lo / up);
        const double e = sqrt(lo * up);
        T[(size_t)q * nz + q + 1] = e;
        T[(size_t)(q + 1) * nz + q] = e;
    }
    for (int q = 0; q < nz; q++) T[(size_t)q * nz + q] = Ap[q + 1] + 2.0 * g.Aw + 2.0 * g.An;
    jacobi_eig(nz, T, Q);
    // forward Wf(k,q) = Q(k,q)/s_k ; inverse Wi(q,k) = s_k Q(k,q)  (column-major nz x nz)
    std::vector<double> Wf((size_t)nz * nz), Wi((size_t)nz * nz), lz(g.nZ, 0.0);
    for (int q = 0; q < nz; q++) {
        lz[q + 1] = T[(size_t)q * nz + q];
        for (int k = 0; k < nz; k++) {
            Wf[k + (size_t)nz * q] = Q[(size_t)k * nz + q] / sc[k];
            Wi[q + (size_t)nz * k] = Q[(size_t)k * nz + q] * sc[k];
        }
    }
    // --- y: real Fourier basis over the period ny (even): column 0 constant,
    // columns 2m-1 / 2m cos / sin of wavenumber m, last column Nyquist
    std::vector<double> Fy((size_t)ny * ny), ly(g.nY, 0.0);
    for (int nn = 0; nn < ny; nn++) {
        const int m = (nn == 0) ? 0 : (nn == ny - 1 ? ny / 2 : (nn + 1) / 2);
        for (int j = 0; j < ny; j++) {
            double v;
            const double th = 2.0 * M_PI * (double)((long)m * j % ny) / ny;
            if (nn == 0) v = 1.0 / sqrt((double)ny);
            else if (nn == ny - 1) v = ((j & 1) ? -1.0 : 1.0) / sqrt((double)ny);
            else if (nn & 1) v = sqrt(2.0 / ny) * cos(th);
            else v = sqrt(2.0 / ny) * sin(th);
            Fy[j + (size_t)ny * nn] = v;
        }
        ly[nn + 1] = 2.0 * g.An * cos(2.0 * M_PI * m / ny) - 2.0 * g.An;
    }
    const size_t es = single ? sizeof(float) : sizeof(double);
    hipError_t e = hipSuccess;
    auto al = [&](void **p, size_t b) { if (hipMalloc(p, b) != hipSuccess) e = hipErrorOutOfMemory; };
    al((void **)&f->ly, sizeof(double) * g.nY);
    al((void **)&f->lz, sizeof(double) * g.nZ);
    al(&f->t1, es * g.N);
    al(&f->t2, es * g.N);
    if (e != hipSuccess) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: hipMalloc failed"); }
    (void)hipMemcpyAsync(f->ly, ly.data(), sizeof(double) * g.nY, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(f->lz, lz.data(), sizeof(double) * g.nZ, hipMemcpyHostToDevice, s);
    (void)hipMemsetAsync(f->t1, 0, es * g.N, s);
    (void)hipMemsetAsync(f->t2, 0, es * g.N, s);
    // transform matrices as b(k, n): C[m, n] = sum_k A[m, k] b(k, n)
    auto wf = [&](int k, int n) { return Wf[k + (size_t)nz * n]; };     // z forward
    auto wi = [&](int k, int n) { return Wi[k + (size_t)nz * n]; };     // z inverse
    auto fy = [&](int k, int n) { return Fy[k + (size_t)ny * n]; };     // y forward
    auto fyt = [&](int k, int n) { return Fy[n + (size_t)ny * k]; };    // y inverse
    auto wit = [&](int k, int n) { return Wi[n + (size_t)nz * k]; };    // transposed z forward
    auto wft = [&](int k, int n) { return Wf[n + (size_t)nz * k]; };    // transposed z inverse
    // the host copies must outlive the asynchronous uploads: synchronised below
    std::vector<std::vector<unsigned short>> hb;
    std::vector<std::vector<double>> hd;
    if (single) {
        // split bf16 transform matrices for the FP32 matrix-core kernel
        hb.push_back(xform_split(nz, nz, wf));
        hb.push_back(xform_split(ny, ny, fy));
        hb.push_back(xform_split(ny, ny, fyt));
        hb.push_back(xform_split(nz, nz, wi));
        hb.push_back(xform_split(nz, nz, wit));
        hb.push_back(xform_split(nz, nz, wft));
        unsigned short **dst[6] = {&f->bz_f, &f->by_f, &f->by_i, &f->bz_i, &f->bz_ft, &f->bz_it};
        for (int q = 0; q < 6; q++) al((void **)dst[q], hb[q].size() * 2);
        if (e != hipSuccess) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: hipMalloc failed"); }
        for (int q = 0; q < 6; q++) (void)hipMemcpyAsync(*dst[q], hb[q].data(), hb[q].size() * 2, hipMemcpyHostToDevice, s);
    } else {
        // FP64 fragment-packed transform matrices for the FP64 matrix-core kernel
        hd.push_back(xform64_pack(nz, nz, wf));
        hd.push_back(xform64_pack(nz, nz, wi));
        hd.push_back(xform64_pack(ny, ny, fy));
        hd.push_back(xform64_pack(ny, ny, fyt));
        hd.push_back(xform64_pack(nz, nz, wit));
        hd.push_back(xform64_pack(nz, nz, wft));
        double **dst[6] = {&f->dz_f, &f->dz_i, &f->dy_f, &f->dy_i, &f->dz_ft, &f->dz_it};
        for (int q = 0; q < 6; q++) al((void **)dst[q], hd[q].size() * sizeof(double));
        if (e != hipSuccess) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: hipMalloc failed"); }
        for (int q = 0; q < 6; q++)
            (void)hipMemcpyAsync(*dst[q], hd[q].data(), hd[q].size() * sizeof(double), hipMemcpyHostToDevice, s);
    }
    if (hipStreamSynchronize(s) != hipSuccess) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: upload failed"); }
    *out = f;
    return 0;
}

void fastpc_destroy(FastPC *f)
{
    if (!f) return;
    void *p[] = {f->ly, f->lz, f->t1, f->t2, f->bz_f, f->by_f, f->by_i, f->bz_i, f->bz_ft, f->bz_it,
                 f->dz_f, f->dz_i, f->dy_f, f->dy_i, f->dz_ft, f->dz_it};
    for (void *q : p)
        if (q) hipFree(q);
    delete f;
}

// Streamwise line solve of the transformed system, one wavefront per line (j, k):
//   ow x(i-1) + dg x(i) + ow x(i+1) = d(i),  i = 0..n-1,  x(-1) = x(n) = 0,
// ow = beta Aw, dg = alpha + beta (-2Aw + ly(j) + lz(k)); for the Poisson
// operator (alpha 0, beta 1) and the Helmholtz operator (alpha 1, beta -dt/Re)
// |dg| > 2|ow| (strictly diagonally dominant).
// Each lane owns M consecutive unknowns (the line padded to 64 M with identity
// rows).  A local elimination writes every unknown of the lane through the
// lane's first and last unknowns (partitioned Thomas); those 128 unknowns form
// a tridiagonal system solved by cyclic reduction across the wave with lane
// shuffles; a last local pass recovers the rest.  The line passes through LDS
// once each way so that the global loads and stores are coalesced.
// The FP32 solve is bound by instruction issue, not by HBM: it takes its
// reciprocals from v_rcp_f32 (1 ulp), one instruction where a correctly
// rounded division is about ten; the solve is a preconditioner, which FP32
// rounding of this size does not change.  FP64 divides exactly.
#define LS_WAVES 2
template <class T> __device__ __forceinline__ T ls_div(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float ls_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <class T, int M>
// FP32: registers allocated for 6 waves per SIMD (80 VGPRs, no spill; the
// compiler's own choice, 81, allows 5)
__global__ __launch_bounds__(64 * LS_WAVES) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
void k_line_solve(Geo g, T *__restrict__ d, const double *__restrict__ ly,
                                                             const double *__restrict__ lz, const int *__restrict__ active,
                                                             double alpha, double beta)
{
    if (active && !*active) return;
    constexpr int P = 64 * M + 64;                    // one pad slot per lane chunk
    __shared__ T sl[LS_WAVES][P];
    const int n = g.nX - 2, ny = g.nY - 2, nz = g.nZ - 2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long L = (long)blockIdx.x * LS_WAVES + wv;
    const bool live = L < (long)ny * nz;
    const int j = live ? (int)(L % ny) + 1 : 1, k = live ? (int)(L / ny) + 1 : 1;
    T *line = d + 1 + (long)g.nX * ((long)j + (long)g.nY * k);
    T *s = sl[wv];
    auto pos = [](int i) { return i + i / M; };
#pragma unroll
    for (int t = 0; t < M; t++) {
        const int i = lane + 64 * t;
        s[pos(i)] = (live && i < n) ? line[i] : T(0);
    }
    __syncthreads();
    const T aw = T(beta * g.Aw);
    const T dg = T(alpha + beta * (-2.0 * g.Aw + ly[j] + lz[k]));
    const int i0 = lane * M;
    auto coef = [&](int gi, T &a, T &b, T &c) {
        const bool in = gi < n;
        a = (in && gi > 0) ? aw : T(0);
        b = in ? dg : T(1);
        c = (in && gi < n - 1) ? aw : T(0);
    };
    T ap[M], cp[M], dp[M];
    T a0, b0, c0;
    coef(i0, a0, b0, c0);
    const T d0 = s[pos(i0)];
    {
        T a, b, c;
        coef(i0 + 1, a, b, c);
        const T r = ls_div(T(1), b);
        ap[1] = a * r; cp[1] = c * r; dp[1] = s[pos(i0 + 1)] * r;
    }
    // forward: rows e >= 1 become a'(e) x(first) + x(e) + c'(e) x(e+1) = d'(e)
#pragma unroll
    for (int e = 2; e < M; e++) {
        T a, b, c;
        coef(i0 + e, a, b, c);
        const T r = ls_div(T(1), b - a * cp[e - 1]);
        ap[e] = -(a * ap[e - 1]) * r;
        cp[e] = c * r;
        dp[e] = (s[pos(i0 + e)] - a * dp[e - 1]) * r;
    }
    // backward: rows 1..M-2 become a''(e) x(first) + x(e) + c''(e) x(last) = d''(e)
#pragma unroll
    for (int e = M - 3; e >= 1; e--) {
        ap[e] = ap[e] - cp[e] * ap[e + 1];
        dp[e] = dp[e] - cp[e] * dp[e + 1];
        cp[e] = -(cp[e] * cp[e + 1]);
    }
    // reduced rows: F (first unknown, sequence index 2 lane) and L (last, 2 lane + 1)
    T AF, BF, CF, DF;
    if (M == 2) { AF = a0; BF = b0; CF = c0; DF = d0; }
    else { AF = a0; BF = b0 - c0 * ap[1]; CF = -(c0 * cp[1]); DF = d0 - c0 * dp[1]; }
    T AL = ap[M - 1], BL = T(1), CL = cp[M - 1], DL = dp[M - 1];
    // cyclic reduction over the 128-row sequence F0 L0 F1 L1 ...
    auto red = [](T &A, T &B, T &C, T &D, T Am, T Bm, T Cm, T Dm, T Ap, T Bp, T Cp, T Dp) {
        const T k1 = ls_div(A, Bm), k2 = ls_div(C, Bp);
        B = B - Cm * k1 - Ap * k2;
        D = D - Dm * k1 - Dp * k2;
        A = -(Am * k1);
        C = -(Cp * k2);
    };
    {
        // stride 1: F(l) sits between L(l-1) and L(l); L(l) between F(l) and F(l+1)
        const T lA = __shfl(AL, lane - 1), lB = __shfl(BL, lane - 1), lC = __shfl(CL, lane - 1), lD = __shfl(DL, lane - 1);
        const T uA = __shfl(AF, lane + 1), uB = __shfl(BF, lane + 1), uC = __shfl(CF, lane + 1), uD = __shfl(DF, lane + 1);
        const bool hl = lane > 0, hu = lane < 63;
        const T fA = AF, fB = BF, fC = CF, fD = DF, gA = AL, gB = BL, gC = CL, gD = DL;
        red(AF, BF, CF, DF, hl ? lA : T(0), hl ? lB : T(1), hl ? lC : T(0), hl ? lD : T(0), gA, gB, gC, gD);
        red(AL, BL, CL, DL, fA, fB, fC, fD, hu ? uA : T(0), hu ? uB : T(1), hu ? uC : T(0), hu ? uD : T(0));
    }
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
        // stride 2h: rows of the same kind h lanes away
        const bool hl = lane >= h, hu = lane + h < 64;
        const int sl_ = hl ? lane - h : lane, su = hu ? lane + h : lane;
        const T fmA = __shfl(AF, sl_), fmB = __shfl(BF, sl_), fmC = __shfl(CF, sl_), fmD = __shfl(DF, sl_);
        const T fpA = __shfl(AF, su), fpB = __shfl(BF, su), fpC = __shfl(CF, su), fpD = __shfl(DF, su);
        const T lmA = __shfl(AL, sl_), lmB = __shfl(BL, sl_), lmC = __shfl(CL, sl_), lmD = __shfl(DL, sl_);
        const T lpA = __shfl(AL, su), lpB = __shfl(BL, su), lpC = __shfl(CL, su), lpD = __shfl(DL, su);
        red(AF, BF, CF, DF, hl ? fmA : T(0), hl ? fmB : T(1), hl ? fmC : T(0), hl ? fmD : T(0),
            hu ? fpA : T(0), hu ? fpB : T(1), hu ? fpC : T(0), hu ? fpD : T(0));
        red(AL, BL, CL, DL, hl ? lmA : T(0), hl ? lmB : T(1), hl ? lmC : T(0), hl ? lmD : T(0),
            hu ? lpA : T(0), hu ? lpB : T(1), hu ? lpC : T(0), hu ? lpD : T(0));
    }
    const T xf = ls_div(DF, BF), xl = ls_div(DL, BL);
    s[pos(i0)] = xf;
    s[pos(i0 + M - 1)] = xl;
#pragma unroll
    for (int e = 1; e < M - 1; e++) s[pos(i0 + e)] = dp[e] - ap[e] * xf - cp[e] * xl;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < M; t++) {
        const int i = lane + 64 * t;
        if (live && i < n) line[i] = s[pos(i)];
    }
}

template <class T>
static void launch_line_solve(const Geo &g, T *d, const double *ly, const double *lz, const int *active, double alpha,
                              double beta, hipStream_t st)
{
    const int n = g.nX - 2;
    const long lines = (long)(g.nY - 2) * (g.nZ - 2);
    const dim3 grid((unsigned)((lines + LS_WAVES - 1) / LS_WAVES)), blk(64 * LS_WAVES);
    const int m = (n + 63) / 64;
    if (m <= 2) hipLaunchKernelGGL((k_line_solve<T, 2>), grid, blk, 0, st, g, d, ly, lz, active, alpha, beta);
    else if (m <= 4) hipLaunchKernelGGL((k_line_solve<T, 4>), grid, blk, 0, st, g, d, ly, lz, active, alpha, beta);
    else if (m <= 8) hipLaunchKernelGGL((k_line_solve<T, 8>), grid, blk, 0, st, g, d, ly, lz, active, alpha, beta);
    else if (m <= 16) hipLaunchKernelGGL((k_line_solve<T, 16>), grid, blk, 0, st, g, d, ly, lz, active, alpha, beta);
    else hipLaunchKernelGGL((k_line_solve<T, 32>), grid, blk, 0, st, g, d, ly, lz, active, alpha, beta);
}

// out = (alpha I + beta A)^{-1} r, or its transpose, on the interior planes
// (face rows of out come out zero).  The transposed system
// (Wi^T Y Wf^T)^T = Wf Y Wi is solved with the z transforms exchanged and
// transposed (Ay and the line matrices are symmetric).
// Rows: the z transforms map all (i,j) rows of planes 1..nz (face rows of r
// are zero: Krylov vectors and residuals are never written there, so they
// map to zero rows); the y transforms map the rows i = 0..nX-1 of planes
// 1..nz, batched over planes (face rows are zero in and out).
// part: the whole solve (FPC_ALL), its first z transform only (FPC_HEAD) or
// the rest (FPC_TAIL), so that the first transform can be queued ahead of a
// host poll.  With active, every launch returns at once when *active is 0.
template <class T>
int run_fastpc_ext(const FastPC *f, const Geo &g, const T *r, T *out, const int *active, double alpha, double beta,
                   bool transpose, hipStream_t s, int part)
{
    const bool head = part != FPC_TAIL, tail = part != FPC_HEAD;
    const int ny = f->ny, nz = f->nz, nX = g.nX;
    const long pst = g.sZ;                     // plane stride nX*nY
    const long y0 = (long)nX + pst;            // (i,j,k) = (0,1,1)
    T *u1 = (T *)f->t1, *u2 = (T *)f->t2;
    if (sizeof(T) == sizeof(float)) {
        if (!f->single) return 1;
        const float *rf = (const float *)(const void *)r;
        float *of = (float *)(void *)out, *v1 = (float *)(void *)u1, *v2 = (float *)(void *)u2;
        const unsigned short *zf = transpose ? f->bz_ft : f->bz_f, *zi = transpose ? f->bz_it : f->bz_i;
        if (head) launch_xform(nz, nz, zf, rf + pst, v1 + pst, pst, pst, 0, pst, pst, s, active);
        if (tail) {
            launch_xform(ny, ny, f->by_f, v1 + y0, v2 + y0, (long)nX * nz, nX, pst, nX, nX, s, active);
            launch_line_solve<float>(g, v2, f->ly, f->lz, active, alpha, beta, s);
            launch_xform(ny, ny, f->by_i, v2 + y0, v1 + y0, (long)nX * nz, nX, pst, nX, nX, s, active);
            launch_xform(nz, nz, zi, v1 + pst, of + pst, pst, pst, 0, pst, pst, s, active);
        }
    } else {
        if (f->single) return 1;
        const double *rd = (const double *)(const void *)r;
        double *od = (double *)(void *)out, *v1 = (double *)(void *)u1, *v2 = (double *)(void *)u2;
        const double *zf = transpose ? f->dz_ft : f->dz_f, *zi = transpose ? f->dz_it : f->dz_i;
        if (head) launch_xform64(nz, nz, zf, rd + pst, v1 + pst, pst, pst, 0, pst, pst, s, active);
        if (tail) {
            launch_xform64(ny, ny, f->dy_f, v1 + y0, v2 + y0, (long)nX * nz, nX, pst, nX, nX, s, active);
            launch_line_solve<double>(g, v2, f->ly, f->lz, active, alpha, beta, s);
            launch_xform64(ny, ny, f->dy_i, v2 + y0, v1 + y0, (long)nX * nz, nX, pst, nX, nX, s, active);
            launch_xform64(nz, nz, zi, v1 + pst, od + pst, pst, pst, 0, pst, pst, s, active);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// out = M^{-1} r, the preconditioner (alpha 0, beta 1)
template <class T>
int run_fastpc(const FastPC *f, const Geo &g, const T *r, T *out, const int *active, hipStream_t s, int part)
{
    return run_fastpc_ext<T>(f, g, r, out, active, 0.0, 1.0, false, s, part);
}

template int run_fastpc<double>(const FastPC *, const Geo &, const double *, double *, const int *, hipStream_t, int);
template int run_fastpc_ext<double>(const FastPC *, const Geo &, const double *, double *, const int *, double, double,
                                    bool, hipStream_t, int);
template int run_fastpc<float>(const FastPC *, const Geo &, const float *, float *, const int *, hipStream_t, int);
template int run_fastpc_ext<float>(const FastPC *, const Geo &, const float *, float *, const int *, double, double,
                                   bool, hipStream_t, int);

}  // namespace tbl
