// tbl_dev.h -- device-side definitions shared by the HIP translation units of
// libtbl_hip.so (gfx950).  Field layout is the reference's: Fortran order,
// i (streamwise) fastest, then j (span, periodic), then k (wall-normal),
// linear offset c = i + nX*(j + nY*k) with 0-based indices here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbl.h"

#define TBL_KMAX_LIMIT 64          // largest GMRES restart length supported
#define TBL_WAVE 64

namespace tbl {

// Grid geometry and the reference's coefficient set (dnssetbl.f90:183-247).
// 1-D arrays are indexed by 0-based k (nZ entries) and live in device memory.
struct Geo {
    int nX, nY, nZ;
    long sZ;                 // nX*nY (plane stride)
    long N;                  // nX*nY*nZ
    double dx, dy, dXi, invdx, invdy, invdXi, c12;
    double Aw, Ae, As, An, df1, df2;
    double dt, Re, Ue, omega;
    const double *x3, *Jac2, *JacInv, *Ab, *At, *Ap, *rAp;
    int scheme;              // TBL_SCHEME_*
    double forc[3];          // body force (TBL_SCHEME_DNSTBL)
    // wall-normal slabs of the reference's decomposition (per-slab orderings):
    // count, thickest slab in planes, 0-based first/last interior plane pairs
    int nslab, nkmax;
    const int *slab;
};

// Periodic span neighbours exactly as the reference's jm1/jm2/jp1/jp2 tables
// (dnssetbl.f90:234-243), for 0-based interior j in [1, nY-2].
__device__ __forceinline__ int jm1_of(int j, int nY) { return j == 1 ? nY - 2 : j - 1; }
__device__ __forceinline__ int jm2_of(int j, int nY) { return j == 2 ? nY - 2 : jm1_of(j, nY) - 1; }
__device__ __forceinline__ int jp1_of(int j, int nY) { return j == nY - 2 ? 1 : j + 1; }
__device__ __forceinline__ int jp2_of(int j, int nY) { return j == nY - 3 ? 1 : jp1_of(j, nY) + 1; }

// XCD-aware block order: the hardware hands consecutive workgroups to the 8
// XCDs round-robin; remapping makes each XCD own one contiguous run of
// logical tiles so that neighbouring tiles (stencil halos) share its L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned L, unsigned T) {
    const unsigned X = 8;
    unsigned q = T / X, r = T % X, x = L % X, idx = L / X;
    return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Raw buffer access: a resource (base, byte range) in scalar registers, a
// 32-bit vector offset, a scalar offset and an immediate; out-of-range loads
// return 0 and out-of-range stores are discarded.
#define BUF_DW3 0x00020000   // raw buffer resource word 3 for gfx9-family parts
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_of(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, BUF_DW3);
}
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double bld64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <class T> __device__ __forceinline__ T bldT(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(T) == 8) return bld64(r, voff, soff);
    else return bld(r, voff, soff);
}

// Plane loops over a register prefetch ring of R slots are unrolled by R so
// that each slot lives in fixed registers: IC<S> names the slot of a step,
// ring_steps runs steps S..R-1 of one ring cycle and ring_tail the steps of
// a partial last cycle.
template <int V> struct IC { static constexpr int value = V; };
template <int S, int R, class F> __device__ __forceinline__ void ring_steps(F &f, int k)
{
    f(IC<S>{}, k);
    if constexpr (S + 1 < R) ring_steps<S + 1, R>(f, k + 1);
}
template <int S, int R, class F> __device__ __forceinline__ void ring_tail(F &f, int k, int k1)
{
    if constexpr (S + 1 < R) {
        if (k < k1) {
            f(IC<S>{}, k);
            ring_tail<S + 1, R>(f, k + 1, k1);
        }
    }
}

// Grid of the reduction kernels: 64 x RT_J (i,j) tiles of RT_T threads, each
// block marching a chunk of k planes, sized so that ~8192 blocks (and as many
// partials) cover the domain (at 1024x256x192: 12-plane chunks).  Eight span
// rows per tile keep the span-halo re-reads of the stencil reductions at 1/4.
#ifndef RT_J
#define RT_J 8
#endif
#define RT_T (64 * RT_J)
struct RTiles {
    int gx, gy, kb, kchunk;
};
__host__ __device__ inline RTiles rtiles_of(int nX, int nY, int nZ) {
    RTiles r;
    r.gx = (nX - 1 + 63) / 64;         // column tiles start at i = 0 (cache-line aligned)
    r.gy = (nY - 2 + RT_J - 1) / RT_J;
    int kb = (8192 + r.gx * r.gy - 1) / (r.gx * r.gy);
    if (kb < 1) kb = 1;
    if (kb > nZ - 2) kb = nZ - 2;
    r.kchunk = (nZ - 2 + kb - 1) / kb;
    r.kb = (nZ - 2 + r.kchunk - 1) / r.kchunk;
    return r;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, TBL_WAVE);
    return v;
}

// Sum over a block of NT threads (linear thread id t); result valid in t == 0.
// Fixed shuffle/LDS order: deterministic.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh, int t) {
    v = wave_sum(v);
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < NT / 64; q++) r += sh[q];
    }
    __syncthreads();
    return r;
}

// GMRES control block, resident in device memory.  Arnoldi, Givens and the
// stopping decisions run on the device; the host polls only `st`.
struct GmStatus {
    int active;              // inner (Arnoldi) loop running
    int outer_done;          // solve finished
    int kit;                 // Arnoldi steps done in the current cycle
    int cycle, total, nan;
    int conv;                // stopped by the tolerance (not by mmax)
    int zero_rhs;            // true-residual mode: ||pRhs|| = 0, the solution is PCorr = 0
    double sumav, true_res, sumnor, rho;
};
struct GmCtrl {
    GmStatus st;
    double h[(TBL_KMAX_LIMIT + 1) * TBL_KMAX_LIMIT];   // h(i,j) at i + (KMAX+1)*j
    double cs[TBL_KMAX_LIMIT + 1], sn[TBL_KMAX_LIMIT + 1];
    double g[TBL_KMAX_LIMIT + 2], y[TBL_KMAX_LIMIT + 1];
    double sigma[TBL_KMAX_LIMIT + 1];                   // scale: v_i = sigma_i * V_i
    double coef[TBL_KMAX_LIMIT + 1];                    // per-step coefficients
    double bnorm;            // ||pRhs||_2 (true-residual mode)
    double inner_ref;        // first (preconditioned, or true in flexible mode) residual of the cycle
    double errtol, inner_tol;
    double cycle_tol;        // true-residual mode: reduction this cycle must reach
    int kmax, mode, mmax;
};
#define GM_H(c, i, j) ((c)->h[(i) + (TBL_KMAX_LIMIT + 1) * (j)])

enum { GM_MODE_REF = 0, GM_MODE_TRUE = 1 };

}  // namespace tbl
