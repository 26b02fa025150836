// tbl_fastpc.hip -- fast-diagonalisation preconditioner for the pressure
// Poisson operator of dnssetbl/dnssetbl.f90:574-583.
//
// The reference's operator is separable:
//   A = Ox (x) I (x) I  +  I (x) Oy (x) I  +  Z,
// Ox: Aw/Ae = 1/dx^2 couplings with homogeneous Dirichlet ends (i = 1, nX),
// Oy: As/An = 1/dy^2 couplings, periodic with period nY-2 (jm1/jp1 tables),
// Z : the k-dependent tridiagonal Ab(k), Ap(k), At(k) (Dirichlet at k = 1, nZ).
// Ox is diagonalised by the orthonormal DST-I (eigenvalues 2 Aw cos(pi m/(nX-1))),
// Oy by the real orthonormal Fourier basis (eigenvalues 2 An cos(2 pi q/(nY-2))),
// so A^{-1} r = Sx Fy^T-transform, one tridiagonal solve in k per (m, q) mode
// with diagonal Ap(k) + mu_x(m) + mu_y(q), and the inverse transforms.
// The four transforms are plain strided-batched GEMMs (rocBLAS, matrix cores);
// the wall-normal solves are a hand-written Thomas kernel, one thread per mode
// column, coalesced across modes.  Used as M^{-1} inside GMRES, which keeps the
// reference's solver and stopping rule while needing only a few iterations.
#include <rocblas/rocblas.h>

#include <math.h>
#include <string.h>

#include <vector>

#include "tbl_dev.h"
#include "tbl_launch.h"

namespace tbl {

struct FastPC {
    rocblas_handle h = nullptr;
    int n = 0, L = 0;                 // interior sizes nX-2, nY-2
    double *Sx = nullptr, *Fy = nullptr, *mux = nullptr, *muy = nullptr;
    float *Sxf = nullptr, *Fyf = nullptr;
    void *t1 = nullptr, *t2 = nullptr, *cp = nullptr;   // work fields (storage type)
    bool single = false;
};

static int setfail(char *err, int len, const char *m)
{
    if (err && len > 0) snprintf(err, len, "%s", m);
    return 1;
}

int fastpc_create(const Geo &g, const double *Ap, hipStream_t s, bool single, FastPC **out, char *err, int errlen)
{
    (void)Ap;
    FastPC *f = new FastPC();
    f->single = single;
    f->n = g.nX - 2;
    f->L = g.nY - 2;
    const int n = f->n, L = f->L;
    // DST-I, orthonormal and symmetric: Sx[i,m] = sqrt(2/(n+1)) sin(pi (i+1)(m+1)/(n+1))
    std::vector<double> Sx((size_t)n * n), Fy((size_t)L * L), mux(g.nX, 0.0), muy(g.nY, 0.0);
    const double sx = sqrt(2.0 / (n + 1));
    for (int m = 0; m < n; m++)
        for (int i = 0; i < n; i++) Sx[i + (size_t)n * m] = sx * sin(M_PI * (double)((long)(i + 1) * (m + 1) % (2 * (n + 1))) / (n + 1));
    for (int m = 0; m < n; m++) mux[m + 1] = 2.0 * g.Aw * cos(M_PI * (m + 1) / (n + 1));
    // real Fourier basis over the period L (L even): column 0 constant, columns
    // 2q-1 / 2q cos / sin of wavenumber q, last column the Nyquist mode
    for (int nn = 0; nn < L; nn++) {
        int q = (nn == 0) ? 0 : (nn == L - 1 ? L / 2 : (nn + 1) / 2);
        for (int j = 0; j < L; j++) {
            double v;
            const double th = 2.0 * M_PI * (double)((long)q * j % L) / L;
            if (nn == 0) v = 1.0 / sqrt((double)L);
            else if (nn == L - 1) v = ((j & 1) ? -1.0 : 1.0) / sqrt((double)L);
            else if (nn & 1) v = sqrt(2.0 / L) * cos(th);
            else v = sqrt(2.0 / L) * sin(th);
            Fy[j + (size_t)L * nn] = v;
        }
        muy[nn + 1] = 2.0 * g.An * cos(2.0 * M_PI * q / L);
    }
    const size_t es = single ? sizeof(float) : sizeof(double);
    hipError_t e = hipSuccess;
    e = (hipError_t)(e | hipMalloc((void **)&f->Sx, sizeof(double) * Sx.size()));
    e = (hipError_t)(e | hipMalloc((void **)&f->Fy, sizeof(double) * Fy.size()));
    e = (hipError_t)(e | hipMalloc((void **)&f->mux, sizeof(double) * g.nX));
    e = (hipError_t)(e | hipMalloc((void **)&f->muy, sizeof(double) * g.nY));
    e = (hipError_t)(e | hipMalloc((void **)&f->Sxf, sizeof(float) * Sx.size()));
    e = (hipError_t)(e | hipMalloc((void **)&f->Fyf, sizeof(float) * Fy.size()));
    e = (hipError_t)(e | hipMalloc(&f->t1, es * g.N));
    e = (hipError_t)(e | hipMalloc(&f->t2, es * g.N));
    e = (hipError_t)(e | hipMalloc(&f->cp, es * g.N));
    if (e != hipSuccess) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: hipMalloc failed"); }
    std::vector<float> Sxf(Sx.begin(), Sx.end()), Fyf(Fy.begin(), Fy.end());
    hipMemcpyAsync(f->Sx, Sx.data(), sizeof(double) * Sx.size(), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(f->Fy, Fy.data(), sizeof(double) * Fy.size(), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(f->mux, mux.data(), sizeof(double) * g.nX, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(f->muy, muy.data(), sizeof(double) * g.nY, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(f->Sxf, Sxf.data(), sizeof(float) * Sxf.size(), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(f->Fyf, Fyf.data(), sizeof(float) * Fyf.size(), hipMemcpyHostToDevice, s);
    hipMemsetAsync(f->t1, 0, es * g.N, s);
    hipMemsetAsync(f->t2, 0, es * g.N, s);
    hipMemsetAsync(f->cp, 0, es * g.N, s);
    if (hipStreamSynchronize(s) != hipSuccess) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: upload failed"); }
    if (rocblas_create_handle(&f->h) != rocblas_status_success) { fastpc_destroy(f); return setfail(err, errlen, "fastpc: rocblas_create_handle failed"); }
    rocblas_set_stream(f->h, s);
    rocblas_set_pointer_mode(f->h, rocblas_pointer_mode_host);
    *out = f;
    return 0;
}

void fastpc_destroy(FastPC *f)
{
    if (!f) return;
    if (f->h) rocblas_destroy_handle(f->h);
    void *p[] = {f->Sx, f->Fy, f->mux, f->muy, f->Sxf, f->Fyf, f->t1, f->t2, f->cp};
    for (void *q : p)
        if (q) hipFree(q);
    delete f;
}

// wall-normal tridiagonal solve of every (m, q) mode in place:
// Ab(k) p(k-1) + (Ap(k) + mu_x(m) + mu_y(q)) p(k) + At(k) p(k+1) = d(k), p(0)=p(nZ-1)=0
template <class T>
__global__ __launch_bounds__(256) void k_thomas_modes(Geo g, T *__restrict__ d, T *__restrict__ cp,
                                                     const double *__restrict__ mux, const double *__restrict__ muy,
                                                     const int *__restrict__ active)
{
    if (active && !*active) return;
    const int nX = g.nX, nY = g.nY, nZ = g.nZ;
    const int t = threadIdx.x;
    const int i = blockIdx.x * 64 + (t & 63) + 1;
    const int j = blockIdx.y * 4 + (t >> 6) + 1;
    if (i > nX - 2 || j > nY - 2) return;
    const long sk = g.sZ;
    const long c0 = (long)i + (long)nX * j;
    const T shift = T(mux[i] + muy[j]);
    T cprev = T(0), dprev = T(0);
    for (int k = 1; k <= nZ - 2; k++) {
        const long c = c0 + sk * k;
        const T a = T(g.Ab[k]), b = T(g.Ap[k]) + shift, cc = T(g.At[k]);
        const T den = b - a * cprev;
        cprev = cc / den;
        dprev = (d[c] - a * dprev) / den;
        cp[c] = cprev;
        d[c] = dprev;
    }
    T pn = d[c0 + sk * (nZ - 2)];
    for (int k = nZ - 3; k >= 1; k--) {
        const long c = c0 + sk * k;
        pn = d[c] - cp[c] * pn;
        d[c] = pn;
    }
}

static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation tb, int m, int n, int k, const double *A, int lda,
                              long sa, const double *B, int ldb, long sb, double *C, int ldc, long sc, int batch)
{
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm_strided_batched(h, rocblas_operation_none, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &zero,
                                         C, ldc, sc, batch);
}
static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation tb, int m, int n, int k, const float *A, int lda,
                              long sa, const float *B, int ldb, long sb, float *C, int ldc, long sc, int batch)
{
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm_strided_batched(h, rocblas_operation_none, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &zero,
                                         C, ldc, sc, batch);
}
template <class T> static const T *mSx(const FastPC *f);
template <> const double *mSx<double>(const FastPC *f) { return f->Sx; }
template <> const float *mSx<float>(const FastPC *f) { return f->Sxf; }
template <class T> static const T *mFy(const FastPC *f);
template <> const double *mFy<double>(const FastPC *f) { return f->Fy; }
template <> const float *mFy<float>(const FastPC *f) { return f->Fyf; }

// out = M^{-1} r on the interior (boundary entries of out are not written)
template <class T>
int run_fastpc(const FastPC *f, const Geo &g, const T *r, T *out, const int *active, hipStream_t s)
{
    const int n = f->n, L = f->L, nX = g.nX;
    const long pst = g.sZ;                               // plane stride
    const long off = 1 + (long)nX + pst;                 // (i,j,k) = (1,1,1)
    const int nb = g.nZ - 2;
    T *t1 = (T *)f->t1, *t2 = (T *)f->t2;
    const T *Sx = mSx<T>(f), *Fy = mFy<T>(f);
    rocblas_status st = rocblas_status_success;
    // forward x: t1_k = Sx * r_k  (rows i, columns j of each interior plane)
    st = (rocblas_status)(st | gemm_sb(f->h, rocblas_operation_none, n, L, n, Sx, n, 0, r + off, nX, pst, t1 + off, nX, pst, nb));
    // forward y: t2_k = t1_k * Fy
    st = (rocblas_status)(st | gemm_sb(f->h, rocblas_operation_none, n, L, L, t1 + off, nX, pst, Fy, L, 0, t2 + off, nX, pst, nb));
    hipLaunchKernelGGL(k_thomas_modes<T>, dim3((n + 63) / 64, (L + 3) / 4), dim3(256), 0, s, g, t2, (T *)f->cp,
                       (const double *)f->mux, (const double *)f->muy, active);
    // inverse y: t1_k = t2_k * Fy^T ; inverse x: out_k = Sx * t1_k
    st = (rocblas_status)(st | gemm_sb(f->h, rocblas_operation_transpose, n, L, L, t2 + off, nX, pst, Fy, L, 0, t1 + off, nX, pst, nb));
    st = (rocblas_status)(st | gemm_sb(f->h, rocblas_operation_none, n, L, n, Sx, n, 0, t1 + off, nX, pst, out + off, nX, pst, nb));
    return st == rocblas_status_success ? 0 : 1;
}

template int run_fastpc<double>(const FastPC *, const Geo &, const double *, double *, const int *, hipStream_t);
template int run_fastpc<float>(const FastPC *, const Geo &, const float *, float *, const int *, hipStream_t);

}  // namespace tbl
