// Ruiz equilibration of a dense matrix (El::RuizEquil, src/lapack_like/equilibrate/
// RuizEquil.cpp:23-60): four sweeps, each scaling the columns by their max norms and then the
// rows by theirs, every scale damped (DampScaling, RuizEquil.cpp:13-21: a zero norm gives 1,
// others at least eps^0.33 with eps = 2^-53), the scales accumulated into dCol and dRow, and
// every entry multiplied by the reciprocal of its scale (DiagonalSolve multiplies by 1/d).
// On return A holds diag(dRow)^{-1} A0 diag(dCol)^{-1}. The symmetric form
// (El::SymmetricRuizEquil) takes the square root of each damped column norm as the one scale
// of that row and column.
//
// The sweeps are HBM-bound passes over A. The row scaling of one sweep is fused with the
// column norms of the next, and the column scaling with the row norms, so a sweep reads and
// writes A twice; every entry still receives its scalings in the reference's order, and a
// max is exact, so the result is bitwise the reference's.
#include <cmath>

#include "elx_common.hpp"
#include "mixed.hpp"

namespace {

constexpr int ENT = 256;
constexpr int ECH = 64;  // columns per row-norm workgroup

// colmax[j] = max_i |A(i,j)| after A(i,j) *= rinv[i] (rinv null: no row scaling)
template <typename T>
__global__ __launch_bounds__(ENT) void ruiz_cols_kernel(i64 m, T* __restrict__ A, i64 lda,
                                                        const T* __restrict__ rinv,
                                                        T* __restrict__ colmax) {
    const i64 j = blockIdx.x;
    T* col = A + j * lda;
    T mx = T(0);
    for (i64 i = threadIdx.x; i < m; i += ENT) {
        T v = col[i];
        if (rinv) {
            v *= rinv[i];
            col[i] = v;
        }
        mx = fmax(mx, fabs(v));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    __shared__ T wm[ENT / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = wm[0];
        for (int w = 1; w < ENT / 64; ++w) r = fmax(r, wm[w]);
        colmax[j] = r;
    }
}

// part[c][i] = max over this chunk's columns j of |A(i,j)| after A(i,j) *= cinv[j]
template <typename T>
__global__ __launch_bounds__(ENT) void ruiz_rows_kernel(i64 m, i64 n, T* __restrict__ A, i64 lda,
                                                        const T* __restrict__ cinv,
                                                        T* __restrict__ part) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    const i64 j0 = (i64)blockIdx.y * ECH;
    if (i >= m) return;
    const i64 j1 = j0 + ECH < n ? j0 + ECH : n;
    T mx = T(0);
    for (i64 j = j0; j < j1; ++j) {
        T* a = A + i + j * lda;
        const T v = *a * cinv[j];
        *a = v;
        mx = fmax(mx, fabs(v));
    }
    part[(i64)blockIdx.y * m + i] = mx;
}

// s = max over the nparts partial norms; s = 1 if s == 0, else max(s, tol); d *= s, inv = 1/s
template <typename T>
__global__ void ruiz_damp_kernel(i64 len, const T* __restrict__ part, int nparts, T tol,
                                 T* __restrict__ d, T* __restrict__ inv) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    if (i >= len) return;
    T s = part[i];
    for (int c = 1; c < nparts; ++c) s = fmax(s, part[(i64)c * len + i]);
    s = (s == T(0)) ? T(1) : (s < tol ? tol : s);
    d[i] *= s;
    inv[i] = T(1) / s;
}

template <typename T>
__global__ void ruiz_rowscale_kernel(i64 m, i64 n, T* __restrict__ A, i64 lda,
                                     const T* __restrict__ rinv) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i >= m) return;
    A[i + j * lda] *= rinv[i];
}

// Symmetric sweeps (El::SymmetricRuizEquil, src/lapack_like/equilibrate/SymmetricRuizEquil.cpp:
// 29-55): colmax[j] = max_i |A(i,j)| after the previous sweep's scaling, applied as the
// reference does it, right then left: A(i,j) = (A(i,j) * sinv[j]) * sinv[i].
__global__ __launch_bounds__(ENT) void ruiz_sym_cols_kernel(i64 n, double* __restrict__ A, i64 lda,
                                                            const double* __restrict__ sinv,
                                                            double* __restrict__ colmax) {
    const i64 j = blockIdx.x;
    double* col = A + j * lda;
    const double sj = sinv ? sinv[j] : 1.0;
    double mx = 0.0;
    for (i64 i = threadIdx.x; i < n; i += ENT) {
        double v = col[i];
        if (sinv) {
            v = v * sj;
            v = v * sinv[i];
            col[i] = v;
        }
        mx = fmax(mx, fabs(v));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    __shared__ double wm[ENT / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = wm[0];
        for (int w = 1; w < ENT / 64; ++w) r = fmax(r, wm[w]);
        colmax[j] = r;
    }
}

// s = sqrt(damp(colmax)); d *= s; sinv = 1/s
__global__ void ruiz_sym_damp_kernel(i64 n, const double* __restrict__ colmax, double tol,
                                     double* __restrict__ d, double* __restrict__ sinv) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    if (i >= n) return;
    double s = colmax[i];
    s = (s == 0.0) ? 1.0 : (s < tol ? tol : s);
    s = sqrt(s);
    d[i] *= s;
    sinv[i] = 1.0 / s;
}

__global__ void ruiz_sym_scale_kernel(i64 n, double* __restrict__ A, i64 lda,
                                      const double* __restrict__ sinv, i64 j0) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    const i64 j = j0 + blockIdx.y;
    if (i >= n) return;
    double* a = A + i + j * lda;
    *a = (*a * sinv[j]) * sinv[i];
}

template <typename T>
__global__ void ruiz_ones_kernel(i64 len, T* __restrict__ d) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    if (i < len) d[i] = T(1);
}

}  // namespace

namespace elx {

double ruiz_tolerance() { return std::pow(std::ldexp(1.0, -53), 0.33); }

elx_status_t launch_ruiz_equil(hipStream_t s, i64 m, i64 n, double* A, i64 lda, double* dRow,
                               double* dCol, double* work) {
    if (m <= 0 || n <= 0) {
        if (m > 0) hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(m, ENT)), dim3(ENT), 0, s, m, dRow);
        if (n > 0) hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, dCol);
        ELX_HIP(hipGetLastError());
        return ELX_SUCCESS;
    }
    ELX_ARG(n < (1ll << 31) && (n + ECH - 1) / ECH < 65536);
    const double tol = ruiz_tolerance();
    const int nch = (int)((n + ECH - 1) / ECH);
    double* colmax = work;        // n
    double* cinv = colmax + n;    // n
    double* rinv = cinv + n;      // m
    double* part = rinv + m;      // nch * m
    hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(m, ENT)), dim3(ENT), 0, s, m, dRow);
    hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, dCol);
    for (int it = 0; it < 4; ++it) {
        hipLaunchKernelGGL(ruiz_cols_kernel<double>, dim3((unsigned)n), dim3(ENT), 0, s, m, A, lda,
                           it ? rinv : nullptr, colmax);
        hipLaunchKernelGGL(ruiz_damp_kernel<double>, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, colmax, 1,
                           tol, dCol, cinv);
        hipLaunchKernelGGL(ruiz_rows_kernel<double>, dim3(cdiv(m, ENT), (unsigned)nch), dim3(ENT), 0, s,
                           m, n, A, lda, cinv, part);
        hipLaunchKernelGGL(ruiz_damp_kernel<double>, dim3(cdiv(m, ENT)), dim3(ENT), 0, s, m, part, nch,
                           tol, dRow, rinv);
    }
    for (i64 j0 = 0; j0 < n; j0 += 65535) {
        const i64 nj = std::min<i64>(65535, n - j0);
        hipLaunchKernelGGL(ruiz_rowscale_kernel<double>, dim3(cdiv(m, ENT), (unsigned)nj), dim3(ENT), 0,
                           s, m, nj, A + j0 * lda, lda, rinv);
    }
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_symmetric_ruiz_equil(hipStream_t s, i64 n, double* A, i64 lda, double* d,
                                         int maxiter, double* work) {
    if (n <= 0) return ELX_SUCCESS;
    ELX_ARG(n < (1ll << 31));
    const double tol = ruiz_tolerance();
    double* colmax = work;     // n
    double* sinv = work + n;   // n
    hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, d);
    for (int it = 0; it < maxiter; ++it) {
        hipLaunchKernelGGL(ruiz_sym_cols_kernel, dim3((unsigned)n), dim3(ENT), 0, s, n, A, lda,
                           it ? sinv : nullptr, colmax);
        hipLaunchKernelGGL(ruiz_sym_damp_kernel, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, colmax, tol,
                           d, sinv);
    }
    if (maxiter > 0)
        for (i64 j0 = 0; j0 < n; j0 += 65535) {
            const i64 nj = std::min<i64>(65535, n - j0);
            hipLaunchKernelGGL(ruiz_sym_scale_kernel, dim3(cdiv(n, ENT), (unsigned)nj), dim3(ENT), 0, s,
                               n, A, lda, sinv, j0);
        }
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

namespace {
__global__ void reciprocal_kernel(i64 n, const double* __restrict__ in, double* __restrict__ out) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    if (i < n) out[i] = 1.0 / in[i];
}
}  // namespace

elx_status_t launch_reciprocal(hipStream_t s, i64 n, const double* in, double* out) {
    if (n <= 0) return ELX_SUCCESS;
    hipLaunchKernelGGL(reciprocal_kernel, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, in, out);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

size_t ruiz_work_doubles(i64 m, i64 n) {
    const i64 nch = (n + ECH - 1) / ECH;
    return (size_t)(2 * n + m + nch * m);
}

}  // namespace elx

extern "C" {

elx_status_t elx_druiz_equil(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                             double* dRow, double* dCol) {
    ELX_ARG(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1));
    ELX_ARG((m == 0 || dRow) && (n == 0 || dCol) && (m == 0 || n == 0 || dA));
    void* w = nullptr;
    ELX_CHECK(elx::scratch_get(elx::kSlotRuizWork,
                               sizeof(double) * std::max<size_t>(elx::ruiz_work_doubles(m, n), 1), &w));
    return elx::launch_ruiz_equil(static_cast<hipStream_t>(stream), m, n, dA, lda, dRow, dCol,
                                  static_cast<double*>(w));
}

elx_status_t elx_dsymmetric_ruiz_equil(void* stream, int64_t n, double* dA, int64_t lda, double* dD,
                                       int32_t maxiter) {
    ELX_ARG(n >= 0 && lda >= std::max<int64_t>(n, 1) && maxiter >= 0);
    ELX_ARG(n == 0 || (dA && dD));
    void* w = nullptr;
    ELX_CHECK(elx::scratch_get(elx::kSlotRuizWork, sizeof(double) * (size_t)std::max<int64_t>(2 * n, 1),
                               &w));
    return elx::launch_symmetric_ruiz_equil(static_cast<hipStream_t>(stream), n, dA, lda, dD, maxiter,
                                            static_cast<double*>(w));
}

}  // extern "C"
