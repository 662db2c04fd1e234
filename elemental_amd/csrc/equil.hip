// Ruiz equilibration of a dense matrix (El::RuizEquil, src/lapack_like/equilibrate/
// RuizEquil.cpp:23-60): four sweeps, each scaling the columns by their max norms and then the
// rows by theirs, every scale damped (DampScaling, RuizEquil.cpp:13-21: a zero norm gives 1,
// others at least eps^0.33 with eps = 2^-53), the scales accumulated into dCol and dRow, and
// every entry multiplied by the reciprocal of its scale (DiagonalSolve multiplies by 1/d).
// On return A holds diag(dRow)^{-1} A0 diag(dCol)^{-1}. Geometric equilibration (El::GeomEquil)
// is here too. The symmetric form
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

// Geometric equilibration (El::GeomEquil, src/lapack_like/equilibrate/GeomEquil.cpp:26-116):
// a column (row) with max |a| = mx > 0 and smallest nonzero |a| = mn is multiplied by 1/s,
// s = max(sqrt(mn * mx), sqrt(1/1000) * mx) (s = mx in the closing column pass), and s goes
// into dCol (dRow); all-zero columns and rows are left alone.
__device__ __forceinline__ double geom_scale(double mx, double mn, double sqrt_damp, int final_pass) {
    if (final_pass) return mx;
    const double prop = sqrt(mn * mx);
    const double floor_ = sqrt_damp * mx;
    return prop < floor_ ? floor_ : prop;
}

// max |v| and the smallest nonzero |v| of a workgroup (mn = +inf when there is none)
__device__ __forceinline__ void block_maxmin(double& mx, double& mn) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off));
        mn = fmin(mn, __shfl_xor(mn, off));
    }
    __shared__ double wx[ENT / 64], wn[ENT / 64];
    if ((threadIdx.x & 63) == 0) {
        wx[threadIdx.x >> 6] = mx;
        wn[threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    mx = wx[0];
    mn = wn[0];
    for (int w = 1; w < ENT / 64; ++w) {
        mx = fmax(mx, wx[w]);
        mn = fmin(mn, wn[w]);
    }
    __syncthreads();
}

// one workgroup per column: statistics, then the column scaled by 1/s
__global__ __launch_bounds__(ENT) void geom_cols_kernel(i64 m, double* __restrict__ A, i64 lda,
                                                        double* __restrict__ dCol, double sqrt_damp,
                                                        int final_pass) {
    const i64 j = blockIdx.x;
    double* col = A + j * lda;
    double mx = 0.0, mn = INFINITY;
    for (i64 i = threadIdx.x; i < m; i += ENT) {
        const double v = fabs(col[i]);
        mx = fmax(mx, v);
        if (v > 0.0) mn = fmin(mn, v);
    }
    block_maxmin(mx, mn);
    if (!(mx > 0.0)) return;
    const double sc = geom_scale(mx, mn, sqrt_damp, final_pass);
    const double inv = 1.0 / sc;
    for (i64 i = threadIdx.x; i < m; i += ENT) col[i] *= inv;
    if (threadIdx.x == 0) dCol[j] *= sc;
}

// partial row statistics over 64-column chunks: pmax[c][i], pmin[c][i]
__global__ __launch_bounds__(ENT) void geom_rowstats_kernel(i64 m, i64 n, const double* __restrict__ A,
                                                            i64 lda, double* __restrict__ pmax,
                                                            double* __restrict__ pmin) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    const i64 j0 = (i64)blockIdx.y * ECH;
    if (i >= m) return;
    const i64 j1 = j0 + ECH < n ? j0 + ECH : n;
    double mx = 0.0, mn = INFINITY;
    for (i64 j = j0; j < j1; ++j) {
        const double v = fabs(A[i + j * lda]);
        mx = fmax(mx, v);
        if (v > 0.0) mn = fmin(mn, v);
    }
    pmax[(i64)blockIdx.y * m + i] = mx;
    pmin[(i64)blockIdx.y * m + i] = mn;
}

// row scales from the partials: rinv = 1/s (1 for an all-zero row), dRow *= s
__global__ void geom_rowscale_kernel(i64 m, int nch, const double* __restrict__ pmax,
                                     const double* __restrict__ pmin, double sqrt_damp,
                                     double* __restrict__ dRow, double* __restrict__ rinv) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    if (i >= m) return;
    double mx = 0.0, mn = INFINITY;
    for (int c = 0; c < nch; ++c) {
        mx = fmax(mx, pmax[(i64)c * m + i]);
        mn = fmin(mn, pmin[(i64)c * m + i]);
    }
    if (mx > 0.0) {
        const double sc = geom_scale(mx, mn, sqrt_damp, 0);
        rinv[i] = 1.0 / sc;
        dRow[i] *= sc;
    } else {
        rinv[i] = 1.0;
    }
}

// A(i,j) *= rinv[i] (columns j0 ..), with the whole matrix's max and smallest nonzero |a|
// folded into out[0], out[1] (as ordered bit patterns: the values are non-negative)
__global__ __launch_bounds__(ENT) void geom_rows_apply_kernel(i64 m, double* __restrict__ A, i64 lda,
                                                              const double* __restrict__ rinv,
                                                              i64 j0, int scale,
                                                              unsigned long long* out) {
    const i64 i = (i64)blockIdx.x * ENT + threadIdx.x;
    const i64 j = j0 + blockIdx.y;
    double mx = 0.0, mn = INFINITY;
    if (i < m) {
        double* a = A + i + j * lda;
        double v = *a;
        if (scale) {
            v *= rinv[i];
            *a = v;
        }
        v = fabs(v);
        mx = v;
        if (v > 0.0) mn = v;
    }
    block_maxmin(mx, mn);
    if (threadIdx.x == 0) {
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
        atomicMin(out + 1, (unsigned long long)__double_as_longlong(mn));
    }
}

__global__ void geom_stats_init_kernel(unsigned long long* out) {
    out[0] = 0ull;
    out[1] = (unsigned long long)__double_as_longlong(INFINITY);
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

// El::GeomEquil: up to six sweeps (at least four when the ratio max/min-nonzero keeps falling
// by 10% or more), then every column normalised to max |a| = 1. The stopping test reads the
// ratio on the host after each sweep.
elx_status_t launch_geom_equil(hipStream_t s, i64 m, i64 n, double* A, i64 lda, double* dRow,
                               double* dCol, double* work) {
    if (m > 0) hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(m, ENT)), dim3(ENT), 0, s, m, dRow);
    if (n > 0) hipLaunchKernelGGL(ruiz_ones_kernel<double>, dim3(cdiv(n, ENT)), dim3(ENT), 0, s, n, dCol);
    ELX_HIP(hipGetLastError());
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    ELX_ARG(n < (1ll << 31) && (n + ECH - 1) / ECH < 65536);
    const int nch = (int)((n + ECH - 1) / ECH);
    double* rinv = work;              // m
    double* pmax = rinv + m;          // nch * m
    double* pmin = pmax + nch * m;    // nch * m
    unsigned long long* st = reinterpret_cast<unsigned long long*>(pmin + nch * m);  // 2
    const double sqrt_damp = std::sqrt(1.0 / 1000.0);
    // max |a| and smallest nonzero |a| of the whole matrix (optionally after a row scaling)
    auto stats = [&](int scale, double* mx, double* mn) -> elx_status_t {
        hipLaunchKernelGGL(geom_stats_init_kernel, dim3(1), dim3(1), 0, s, st);
        for (i64 j0 = 0; j0 < n; j0 += 65535) {
            const i64 nj = std::min<i64>(65535, n - j0);
            hipLaunchKernelGGL(geom_rows_apply_kernel, dim3(cdiv(m, ENT), (unsigned)nj), dim3(ENT), 0, s,
                               m, A, lda, rinv, j0, scale, st);
        }
        ELX_HIP(hipGetLastError());
        unsigned long long h[2];
        ELX_HIP(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, s));
        ELX_HIP(hipStreamSynchronize(s));
        *mx = __builtin_bit_cast(double, h[0]);
        *mn = __builtin_bit_cast(double, h[1]);
        return ELX_SUCCESS;
    };
    double mx = 0.0, mn = 0.0;
    ELX_CHECK(stats(0, &mx, &mn));
    if (mx == 0.0) return ELX_SUCCESS;
    double ratio = mx / std::min(mn, mx);
    for (int it = 0; it < 6; ++it) {
        hipLaunchKernelGGL(geom_cols_kernel, dim3((unsigned)n), dim3(ENT), 0, s, m, A, lda, dCol,
                           sqrt_damp, 0);
        hipLaunchKernelGGL(geom_rowstats_kernel, dim3(cdiv(m, ENT), (unsigned)nch), dim3(ENT), 0, s, m,
                           n, A, lda, pmax, pmin);
        hipLaunchKernelGGL(geom_rowscale_kernel, dim3(cdiv(m, ENT)), dim3(ENT), 0, s, m, nch, pmax, pmin,
                           sqrt_damp, dRow, rinv);
        ELX_CHECK(stats(1, &mx, &mn));
        const double new_ratio = mx / std::min(mn, mx);
        if (it >= 3 && new_ratio >= ratio * (9.0 / 10.0)) break;
        ratio = new_ratio;
    }
    hipLaunchKernelGGL(geom_cols_kernel, dim3((unsigned)n), dim3(ENT), 0, s, m, A, lda, dCol, sqrt_damp, 1);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

size_t geom_work_doubles(i64 m, i64 n) {
    const i64 nch = (n + ECH - 1) / ECH;
    return (size_t)(m + 2 * nch * m + 2);
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

elx_status_t elx_dgeom_equil(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                             double* dRow, double* dCol) {
    ELX_ARG(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1));
    ELX_ARG((m == 0 || dRow) && (n == 0 || dCol) && (m == 0 || n == 0 || dA));
    void* w = nullptr;
    ELX_CHECK(elx::scratch_get(elx::kSlotRuizWork,
                               sizeof(double) * std::max<size_t>(elx::geom_work_doubles(m, n), 1), &w));
    return elx::launch_geom_equil(static_cast<hipStream_t>(stream), m, n, dA, lda, dRow, dCol,
                                  static_cast<double*>(w));
}

}  // extern "C"
