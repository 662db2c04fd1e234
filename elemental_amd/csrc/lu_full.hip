// LU with full (complete) pivoting: El::LU(Matrix<F>&, Permutation& P, Permutation& Q), i.e.
// lu::Full (src/lapack_like/factor/LU/Full.hpp:15-58), and its SolveAfter
// (src/lapack_like/factor/LU/SolveAfter.hpp:123-151).
//
// Per step k the reference finds MaxAbsLoc of the trailing block A(k:,k:) -- the first entry
// of largest magnitude in column-major order (src/blas_like/level1/MaxAbsLoc.cpp:131-163,
// strict '>' from a start of 0 at the top-left entry) -- swaps row k with the pivot's row and
// column k with its column over the whole matrix, scales the column below the pivot by its
// reciprocal and applies the rank-1 update (dger: columns whose multiplier-row entry is zero
// are skipped). On the GPU every step is two launches:
//   update_argmax  the rank-1 update of step k-1 over the trailing block, fused with each
//                  workgroup's candidate for step k: the block is read and written once per
//                  step;
//   pivot_swap     one workgroup reduces the candidates (largest magnitude, ties to the
//                  smallest column-major index, NaN never chosen), records the swaps, moves
//                  row k and column k, and scales the column below the pivot.
// An exactly zero pivot records info = k+1 after the swaps, as the reference leaves A when it
// throws; the remaining launches then change nothing.
#include <algorithm>
#include <cmath>

#include "elx_common.hpp"

namespace {

constexpr int FT = 256;     // threads per workgroup
constexpr int FCOLS = 8;    // trailing-block columns per workgroup (one row per thread)

struct Cand {
    double v;   // |a|, -1 for NaN
    long long idx;  // column-major index in the trailing block
};

__device__ __forceinline__ bool cand_better(double v, long long i, double bv, long long bi) {
    return v > bv || (v == bv && i < bi);
}

__device__ __forceinline__ void wave_reduce(double& v, long long& i) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const long long oi = __shfl_xor(i, off);
        if (cand_better(ov, oi, v, i)) {
            v = ov;
            i = oi;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(FT) void update_argmax_kernel(T* A, i64 m, i64 n, i64 lda, i64 k,
                                                           const int32_t* __restrict__ info,
                                                           Cand* part) {
    __shared__ double sv[FT / 64];
    __shared__ long long si[FT / 64];
    if (*info != 0) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const i64 rows = m - k, cols = n - k;
    const i64 i = k + (i64)blockIdx.x * FT + t;  // this thread's row
    const i64 j0 = k + (i64)blockIdx.y * FCOLS;
    double bv = -1.0;
    long long bi = LLONG_MAX;
    if (i < m) {
        // finish step k-1: the rank-1 update with row i's multiplier (column k-1 was scaled by
        // the pivot launch of step k-1)
        const bool upd = k > 0;
        const T l = upd ? A[i + (k - 1) * lda] : T(0);
        for (int c = 0; c < FCOLS; ++c) {
            const i64 j = j0 + c;
            if (j >= n) break;
            T a = A[i + j * lda];
            if (upd) {
                const T u = A[(k - 1) + j * lda];
                if (u != T(0)) {
                    a = __builtin_fma(l, -u, a);
                    A[i + j * lda] = a;
                }
            }
            const double v = (a != a) ? -1.0 : (double)fabs(a);
            const long long idx = (long long)(j - k) * rows + (i - k);
            if (cand_better(v, idx, bv, bi)) {
                bv = v;
                bi = idx;
            }
        }
    }
    wave_reduce(bv, bi);
    if (lane == 0) {
        sv[wave] = bv;
        si[wave] = bi;
    }
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < FT / 64; ++w)
            if (cand_better(sv[w], si[w], bv, bi)) {
                bv = sv[w];
                bi = si[w];
            }
        part[blockIdx.y * gridDim.x + blockIdx.x] = Cand{bv, bi};
    }
    (void)cols;
}

template <typename T>
__global__ __launch_bounds__(1024) void pivot_swap_kernel(T* A, i64 m, i64 n, i64 lda, i64 k,
                                                          const Cand* __restrict__ part,
                                                          int nparts, int32_t* p, int32_t* q,
                                                          int32_t* info) {
    __shared__ double sv[16];
    __shared__ long long si[16];
    __shared__ long long s_i, s_j;
    if (*info != 0) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double bv = -1.0;
    long long bi = LLONG_MAX;
    for (int e = t; e < nparts; e += blockDim.x) {
        const Cand c = part[e];
        if (cand_better(c.v, c.idx, bv, bi)) {
            bv = c.v;
            bi = c.idx;
        }
    }
    wave_reduce(bv, bi);
    if (lane == 0) {
        sv[wave] = bv;
        si[wave] = bi;
    }
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); ++w)
            if (cand_better(sv[w], si[w], bv, bi)) {
                bv = sv[w];
                bi = si[w];
            }
        // MaxAbsLoc starts from the top-left entry with value 0: an all-zero (or all-NaN) block
        // gives index 0
        if (!(bv > 0.0)) bi = 0;
        const i64 rows = m - k;
        s_i = k + bi % rows;
        s_j = k + bi / rows;
        p[k] = (int32_t)s_i;
        q[k] = (int32_t)s_j;
    }
    __syncthreads();
    const i64 ip = s_i, jp = s_j;
    // RowSwap(A, k, iPiv) over all columns, then ColSwap(A, k, jPiv) over all rows
    if (ip != k)
        for (i64 j = t; j < n; j += blockDim.x) {
            const T a = A[k + j * lda], b = A[ip + j * lda];
            A[k + j * lda] = b;
            A[ip + j * lda] = a;
        }
    __syncthreads();
    if (jp != k)
        for (i64 i = t; i < m; i += blockDim.x) {
            const T a = A[i + k * lda], b = A[i + jp * lda];
            A[i + k * lda] = b;
            A[i + jp * lda] = a;
        }
    __syncthreads();
    // a21 *= 1/alpha (blas::Scal); an exactly zero pivot stops here, as the reference throws
    const T alpha = A[k + k * lda];
    if (alpha == T(0)) {
        if (t == 0) atomicCAS(info, 0, (int32_t)(k + 1));
        return;
    }
    const T inv = T(1) / alpha;
    for (i64 i = k + 1 + t; i < m; i += blockDim.x) A[i + k * lda] = inv * A[i + k * lda];
}

template <typename T>
elx_status_t getrf_full_impl(hipStream_t s, i64 m, i64 n, T* A, i64 lda, int32_t* p, int32_t* q,
                             int32_t* info) {
    const i64 minDim = std::min(m, n);
    ELX_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), s));
    if (minDim == 0) return ELX_SUCCESS;
    const i64 gx0 = (m + FT - 1) / FT, gy0 = (n + FCOLS - 1) / FCOLS;
    ELX_ARG(gy0 <= 65535);
    void* pv = nullptr;
    ELX_CHECK(elx::scratch_get(elx::kSlotFullPiv, sizeof(Cand) * (size_t)(gx0 * gy0), &pv));
    Cand* part = static_cast<Cand*>(pv);
    for (i64 k = 0; k < minDim; ++k) {
        const i64 rows = m - k, cols = n - k;
        const unsigned gx = (unsigned)((rows + FT - 1) / FT), gy = (unsigned)((cols + FCOLS - 1) / FCOLS);
        hipLaunchKernelGGL(update_argmax_kernel<T>, dim3(gx, gy), dim3(FT), 0, s, A, m, n, lda, k,
                           info, part);
        ELX_HIP(hipGetLastError());
        hipLaunchKernelGGL(pivot_swap_kernel<T>, dim3(1), dim3(1024), 0, s, A, m, n, lda, k, part,
                           (int)(gx * gy), p, q, info);
        ELX_HIP(hipGetLastError());
    }
    // after the last pivot nothing is left to update: its column was scaled by its pivot launch
    // and no column lies to the right of it (m >= n) or no row below it (m < n)
    return ELX_SUCCESS;
}

}  // namespace

namespace elx {

elx_status_t getrf_full(hipStream_t s, i64 m, i64 n, double* A, i64 lda, int32_t* p, int32_t* q,
                        int32_t* info) {
    return getrf_full_impl<double>(s, m, n, A, lda, p, q, info);
}
elx_status_t getrf_full(hipStream_t s, i64 m, i64 n, float* A, i64 lda, int32_t* p, int32_t* q,
                        int32_t* info) {
    return getrf_full_impl<float>(s, m, n, A, lda, p, q, info);
}

}  // namespace elx
