// FP64 GEMM / SYRK on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// C := alpha op(A) op(B) + beta C for column-major operands. This is the trailing update
// of the blocked LU (Gemm(NORMAL,NORMAL,-1,A21,A12,1,A22), src/lapack_like/factor/LU.cpp:140)
// and, with tri != 0, the Herk/Syrk update of the blocked Cholesky
// (src/lapack_like/factor/Cholesky/LowerVariant3.hpp:66, UpperVariant3.hpp:71) where only one
// triangle of C is touched.
//
// Tiling: a 256-thread workgroup owns a 128x128 tile of C; its four waves each own a 64x64
// sub-tile held as 4x4 MFMA accumulators (64 doubles per lane). K advances 16 at a time
// through LDS, with the next K-slab prefetched into registers while the current one feeds
// the MFMAs. The MFMA operands are fed as (op(B)^T, op(A)^T) so that each lane's four
// accumulator values are one column-contiguous run of C: stores go out as 128-byte rows.
#include <algorithm>

#include "elx_common.hpp"

namespace {

constexpr int BM = 128, BN = 128, BK = 16, PAD = 4, NT = 256;
typedef double d4 __attribute__((ext_vector_type(4)));

// Loads of one BK-slab of op(A) (BM x BK) or op(B) (BK x BN) into 8 registers per thread.
// "mn-contiguous" operands (op(A) = A, op(B) = B^T) have 128 contiguous doubles per k;
// "k-contiguous" operands (op(A) = A^T, op(B) = B) have 16 contiguous doubles per m/n.
template <bool KCONTIG>
__device__ __forceinline__ void load_slab(const double* __restrict__ P, i64 ld, i64 mn0,
                                          i64 MN, i64 k0, i64 K, double (&r)[8]) {
    const int t = threadIdx.x;
    if (!KCONTIG) {
        const int kk = t >> 4;          // 0..15
        const int ii = (t & 15) * 8;    // 0..120
        const i64 k = k0 + kk;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const i64 i = mn0 + ii + q;
            r[q] = (k < K && i < MN) ? P[i + k * ld] : 0.0;
        }
    } else {
        const int ii = t >> 1;          // 0..127
        const int kk = (t & 1) * 8;     // 0 or 8
        const i64 i = mn0 + ii;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const i64 k = k0 + kk + q;
            r[q] = (k < K && i < MN) ? P[k + i * ld] : 0.0;
        }
    }
}

template <bool KCONTIG>
__device__ __forceinline__ void store_slab(double (*S)[BM + PAD], const double (&r)[8],
                                           double sgn = 1.0) {
    const int t = threadIdx.x;
    if (!KCONTIG) {
        const int kk = t >> 4, ii = (t & 15) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) S[kk][ii + q] = sgn * r[q];
    } else {
        const int ii = t >> 1, kk = (t & 1) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) S[kk + q][ii] = sgn * r[q];
    }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(NT, 2) void dgemm_kernel(i64 M, i64 N, i64 K, double alpha,
                                                   const double* __restrict__ A, i64 lda,
                                                   const double* __restrict__ B, i64 ldb,
                                                   double beta, double* __restrict__ C,
                                                   i64 ldc, int tri, int tiles_m, int tiles_n) {
    __shared__ double As[BK][BM + PAD];  // As[k][i] = op(A)(m0+i, k0+k)
    __shared__ double Bs[BK][BN + PAD];  // Bs[k][j] = op(B)(k0+k, n0+j)

    // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous run of
    // tiles so that tiles sharing A rows / B columns share that XCD's L2.
    const int total = tiles_m * tiles_n;
    const int per_xcd = (total + 7) / 8;
    const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (lin >= total) return;
    // Grouped ordering: walk 8 tile-rows at a time down the columns.
    constexpr int GROUP = 8;
    const int group = lin / (GROUP * tiles_n);
    const int first_m = group * GROUP;
    const int gsize = min(tiles_m - first_m, GROUP);
    const int tm = first_m + (lin % (GROUP * tiles_n)) % gsize;
    const int tn = (lin % (GROUP * tiles_n)) / gsize;
    const i64 m0 = (i64)tm * BM, n0 = (i64)tn * BN;
    if (tri == 1 && n0 > m0 + BM - 1) return;  // tile strictly above the diagonal
    if (tri == 2 && m0 > n0 + BN - 1) return;  // tile strictly below the diagonal

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    const int r16 = lane & 15, kq = lane >> 4;

    // With beta == 1 and alpha == +-1 (every update of the factorizations) the accumulators
    // start from C and alpha is folded into the A operand (negation is exact): the C tile is
    // read while the first slabs load, and the epilogue is a plain store.
    const bool fold = (beta == 1.0) && (alpha == 1.0 || alpha == -1.0);
    const double sgn = fold ? alpha : 1.0;
    d4 acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const i64 m = m0 + wm + mi * 16 + r16;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + ni * 16 + kq + 4 * r;
                const bool ok = fold && m < M && n < N && !(tri == 1 && m < n) && !(tri == 2 && m > n);
                acc[mi][ni][r] = ok ? C[m + n * ldc] : 0.0;
            }
    }

    double ra[8], rb[8];
    load_slab<TA>(A, lda, m0, M, 0, K, ra);
    load_slab<!TB>(B, ldb, n0, N, 0, K, rb);
    for (i64 k0 = 0; k0 < K; k0 += BK) {
        store_slab<TA>(As, ra, sgn);
        store_slab<!TB>(Bs, rb);
        __syncthreads();
        if (k0 + BK < K) {
            load_slab<TA>(A, lda, m0, M, k0 + BK, K, ra);
            load_slab<!TB>(B, ldb, n0, N, k0 + BK, K, rb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[4], bf[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                af[q] = As[kk + kq][wm + q * 16 + r16];
                bf[q] = Bs[kk + kq][wn + q * 16 + r16];
            }
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    acc[mi][ni] =
                        __builtin_amdgcn_mfma_f64_16x16x4f64(bf[ni], af[mi], acc[mi][ni], 0, 0, 0);
        }
        __syncthreads();
    }

    // v_mfma_f64_16x16x4f64 leaves D(i,j) with i = lane/16 + 4r, j = lane%16 in register r,
    // so the lane holds C(m0+wm+mi*16+r16, n0+wn+ni*16+kq+4r), r = 0..3.
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const i64 m = m0 + wm + mi * 16 + r16;
        if (m >= M) continue;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + ni * 16 + kq + 4 * r;
                if (n >= N) continue;
                if (tri == 1 && m < n) continue;
                if (tri == 2 && m > n) continue;
                double* c = C + m + n * ldc;
                const double v = acc[mi][ni][r];
                if (fold) *c = v;
                else *c = (beta == 0.0) ? alpha * v : __builtin_fma(alpha, v, beta * *c);
            }
        }
    }
}

// C := beta C on the touched triangle (k == 0 or alpha == 0). Grid-stride over all entries.
__global__ void dscale_kernel(i64 M, i64 N, double beta, double* C, i64 ldc, int tri) {
    const i64 total = M * N;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (i64)gridDim.x * blockDim.x) {
        const i64 i = idx % M, j = idx / M;
        if (tri == 1 && i < j) continue;
        if (tri == 2 && i > j) continue;
        double* c = C + i + j * ldc;
        *c = (beta == 0.0) ? 0.0 : beta * *c;
    }
}

}  // namespace

namespace elx {

elx_status_t launch_dgemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                          const double* A, i64 lda, const double* B, i64 ldb, double beta,
                          double* C, i64 ldc, int tri) {
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    if (k <= 0 || alpha == 0.0) {
        if (beta == 1.0) return ELX_SUCCESS;
        const unsigned blocks = (unsigned)std::min<i64>((m * n + 255) / 256, 65536);
        hipLaunchKernelGGL(dscale_kernel, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc, tri);
        ELX_HIP(hipGetLastError());
        return ELX_SUCCESS;
    }
    // Large products are plain library GEMMs: rocBLAS runs the FP64 rank-128 trailing updates
    // of the LU at 46-47 TFLOP/s against 34-35 for dgemm_kernel (m = n = 8192..32640); the
    // symmetric rank-k update of the Cholesky (A == B, one operand transposed, one triangle
    // of C) goes to dsyrk.
    if ((double)m * (double)n * (double)k >= (double)(1 << 24)) {
        ProfScope prof_(s, kKDgemm, (tri ? 1.0 : 2.0) * (double)m * (double)n * (double)k);
        if (tri == 0) return library_dgemm(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
        if (A == B && lda == ldb && m == n && ta != tb && (tri == 1) == tb)
            return library_dsyrk(s, tri == 2, ta, n, k, alpha, A, lda, beta, C, ldc);
    }
    const i64 tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    ELX_ARG(tm * tn < (i64)1 << 30);
    ProfScope prof_(s, kKDgemm, (tri ? 1.0 : 2.0) * (double)m * (double)n * (double)k);
    const unsigned blocks = (unsigned)(((tm * tn + 7) / 8) * 8);
    const int itm = (int)tm, itn = (int)tn;
    if (!ta && !tb)
        hipLaunchKernelGGL((dgemm_kernel<false, false>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn);
    else if (!ta && tb)
        hipLaunchKernelGGL((dgemm_kernel<false, true>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn);
    else if (ta && !tb)
        hipLaunchKernelGGL((dgemm_kernel<true, false>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn);
    else
        hipLaunchKernelGGL((dgemm_kernel<true, true>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

}  // namespace elx
