// FP64 / FP32 GEMM and SYRK on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64,
// v_mfma_f32_16x16x4_f32).
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
// accumulator values lie in one row of C (m fixed), 16 lanes covering 16 consecutive rows.
//
// In a factorization sequence the kernel takes the sequence's info word (`stop`): once an
// earlier step has recorded a failure, every workgroup returns at once and C is left as the
// reference leaves it when it throws.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "elx_common.hpp"

namespace {

constexpr int BM = 128, BN = 128, BK = 16, PAD = 4, NT = 256;
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// v_mfma_f64_16x16x4f64: register r of lane l holds D(l/16 + 4r, l%16)
template <>
struct Mfma<double> {
    typedef d4 acc;
    __device__ static acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int kq, int r) { return kq + 4 * r; }
};
// v_mfma_f32_16x16x4f32: register r of lane l holds D(4(l/16) + r, l%16)
template <>
struct Mfma<float> {
    typedef f4 acc;
    __device__ static acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int kq, int r) { return 4 * kq + r; }
};

// Loads of one BK-slab of op(A) (BM x BK) or op(B) (BK x BN) into 8 registers per thread.
// "mn-contiguous" operands (op(A) = A, op(B) = B^T) have 128 contiguous entries per k;
// "k-contiguous" operands (op(A) = A^T, op(B) = B) have 16 contiguous entries per m/n.
template <bool KCONTIG, typename T>
__device__ __forceinline__ void load_slab(const T* __restrict__ P, i64 ld, i64 mn0,
                                          i64 MN, i64 k0, i64 K, T (&r)[8]) {
    const int t = threadIdx.x;
    if (!KCONTIG) {
        const int kk = t >> 4;          // 0..15
        const int ii = (t & 15) * 8;    // 0..120
        const i64 k = k0 + kk;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const i64 i = mn0 + ii + q;
            r[q] = (k < K && i < MN) ? P[i + k * ld] : T(0);
        }
    } else {
        const int ii = t >> 1;          // 0..127
        const int kk = (t & 1) * 8;     // 0 or 8
        const i64 i = mn0 + ii;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const i64 k = k0 + kk + q;
            r[q] = (k < K && i < MN) ? P[k + i * ld] : T(0);
        }
    }
}

template <bool KCONTIG, typename T>
__device__ __forceinline__ void store_slab(T (*S)[BM + PAD], const T (&r)[8], T sgn = T(1)) {
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

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(NT, 2) void gemm_kernel(i64 M, i64 N, i64 K, T alpha,
                                                  const T* __restrict__ A, i64 lda,
                                                  const T* __restrict__ B, i64 ldb, T beta,
                                                  T* __restrict__ C, i64 ldc, int tri,
                                                  int tiles_m, int tiles_n,
                                                  const int32_t* __restrict__ stop) {
    typedef Mfma<T> MF;
    typedef typename MF::acc acc_t;
    __shared__ T As[BK][BM + PAD];  // As[k][i] = op(A)(m0+i, k0+k)
    __shared__ T Bs[BK][BN + PAD];  // Bs[k][j] = op(B)(k0+k, n0+j)
    if (stop && *stop != 0) return;

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
    const bool fold = (beta == T(1)) && (alpha == T(1) || alpha == T(-1));
    const T sgn = fold ? alpha : T(1);
    acc_t acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const i64 m = m0 + wm + mi * 16 + r16;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + ni * 16 + MF::row(kq, r);
                const bool ok = fold && m < M && n < N && !(tri == 1 && m < n) && !(tri == 2 && m > n);
                acc[mi][ni][r] = ok ? C[m + n * ldc] : T(0);
            }
    }

    T ra[8], rb[8];
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
            T af[4], bf[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                af[q] = As[kk + kq][wm + q * 16 + r16];
                bf[q] = Bs[kk + kq][wn + q * 16 + r16];
            }
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = MF::run(bf[ni], af[mi], acc[mi][ni]);
        }
        __syncthreads();
    }

    // With the operands fed as (op(B)^T, op(A)^T) the MFMA result is the transpose of the C
    // tile: the lane holds C(m0+wm+mi*16+r16, n0+wn+ni*16+MF::row(kq, r)), r = 0..3.
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const i64 m = m0 + wm + mi * 16 + r16;
        if (m >= M) continue;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + ni * 16 + MF::row(kq, r);
                if (n >= N) continue;
                if (tri == 1 && m < n) continue;
                if (tri == 2 && m > n) continue;
                T* c = C + m + n * ldc;
                const T v = acc[mi][ni][r];
                if (fold) *c = v;
                else *c = (beta == T(0)) ? alpha * v : __builtin_fma(alpha, v, beta * *c);
            }
        }
    }
}

// One 128x128 tile of dgemm_fast_kernel; FULL: the tile lies inside C, every load and store
// unchecked (a compile-time branch, so the K loop holds no per-tile conditions).
template <bool TA, bool TB, bool FULL>
__device__ __forceinline__ void dgemm_fast_tile(i64 M, i64 N, i64 K, double alpha,
                                                const double* __restrict__ A, i64 lda,
                                                const double* __restrict__ B, i64 ldb,
                                                double beta, double* __restrict__ C, i64 ldc,
                                                int tri, i64 m0, i64 n0,
                                                double (*As)[BK][BM + PAD],
                                                double (*Bs)[BK][BN + PAD]) {
    typedef double d2 __attribute__((ext_vector_type(2)));

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    const int r16 = lane & 15, kq = lane >> 4;
    const bool fold = (beta == 1.0) && (alpha == 1.0 || alpha == -1.0);
    const bool neg = fold && alpha < 0.0;  // alpha folded into the A operand (exact)
    d4 acc[4][4];
    // slab loads: A as 4 pairs along m per thread (16 k x 64 pairs); B as 4 pairs along k
    // (NN: 128 n x 8 pairs) or along n (NT: 16 k x 64 pairs)
    d2 ra[4], rb[4];
    auto gload = [&](i64 k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = t + NT * u;
            if (!TA) {
                const int k = idx >> 6, i2 = (idx & 63) * 2;
                const i64 m = m0 + i2, kk = k0 + k;
                if (FULL) {
                    ra[u] = *reinterpret_cast<const d2*>(A + m + kk * lda);
                } else {
                    ra[u][0] = (m < M) ? A[m + kk * lda] : 0.0;
                    ra[u][1] = (m + 1 < M) ? A[m + 1 + kk * lda] : 0.0;
                }
            } else {
                const int i = idx >> 3, k2 = (idx & 7) * 2;
                const i64 m = m0 + i, kk = k0 + k2;
                if (FULL) {
                    ra[u] = *reinterpret_cast<const d2*>(A + kk + m * lda);
                } else {
                    ra[u][0] = (m < M) ? A[kk + m * lda] : 0.0;
                    ra[u][1] = (m < M) ? A[kk + 1 + m * lda] : 0.0;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = t + NT * u;
            if (!TB) {
                const int n = idx >> 3, k2 = (idx & 7) * 2;
                const i64 nn = n0 + n, kk = k0 + k2;
                if (FULL) {
                    rb[u] = *reinterpret_cast<const d2*>(B + kk + nn * ldb);
                } else {
                    rb[u][0] = (nn < N) ? B[kk + nn * ldb] : 0.0;
                    rb[u][1] = (nn < N) ? B[kk + 1 + nn * ldb] : 0.0;
                }
            } else {
                const int k = idx >> 6, j2 = (idx & 63) * 2;
                const i64 nn = n0 + j2, kk = k0 + k;
                if (FULL) {
                    rb[u] = *reinterpret_cast<const d2*>(B + nn + kk * ldb);
                } else {
                    rb[u][0] = (nn < N) ? B[nn + kk * ldb] : 0.0;
                    rb[u][1] = (nn + 1 < N) ? B[nn + 1 + kk * ldb] : 0.0;
                }
            }
        }
    };
    auto sstore = [&](int b) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = t + NT * u;
            if (!TA) {
                const int k = idx >> 6, i2 = (idx & 63) * 2;
                *reinterpret_cast<d2*>(&As[b][k][i2]) = neg ? -ra[u] : ra[u];
            } else {
                const int i = idx >> 3, k2 = (idx & 7) * 2;
                As[b][k2][i] = neg ? -ra[u][0] : ra[u][0];
                As[b][k2 + 1][i] = neg ? -ra[u][1] : ra[u][1];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = t + NT * u;
            if (!TB) {
                const int n = idx >> 3, k2 = (idx & 7) * 2;
                Bs[b][k2][n] = rb[u][0];
                Bs[b][k2 + 1][n] = rb[u][1];
            } else {
                const int k = idx >> 6, j2 = (idx & 63) * 2;
                *reinterpret_cast<d2*>(&Bs[b][k][j2]) = rb[u];
            }
        }
    };
    // the first slab's loads go out before C's, so that storing the slab waits for them alone
    // (loads complete in order) and C's 64 loads per thread stay in flight into the first steps
    gload(0);
    if (FULL && fold && tri == 0) {
        // the whole tile, no per-entry conditions (the LU's trailing update)
        const double* c0 = C + m0 + wm + r16 + (n0 + wn + kq) * ldc;
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[mi][ni][r] = c0[mi * 16 + (ni * 16 + 4 * r) * ldc];
    } else {
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const i64 m = m0 + wm + mi * 16 + r16;
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const i64 n = n0 + wn + ni * 16 + kq + 4 * r;
                    const bool ok = fold && (FULL || (m < M && n < N)) &&
                                    !(tri == 1 && m < n) && !(tri == 2 && m > n);
                    acc[mi][ni][r] = ok ? C[m + n * ldc] : 0.0;
                }
        }
    }
    sstore(0);
    __syncthreads();
    int cur = 0;
    for (i64 k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[4], bfv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                af[q] = As[cur][kk + kq][wm + q * 16 + r16];
                bfv[q] = Bs[cur][kk + kq][wn + q * 16 + r16];
            }
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bfv[ni], af[mi], acc[mi][ni],
                                                                       0, 0, 0);
        }
        if (more) sstore(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const i64 m = m0 + wm + mi * 16 + r16;
        if (!FULL && m >= M) continue;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + ni * 16 + kq + 4 * r;
                if (!FULL && n >= N) continue;
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


// FP64 C := alpha op(A) op(B) + beta C (the LU's Gemm(NORMAL, NORMAL) update, the Cholesky's
// Herk(LOWER, NORMAL) = A A^T and Herk(UPPER, ADJOINT) = A^T A updates on one triangle), for K a
// multiple of 16 and operands whose columns start on 16-byte boundaries (even leading
// dimensions, aligned bases).
// Same 128x128 tile, 4x4 MFMA accumulators per wave and XCD-aware tile order as gemm_kernel,
// but with 16-byte operand loads and double-buffered LDS slabs: the next slab is fetched into
// registers while the current one feeds the MFMAs and is then stored into the other buffer, so
// each 16-deep step costs one barrier. Edge tiles (partial in m or n) load element by element.
template <bool TA, bool TB, bool FULL>
__global__ __launch_bounds__(NT, 2) void dgemm_fast_kernel(i64 M, i64 N, i64 K, double alpha,
                                                        const double* __restrict__ A, i64 lda,
                                                        const double* __restrict__ B, i64 ldb,
                                                        double beta, double* __restrict__ C,
                                                        i64 ldc, int tri, int tiles_m, int tiles_n,
                                                        const int32_t* __restrict__ stop) {
    if (stop && *stop != 0) return;
    const int total = tiles_m * tiles_n;
    const int per_xcd = (total + 7) / 8;
    const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (lin >= total) return;
    constexpr int GROUP = 8;
    const int group = lin / (GROUP * tiles_n);
    const int first_m = group * GROUP;
    const int gsize = min(tiles_m - first_m, GROUP);
    const int tm = first_m + (lin % (GROUP * tiles_n)) % gsize;
    const int tn = (lin % (GROUP * tiles_n)) / gsize;
    const i64 m0 = (i64)tm * BM, n0 = (i64)tn * BN;
    if (tri == 1 && n0 > m0 + BM - 1) return;
    if (tri == 2 && m0 > n0 + BN - 1) return;
    // As[b][k][i] = sgn op(A)(m0+i, k0+k), Bs[b][k][j] = op(B)(k0+k, n0+j), double-buffered
    __shared__ __align__(16) double As[2][BK][BM + PAD];
    __shared__ __align__(16) double Bs[2][BK][BN + PAD];
    // FULL: launched over the tiles inside C only; otherwise over every tile, the edge tiles
    // done here and the full ones left to the FULL launch
    if (!FULL && (m0 + BM <= M) && (n0 + BN <= N)) return;
    dgemm_fast_tile<TA, TB, FULL>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri, m0, n0, As, Bs);
}

// The large FP64 products on rocBLAS instead of the hand-written kernels: environment
// ELX_DGEMM=library, read at every call.
bool dgemm_own_selected() {
    const char* v = std::getenv("ELX_DGEMM");
    return !(v && std::strcmp(v, "library") == 0);
}

// C := beta C on the touched triangle (k == 0 or alpha == 0). Grid-stride over all entries.
template <typename T>
__global__ void scale_kernel(i64 M, i64 N, T beta, T* C, i64 ldc, int tri,
                             const int32_t* __restrict__ stop) {
    if (stop && *stop != 0) return;
    const i64 total = M * N;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (i64)gridDim.x * blockDim.x) {
        const i64 i = idx % M, j = idx / M;
        if (tri == 1 && i < j) continue;
        if (tri == 2 && i > j) continue;
        T* c = C + i + j * ldc;
        *c = (beta == T(0)) ? T(0) : beta * *c;
    }
}

// alpha and beta of a library call in a factorization sequence: (alpha, beta) while no step
// has failed, (0, 1) -- C unchanged -- once one has.
__global__ void gate_kernel(const int32_t* __restrict__ stop, double alpha, double beta,
                            double* out) {
    const bool halt = *stop != 0;
    out[0] = halt ? 0.0 : alpha;
    out[1] = halt ? 1.0 : beta;
}

template <typename T>
elx_status_t gemm_impl(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, T alpha,
                       const T* A, i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc, int tri,
                       const int32_t* stop, bool allow_library) {
    using namespace elx;
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    if (k <= 0 || alpha == T(0)) {
        if (beta == T(1)) return ELX_SUCCESS;
        const unsigned blocks = (unsigned)std::min<i64>((m * n + 255) / 256, 65536);
        hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc, tri,
                           stop);
        ELX_HIP(hipGetLastError());
        return ELX_SUCCESS;
    }
    const double flops = (tri ? 1.0 : 2.0) * (double)m * (double)n * (double)k;
    if constexpr (sizeof(T) == 8) {
        // Large FP64 products go to rocBLAS (dgemm, and dsyrk for the symmetric rank-k update:
        // A == B, one operand transposed, one triangle of C) with ELX_DGEMM=library, and so do
        // the narrow ones with a long K (fewer than 128 rows or columns and K >= 4096, e.g. the
        // residual of a block of right-hand sides, 32768 x 64 x 32768): there every 128 x 128
        // tile of the hand-written kernel is an edge tile, and only m/128 workgroups each run
        // the whole K extent (1.4 ms more per residual at that shape). In a factorization
        // sequence the coefficients come from the device (gate_kernel), so a recorded failure
        // leaves C as it is.
        const bool narrow = std::min(m, n) < BM && k >= 4096;
        if (allow_library && (!dgemm_own_selected() || narrow) &&
            (double)m * (double)n * (double)k >= (double)(1 << 24)) {
            const bool syrk = tri != 0 && A == B && lda == ldb && m == n && ta != tb && (tri == 1) == tb;
            if (tri == 0 || syrk) {
                const double* coef = nullptr;
                if (stop) {
                    double* g = nullptr;
                    ELX_CHECK(gate_slot(&g));
                    hipLaunchKernelGGL(gate_kernel, dim3(1), dim3(1), 0, s, stop, alpha, beta, g);
                    ELX_HIP(hipGetLastError());
                    coef = g;
                }
                ProfScope prof_(s, kKDgemm, flops);
                if (!syrk)
                    return library_dgemm(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, coef);
                return library_dsyrk(s, tri == 2, ta, n, k, alpha, A, lda, beta, C, ldc, coef);
            }
        }
        // 16-byte operand loads and double-buffered slabs where the operands allow them
        const bool vec = (k % BK == 0) && (lda % 2 == 0) && (ldb % 2 == 0) &&
                         (reinterpret_cast<uintptr_t>(A) % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(B) % 16 == 0);
        if (vec) {
            ProfScope prof_(s, kKDgemm, flops);
            const double* dA = reinterpret_cast<const double*>(A);
            const double* dB = reinterpret_cast<const double*>(B);
            double* dC = reinterpret_cast<double*>(C);
            const double al = (double)alpha, be = (double)beta;
            // interior tiles (unchecked), then, if C has partial edge tiles, the checked launch
            const i64 fm = m / BM, fn = n / BN;
            const i64 tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
            ELX_ARG(tm * tn < (i64)1 << 30);
            auto go = [&](auto full_tag, i64 gm, i64 gn) {
                constexpr bool F = decltype(full_tag)::value;
                if (gm <= 0 || gn <= 0) return;
                const unsigned blocks = (unsigned)(((gm * gn + 7) / 8) * 8);
                const int igm = (int)gm, ign = (int)gn;
                if (!ta && !tb)
                    hipLaunchKernelGGL((dgemm_fast_kernel<false, false, F>), dim3(blocks), dim3(NT), 0,
                                       s, m, n, k, al, dA, lda, dB, ldb, be, dC, ldc, tri, igm, ign, stop);
                else if (!ta && tb)
                    hipLaunchKernelGGL((dgemm_fast_kernel<false, true, F>), dim3(blocks), dim3(NT), 0,
                                       s, m, n, k, al, dA, lda, dB, ldb, be, dC, ldc, tri, igm, ign, stop);
                else if (ta && !tb)
                    hipLaunchKernelGGL((dgemm_fast_kernel<true, false, F>), dim3(blocks), dim3(NT), 0,
                                       s, m, n, k, al, dA, lda, dB, ldb, be, dC, ldc, tri, igm, ign, stop);
                else
                    hipLaunchKernelGGL((dgemm_fast_kernel<true, true, F>), dim3(blocks), dim3(NT), 0,
                                       s, m, n, k, al, dA, lda, dB, ldb, be, dC, ldc, tri, igm, ign, stop);
            };
            go(std::integral_constant<bool, true>{}, fm, fn);
            if (fm != tm || fn != tn) go(std::integral_constant<bool, false>{}, tm, tn);
            ELX_HIP(hipGetLastError());
            return ELX_SUCCESS;
        }
    }
    const i64 tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    ELX_ARG(tm * tn < (i64)1 << 30);
    ProfScope prof_(s, kKDgemm, flops);
    const unsigned blocks = (unsigned)(((tm * tn + 7) / 8) * 8);
    const int itm = (int)tm, itn = (int)tn;
    if (!ta && !tb)
        hipLaunchKernelGGL((gemm_kernel<T, false, false>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn, stop);
    else if (!ta && tb)
        hipLaunchKernelGGL((gemm_kernel<T, false, true>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn, stop);
    else if (ta && !tb)
        hipLaunchKernelGGL((gemm_kernel<T, true, false>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn, stop);
    else
        hipLaunchKernelGGL((gemm_kernel<T, true, true>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, B, ldb, beta, C, ldc, tri, itm, itn, stop);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

}  // namespace

namespace elx {

elx_status_t launch_gemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                         const double* A, i64 lda, const double* B, i64 ldb, double beta,
                         double* C, i64 ldc, int tri, const int32_t* stop, bool allow_library) {
    return gemm_impl<double>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, stop,
                             allow_library);
}

elx_status_t launch_gemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, float alpha,
                         const float* A, i64 lda, const float* B, i64 ldb, float beta, float* C,
                         i64 ldc, int tri, const int32_t* stop, bool allow_library) {
    return gemm_impl<float>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, stop,
                            allow_library);
}

elx_status_t launch_dgemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                          const double* A, i64 lda, const double* B, i64 ldb, double beta,
                          double* C, i64 ldc, int tri) {
    return gemm_impl<double>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, nullptr,
                             true);
}

}  // namespace elx
