// Batched variable-size Cholesky factorization and solve (lower, FP64 and FP32).
//
// Each matrix runs the blocked lower Cholesky of El::Cholesky (variant 3,
// src/lapack_like/factor/Cholesky/LowerVariant3.hpp:15-68) with NB-column blocks (NB = 16 or
// 32), in the left-looking order: a block column first receives the updates of every earlier
// block column (one GEMM with K = the number of earlier columns, on MFMA), then its NB x NB
// diagonal block is factored unblocked (alpha = sqrt(a_jj), a_jj <= 0 -> not HPD,
// a21 *= 1/alpha, then the dsyr rank-1 update, as LowerVariant3Unblocked does) and the rows
// below are solved against it (X L11^T = A21, here as the MFMA product of A21 with the
// transposed inverse of L11, which differs from the reference's substitution only in
// rounding). The solve follows cholesky::SolveAfter
// (src/lapack_like/factor/Cholesky/SolveAfter.hpp:15-43): L y = b, then L^T x = y, blocked by
// 32 rows, up to 8 right-hand sides at a time.
//
// Design for MI355X: matrices are sorted largest first and split into size bins (n <= 128,
// <= 256, <= 512). Each bin has its own persistent kernel whose LDS holds one block column of
// the bin's largest order, so smaller matrices run with more workgroups per CU (3, 2, 1 for
// FP64): while one workgroup is in the serial diagonal factor, another keeps the matrix cores
// busy. Within a bin, workgroups pull matrices from a work queue. The block column being
// factored is accumulated in MFMA registers and staged in LDS; the earlier columns stream from
// memory once per block column. The block width and the workgroups per CU of each bin come
// from a measured table (batched_tuning.hpp). A factor + solve sequence can be captured once
// into a HIP graph and replayed (elx_?posv_vbatched_graph).
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "batched_tuning.hpp"
#include "elx_common.hpp"

namespace {

constexpr int BNT = 256;    // threads per workgroup
constexpr int BNMAX = 512;  // largest order
constexpr int NBIN = 3;
constexpr int kBinMax[NBIN] = {128, 256, 512};
constexpr int kQueueFactor = 0;  // queue words: one per factor bin, then the solve's
constexpr int kQueueSolve = NBIN;
constexpr int kQueueWords = 8;

template <typename T>
struct Mma;
template <>
struct Mma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int drow(int l, int r) { return l / 16 + 4 * r; }
};
template <>
struct Mma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int drow(int l, int r) { return 4 * (l / 16) + r; }
};

struct BatchItem {
    int32_t n;
    int32_t lda;
    int32_t index;  // position in the caller's arrays
    int32_t nrhs;
    int32_t ldb;
    int32_t pad;
};

template <typename T>
__device__ __forceinline__ T fsqrt(T x) {
    return __builtin_sqrt(x);
}
template <>
__device__ __forceinline__ float fsqrt<float>(float x) {
    return __builtin_sqrtf(x);
}

// Broadcast of lane `src`'s value (src uniform) through the scalar unit.
__device__ __forceinline__ double bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// A matrix in global memory. The matrices' pointers come from a device array, so without the
// address space the compiler emits flat loads, which also count as LDS operations: every wait
// for an LDS read would then wait for the next chunk's loads still in flight as well.
template <typename T>
using gmem = __attribute__((address_space(1))) T;

// Compile-time loop: f(std::integral_constant<int, i>) for i = B .. E-1.
template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// S[kk * SM + r] = A(k0 + r, c0 + kk) for r < m, kk < kc, zero for kc <= kk < kcp: threads
// walk the (row, column) pairs column by column (coalesced), 16 loads in flight per thread.
template <typename T>
__device__ __forceinline__ void stage_cols(const gmem<T>* A, int lda, int k0, int m, int c0, int kc,
                                           int kcp, int SM, T* S) {
    const int total = m * kcp;
    for (int base = threadIdx.x; base < total; base += 16 * BNT) {
        T v[16];
        int off[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * BNT;
            const int kk = idx / m, r = idx - kk * m;
            off[u] = (idx < total) ? kk * SM + r : -1;
            v[u] = (idx < total && kk < kc) ? A[(k0 + r) + (size_t)(c0 + kk) * lda] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (off[u] >= 0) S[off[u]] = v[u];
    }
}

// One workgroup: the lower Cholesky factor of one matrix of order n <= NM, in place. P is
// NM x (NB+1) values of LDS (the block column, also the staging area), Minv (NB+1) x NB.
template <typename T, int NM, int NB>
__device__ void potrf_one(int n, gmem<T>* A, int lda, int32_t* info, T* P, T* Minv) {
    typedef typename Mma<T>::v4 v4;
    constexpr int MT = NM / 64;     // row tiles per wave (NM/16 tiles over 4 waves)
    constexpr int CT = NB / 16;     // column tiles of a block column
    constexpr int LD = NB + 1;      // odd row stride of P and Minv
    constexpr int LPC = BNT / NB;   // lanes per column of the inverse of L11
    constexpr int RPL = NB / LPC;   // rows per lane of that inverse
    constexpr int SCAP = NM * LD;   // staging capacity (values)
    // staged values per thread per chunk (16 where the registers are halved for two
    // workgroups per CU in the largest bin)
    constexpr int PV = (NM <= 128 || (NM > 256 && NB == 16)) ? 16 : 32;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = lane >> 4;
    T* S = P;
    __shared__ int s_fail;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0), m = n - k0;
        const int nrt = (m + 15) >> 4;
        const int SM = nrt * 16;
        // chunk width: at most PV staged values per thread and the staging capacity, a
        // multiple of 4 columns
        const int kcw = min(256, max(4, (min(PV * BNT, SCAP) / SM) & ~3));
        const int cq = BNT / m, cr = BNT - cq * m;  // BNT = cq*m + cr: the walk over (row, column)
        // (1) C = A(k0:n, k0:k0+kb) - L(k0:n, 0:k0) L(k0:k0+kb, 0:k0)^T: the block column and
        // then chunks of L to its left pass through LDS; wave w accumulates row tiles w, w+4, ...
        // (all of them at once, in MFMA registers). Rows >= m of a staged tile and columns
        // >= kb of the block only reach accumulator entries that are never used.
        stage_cols(A, lda, k0, m, k0, kb, NB, SM, S);
        __syncthreads();
        v4 acc[MT][CT];
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (wave + 4 * u) * 16 + Mma<T>::drow(lane, r);
                    const int col = ct * 16 + lr;
                    acc[u][ct][r] = (row < m) ? S[col * SM + row] : T(0);
                }
        // chunk pipeline: the loads of chunk i+1 are in flight while chunk i is multiplied
        T pv[PV];
        auto load_chunk = [&](int kc, int w) {
            const int total = m * w;
            int kk = t / m, r = t - kk * m;
#pragma unroll
            for (int u = 0; u < PV; ++u) {
                const int idx = t + u * BNT;
                pv[u] = (idx < total) ? A[(k0 + r) + (size_t)(kc + kk) * lda] : T(0);
                r += cr;
                kk += cq;
                if (r >= m) {
                    r -= m;
                    ++kk;
                }
            }
        };
        auto store_chunk = [&](int w) {
            const int total = m * w;
            int kk = t / m, r = t - kk * m;
#pragma unroll
            for (int u = 0; u < PV; ++u) {
                const int idx = t + u * BNT;
                if (idx < total) S[kk * SM + r] = pv[u];
                r += cr;
                kk += cq;
                if (r >= m) {
                    r -= m;
                    ++kk;
                }
            }
        };
        int wnext = min(kcw, k0);
        if (k0 > 0) load_chunk(0, wnext);
        for (int kc = 0; kc < k0;) {
            const int w = wnext;
            __syncthreads();
            store_chunk(w);
            __syncthreads();
            const int kcn = kc + w;
            wnext = min(kcw, k0 - kcn);
            if (kcn < k0) load_chunk(kcn, wnext);
#pragma unroll 4
            for (int kk = 0; kk < w; kk += 4) {
                const T* Sk = S + (kk + lk) * SM;
                T b[CT];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) b[ct] = Sk[ct * 16 + lr];
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    const int rt = wave + 4 * u;
                    if (rt < nrt) {
                        const T a = -Sk[rt * 16 + lr];
#pragma unroll
                        for (int ct = 0; ct < CT; ++ct) acc[u][ct] = Mma<T>::mma(a, b[ct], acc[u][ct]);
                    }
                }
            }
            kc = kcn;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (wave + 4 * u) * 16 + Mma<T>::drow(lane, r);
                    const int col = ct * 16 + lr;
                    if (row < m) P[row * LD + col] = acc[u][ct][r];
                }
        __syncthreads();
        // (2) the diagonal block, unblocked, by wave 0: lane i holds row i. Each step
        // broadcasts the pivot and the scaled column through the scalar unit; the dsyr update
        // is a select, not a branch. The factored block stays in rows 0..kb-1 of P (L11).
        if (wave == 0) {
            T rw[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) rw[c] = (lane < kb) ? P[lane * LD + c] : T(0);
            int fail = -1;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j < kb && fail < 0) {
                    const T ajj = bcast(rw[j], j);
                    if (ajj <= T(0)) {
                        fail = j;
                    } else {
                        const T alpha = fsqrt(ajj);
                        const T inv = T(1) / alpha;
                        rw[j] = (lane == j) ? alpha : ((lane > j) ? rw[j] * inv : rw[j]);
                        // dsyr on the trailing lower triangle. Lanes < l update upper-triangle
                        // entries that are never read, so no lane mask is needed; the update is
                        // applied for a zero x_l too (the reference BLAS skips it), which can
                        // only change the sign of a zero entry.
                        const T xi = rw[j];
#pragma unroll
                        for (int l = j + 1; l < NB; ++l) rw[l] = __builtin_fma(xi, -bcast(rw[j], l), rw[l]);
                    }
                }
            }
            if (lane < kb)
#pragma unroll
                for (int c = 0; c < NB; ++c) P[lane * LD + c] = rw[c];
            if (lane == 0) s_fail = fail;
        }
        __syncthreads();
        const int fail = s_fail;
        if (fail < 0 && m > NB) {
            // inv(L11) (rows below exist, so kb == NB) by all threads: LPC lanes per column c,
            // lane g of the group holds rows g, g+LPC, ... of y = L11^{-1} e_c, formed in the
            // column order (y_k /= l_kk, then y_i -= l_ik y_k); stored as row c of Minv, so
            // Minv(k, j) = inv(L11)(j, k)
            const int c = t / LPC, g = t % LPC;
            T y[RPL];
#pragma unroll
            for (int q = 0; q < RPL; ++q) y[q] = (g + LPC * q == c) ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int qk = k / LPC, gk = k % LPC;
                if (g == gk) y[qk] = y[qk] / P[k * LD + k];
                const T yk = __shfl(y[qk], (lane & ~(LPC - 1)) | gk);
#pragma unroll
                for (int q = 0; q < RPL; ++q) {
                    const int i = g + LPC * q;
                    if (i > k) y[q] = __builtin_fma(-P[i * LD + k], yk, y[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < RPL; ++q) Minv[c * LD + g + LPC * q] = y[q];
            __syncthreads();
        }
        if (fail >= 0) {
            // NonHPDMatrixException at global column k0+fail. The block column as the blocked
            // reference holds it when it throws: the diagonal block's lower triangle after the
            // finished steps, the rows below updated by the earlier block columns
            if (t == 0) *info = k0 + fail + 1;
            for (int idx = t; idx < m * kb; idx += BNT) {
                const int i = idx % m, c = idx / m;
                if (i >= c) A[(k0 + i) + (size_t)(k0 + c) * lda] = P[i * LD + c];
            }
            __syncthreads();
            return;
        }
        // (3) rows below (present only when kb == NB): X = A21 L11^{-T} as one MFMA product
        // with the inverse, Minv(k, j) = inv(L11)(j, k); results go straight to A
        for (int rt = CT + wave; rt < nrt; rt += 4) {
            v4 acc2[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc2[ct] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kk = 0; kk < NB; kk += 4) {
                const T a = P[(rt * 16 + lr) * LD + kk + lk];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                    acc2[ct] = Mma<T>::mma(a, Minv[(kk + lk) * LD + ct * 16 + lr], acc2[ct]);
            }
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rt * 16 + Mma<T>::drow(lane, r);
                    if (row < m) A[(k0 + row) + (size_t)(k0 + ct * 16 + lr) * lda] = acc2[ct][r];
                }
        }
        // the diagonal block's lower triangle
        for (int idx = t; idx < kb * kb; idx += BNT) {
            const int i = idx % kb, c = idx / kb;
            if (i >= c) A[(k0 + i) + (size_t)(k0 + c) * lda] = P[i * LD + c];
        }
        __syncthreads();
    }
}

// Register-direct form (variant kind 1) of potrf_one for the larger bins: the left-looking
// update of a block column reads L straight from memory into MFMA operand registers, DEPTH
// 4-column steps ahead, with no LDS staging and no barrier inside the product; the updated
// block column goes back to A, the diagonal block is factored by wave 0 from there, and the
// rows below are read back for their product with the inverse of L11. LDS holds only L11 and
// its inverse, so several workgroups share a CU and one's serial diagonal factor overlaps
// another's products.
template <typename T, int NM, int NB>
__device__ void potrf_rd_one(int n, gmem<T>* A, int lda, int32_t* info) {
    typedef typename Mma<T>::v4 v4;
    constexpr int MT = NM / 64;   // row tiles per wave
    constexpr int CT = NB / 16;   // column tiles of a block column
    constexpr int LD = NB + 1;
    constexpr int LPC = BNT / NB;
    constexpr int RPL = NB / LPC;
    // 4-column steps in flight (block columns are multiples of 16); 2 where the accumulators
    // take half the registers (FP64, 512 rows x 32 columns)
    constexpr int DEPTH = (NM * NB * (int)sizeof(T) >= 512 * 32 * 8) ? 2 : 4;
    __shared__ T Dg[NB * LD];     // L11
    __shared__ T Minv[NB * LD];   // Minv(k, j) = inv(L11)(j, k)
    __shared__ int s_fail;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0), m = n - k0;
        const int nrt = (m + 15) >> 4;
        // (1) C = A(k0:n, k0:k0+kb) - L(k0:n, 0:k0) L(k0:k0+kb, 0:k0)^T; wave w owns row tiles
        // w, w+4, ...
        v4 acc[MT][CT];
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (wave + 4 * u) * 16 + Mma<T>::drow(lane, r);
                    const int col = ct * 16 + lr;
                    acc[u][ct][r] = (row < m && col < kb) ? A[(k0 + row) + (size_t)(k0 + col) * lda]
                                                          : T(0);
                }
        if (k0 > 0) {
            T ra[DEPTH][MT], rb[DEPTH][CT];
            auto load = [&](auto dc, int kk) {
                constexpr int d = decltype(dc)::value;
                const gmem<T>* col = A + (size_t)(kk + lk) * lda + k0;
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    const int row = (wave + 4 * u) * 16 + lr;
                    ra[d][u] = (row < m) ? -col[row] : T(0);
                }
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    const int c = ct * 16 + lr;
                    rb[d][ct] = (c < kb) ? col[c] : T(0);
                }
            };
            static_for<0, DEPTH>([&](auto dc) {
                if (4 * decltype(dc)::value < k0) load(dc, 4 * decltype(dc)::value);
            });
            for (int kk = 0; kk < k0; kk += 4 * DEPTH) {
                static_for<0, DEPTH>([&](auto dc) {
                    constexpr int d = decltype(dc)::value;
                    const int ks = kk + 4 * d;
                    if (ks < k0) {
#pragma unroll
                        for (int u = 0; u < MT; ++u) {
                            if (wave + 4 * u < nrt) {
#pragma unroll
                                for (int ct = 0; ct < CT; ++ct)
                                    acc[u][ct] = Mma<T>::mma(ra[d][u], rb[d][ct], acc[u][ct]);
                            }
                        }
                        if (ks + 4 * DEPTH < k0) load(dc, ks + 4 * DEPTH);
                    }
                });
            }
        }
        // (2) the updated block column back to A (on and below the diagonal only)
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (wave + 4 * u) * 16 + Mma<T>::drow(lane, r);
                    const int col = ct * 16 + lr;
                    if (row < m && col < kb && row >= col)
                        A[(k0 + row) + (size_t)(k0 + col) * lda] = acc[u][ct][r];
                }
        __syncthreads();
        // (3) the diagonal block, unblocked, by wave 0 (lane i holds row i; entries above the
        // diagonal are never read into the result), as in potrf_one
        if (wave == 0) {
            T rw[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c)
                rw[c] = (lane < kb && c <= lane) ? A[(k0 + lane) + (size_t)(k0 + c) * lda] : T(0);
            int fail = -1;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j < kb && fail < 0) {
                    const T ajj = bcast(rw[j], j);
                    if (ajj <= T(0)) {
                        fail = j;
                    } else {
                        const T alpha = fsqrt(ajj);
                        const T inv = T(1) / alpha;
                        rw[j] = (lane == j) ? alpha : ((lane > j) ? rw[j] * inv : rw[j]);
                        const T xi = rw[j];
#pragma unroll
                        for (int l = j + 1; l < NB; ++l) rw[l] = __builtin_fma(xi, -bcast(rw[j], l), rw[l]);
                    }
                }
            }
            if (lane < kb) {
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    Dg[lane * LD + c] = rw[c];
                    if (c <= lane) A[(k0 + lane) + (size_t)(k0 + c) * lda] = rw[c];
                }
            }
            if (lane == 0) s_fail = fail;
        }
        __syncthreads();
        const int fail = s_fail;
        if (fail >= 0) {
            // NonHPDMatrixException at global column k0+fail: the diagonal block's lower
            // triangle after the finished steps, the rows below updated by the earlier block
            // columns (both already in A)
            if (t == 0) *info = k0 + fail + 1;
            __syncthreads();
            return;
        }
        if (m > NB) {
            // (4) inv(L11) by all threads, as in potrf_one
            const int c = t / LPC, g = t % LPC;
            T y[RPL];
#pragma unroll
            for (int q = 0; q < RPL; ++q) y[q] = (g + LPC * q == c) ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int qk = k / LPC, gk = k % LPC;
                if (g == gk) y[qk] = y[qk] / Dg[k * LD + k];
                const T yk = __shfl(y[qk], (lane & ~(LPC - 1)) | gk);
#pragma unroll
                for (int q = 0; q < RPL; ++q) {
                    const int i = g + LPC * q;
                    if (i > k) y[q] = __builtin_fma(-Dg[i * LD + k], yk, y[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < RPL; ++q) Minv[c * LD + g + LPC * q] = y[q];
            __syncthreads();
            // (5) rows below: X = A21 L11^{-T} as one MFMA product with the inverse; each wave
            // reads its row tiles completely before it overwrites them
            for (int rt = CT + wave; rt < nrt; rt += 4) {
                v4 acc2[CT];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) acc2[ct] = v4{T(0), T(0), T(0), T(0)};
                const int row = rt * 16 + lr;
                T a[NB / 4];
#pragma unroll
                for (int kk = 0; kk < NB; kk += 4)
                    a[kk / 4] = (row < m) ? A[(k0 + row) + (size_t)(k0 + kk + lk) * lda] : T(0);
#pragma unroll
                for (int kk = 0; kk < NB; kk += 4)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct)
                        acc2[ct] = Mma<T>::mma(a[kk / 4], Minv[(kk + lk) * LD + ct * 16 + lr], acc2[ct]);
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int rw2 = rt * 16 + Mma<T>::drow(lane, r);
                        if (rw2 < m) A[(k0 + rw2) + (size_t)(k0 + ct * 16 + lr) * lda] = acc2[ct][r];
                    }
            }
        }
        __syncthreads();
    }
}

// Wide-panel form (variant kind 2) for the larger bins. The 16- or 32-column block columns
// above stream every earlier column once per block column, at 4-8 flops per byte read, which
// bounds the n <= 512 bin by memory; here the left-looking update runs over 64-column panels
// (16 flops per byte). A panel is updated in passes of up to 256 rows, the pass holding the
// diagonal block last: a 16-column chunk of the earlier columns, for the pass's rows and for the
// panel's first 64 rows, goes through LDS while the next chunk's loads are in flight in
// registers. The accumulators hold the panel transposed (MFMA rows = panel columns), which is
// the operand layout of the product with the inverse, so the last pass's rows below the
// diagonal block never leave the registers. The 64 x 64 diagonal block is factored in LDS in
// 16-column steps (wave 0 factors each 16 x 16 block as potrf_one does, all threads invert it,
// the other blocks are updated with MFMA); its inverse is assembled from the four block
// inverses by MFMA (wave j forms block column j), and the rows below are multiplied by the
// transposed inverse, as the narrower forms do with their diagonal blocks.
template <typename T>
struct Wide {
    static constexpr int W = 64, H = 256, KC = 8;
    static constexpr int LSA = H + 4, LSB = W + 4;  // strides of the staged chunk (k-major)
    static constexpr int LDD = W + 1;               // row stride of L11 and its inverse
    static constexpr int L16 = 17;                  // row stride of the 16 x 16 block inverses
    static constexpr int buf = KC * (LSA + LSB);    // one chunk buffer (two of them)
    static constexpr int stage = 2 * buf;
    static constexpr int diag = 2 * W * LDD + 4 * 16 * L16;
    static constexpr size_t lds = sizeof(T) * (size_t)(stage > diag ? stage : diag);
};

template <typename T, int NM>
__device__ void potrf_wide_one(int n, gmem<T>* A, int lda, int32_t* info, T* smem) {
    typedef typename Mma<T>::v4 v4;
    typedef Wide<T> C;
    constexpr int W = C::W, H = C::H, KC = C::KC, LSA = C::LSA, LSB = C::LSB, LDD = C::LDD,
                  L16 = C::L16;
    constexpr int MTH = H / 64;       // row tiles per wave in a pass
    constexpr int CTW = W / 16;       // column tiles of a panel
    constexpr int BV = W * KC / BNT;  // staged values per thread of the panel's first rows
    static_assert(H == BNT && NM <= 2 * H, "one pass row per thread, at most two passes");
    // chunk buffer b: the pass rows at smem + b * buf (kk * LSA + r), then the panel's first W
    // rows (kk * LSB + r)
    T* D = smem;                  // L11 at D[i * LDD + j] (once the chunks are done)
    T* Minv = smem + W * LDD;     // Minv[k * LDD + j] = inv(L11)(j, k)
    T* M16 = smem + 2 * W * LDD;  // M16[b * 16 * L16 + k * L16 + j] = inv(L_bb)(j, k)
    __shared__ int s_fail;
    for (int K0 = 0; K0 < n; K0 += W) {
        // the thread index through an opaque copy, once per panel: the many per-lane offsets
        // and conditions derived from it are otherwise hoisted out of the panel loop to the
        // kernel's start, where they do not fit in registers
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
        const int lr = lane & 15, lk = lane >> 4;
        const int kw = min(W, n - K0), m = n - K0;
        gmem<T>* Ap = A + K0 + (size_t)K0 * lda;  // Ap[r + c * lda] = A(K0 + r, K0 + c)
        // (1) the panel minus L(K0:n, 0:K0) L(K0:K0+kw, 0:K0)^T, pass by pass (only the last
        // pass for the first panel: the rows below are A's). acc[ct][u] holds the entries
        // (row r0 + (wave + 4u)*16 + lr, column ct*16 + drow(lane, r)); entries above the
        // diagonal start at zero (never read into the result).
        v4 acc[CTW][MTH];
        for (int p = K0 > 0 ? (m - 1) / H : 0; p >= 0; --p) {
            const int r0 = p * H;
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                for (int u = 0; u < MTH; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = r0 + (wave + 4 * u) * 16 + lr;
                        const int col = ct * 16 + Mma<T>::drow(lane, r);
                        acc[ct][u][r] = (row < m && col < kw && row >= col)
                                            ? Ap[row + (size_t)col * lda] : T(0);
                    }
            if (K0 > 0) {
                const int arow = r0 + t, brow = lane, bc = wave * BV;
                // chunk c (columns c*KC ..) is loaded into registers R[c & 1] two chunks
                // ahead, stored to buffer c & 1 after chunk c-1 is multiplied: one barrier per
                // chunk, two chunks of loads in flight
                T va[2][KC], vb[2][BV];
                auto load = [&](auto pc, int c) {
                    constexpr int P = decltype(pc)::value;
#pragma unroll
                    for (int j = 0; j < KC; ++j) {
                        // column bases are the same for every lane (scalar registers)
                        const gmem<T>* col = A + K0 + (size_t)(c * KC + j) * lda;
                        va[P][j] = (arow < m) ? col[arow] : T(0);
                    }
#pragma unroll
                    for (int j = 0; j < BV; ++j) {
                        const gmem<T>* col = A + K0 + (size_t)(c * KC + bc + j) * lda;
                        vb[P][j] = (brow < kw) ? col[brow] : T(0);
                    }
                };
                auto store = [&](auto pc) {
                    constexpr int P = decltype(pc)::value;
                    T* SA = smem + P * C::buf;
                    T* SB = SA + KC * LSA;
#pragma unroll
                    for (int j = 0; j < KC; ++j) SA[j * LSA + t] = va[P][j];
#pragma unroll
                    for (int j = 0; j < BV; ++j) SB[(bc + j) * LSB + brow] = vb[P][j];
                };
                auto multiply = [&](auto pc) {
                    constexpr int P = decltype(pc)::value;
                    const T* SA = smem + P * C::buf;
                    const T* SB = SA + KC * LSA;
#pragma unroll
                    for (int kk = 0; kk < KC; kk += 4) {
                        T a[CTW];
#pragma unroll
                        for (int ct = 0; ct < CTW; ++ct) a[ct] = -SB[(kk + lk) * LSB + ct * 16 + lr];
#pragma unroll
                        for (int u = 0; u < MTH; ++u) {
                            const int rt = wave + 4 * u;
                            if (r0 + rt * 16 < m) {
                                const T b = SA[(kk + lk) * LSA + rt * 16 + lr];
#pragma unroll
                                for (int ct = 0; ct < CTW; ++ct)
                                    if (ct * 16 < kw) acc[ct][u] = Mma<T>::mma(a[ct], b, acc[ct][u]);
                            }
                        }
                    }
                };
                const int nch = K0 / KC;  // K0 is a multiple of W, so at least 8 chunks
                load(std::integral_constant<int, 0>{}, 0);
                load(std::integral_constant<int, 1>{}, 1);
                store(std::integral_constant<int, 0>{});
                __syncthreads();
                // step c: chunk c+2 into registers, chunk c multiplied from buffer c & 1,
                // chunk c+1 to the other buffer (multiplied by every wave at step c-1)
                auto step = [&](auto pc, int c) {
                    if (c + 2 < nch) load(pc, c + 2);
                    multiply(pc);
                    if (c + 1 < nch) store(std::integral_constant<int, 1 - decltype(pc)::value>{});
                    __syncthreads();
                };
                for (int c = 0; c < nch; c += 2) {
                    step(std::integral_constant<int, 0>{}, c);
                    if (c + 1 < nch) step(std::integral_constant<int, 1>{}, c + 1);
                }
                // (the last step's barrier: the chunk area becomes L11)
            }
            // the diagonal block's rows to LDS; a pass below the last one back to A, through
            // an opaque copy of the panel pointer (the compiler would otherwise keep the 64
            // load addresses live across the product instead of recomputing them)
            gmem<T>* Aw = Ap;
            asm volatile("" : "+s"(Aw));
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                for (int u = 0; u < MTH; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = r0 + (wave + 4 * u) * 16 + lr;
                        const int col = ct * 16 + Mma<T>::drow(lane, r);
                        if (row < W) D[row * LDD + col] = acc[ct][u][r];
                        else if (p > 0 && row < m && col < kw) Aw[row + (size_t)col * lda] = acc[ct][u][r];
                    }
        }
        __syncthreads();
        // (2) L11 in 16-column steps: wave 0 factors the 16 x 16 diagonal block (lane i holds
        // row i) and inverts it into M16 (lane j forms column j of the inverse by forward
        // substitution); all four inverses are kept when rows below exist (they make the
        // inverse of L11)
        const bool below = m > kw;  // then kw == W
        int fail = -1;
        for (int c = 0; c < kw; c += 16) {
            const int cb = min(16, kw - c);
            const bool last = c + 16 >= kw;
            T* Mb = M16 + (c >> 4) * 16 * L16;
            if (wave == 0) {
                T rw[16], dinv[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) rw[j] = (lane < cb && j < cb) ? D[(c + lane) * LDD + c + j] : T(0);
                int f = -1;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    dinv[j] = T(0);
                    if (j < cb && f < 0) {
                        const T ajj = bcast(rw[j], j);
                        if (ajj <= T(0)) {
                            f = j;
                        } else {
                            const T alpha = fsqrt(ajj);
                            const T inv = T(1) / alpha;
                            dinv[j] = inv;
                            rw[j] = (lane == j) ? alpha : ((lane > j) ? rw[j] * inv : rw[j]);
                            const T xi = rw[j];
#pragma unroll
                            for (int l = j + 1; l < 16; ++l) rw[l] = __builtin_fma(xi, -bcast(rw[j], l), rw[l]);
                        }
                    }
                }
                if (lane < cb)
#pragma unroll
                    for (int j = 0; j < 16; ++j) D[(c + lane) * LDD + c + j] = rw[j];
                if (f < 0 && (!last || below)) {
                    // y = inv(L_cc) e_lane: y_k = (e_k - sum_{i<k} l_ki y_i) / l_kk
                    T y[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        T sk = (k == lane) ? T(1) : T(0);
#pragma unroll
                        for (int i = 0; i < k; ++i) sk = __builtin_fma(-bcast(rw[i], k), y[i], sk);
                        y[k] = sk * dinv[k];
                    }
                    if (lane < 16)
#pragma unroll
                        for (int k = 0; k < 16; ++k) Mb[lane * L16 + k] = y[k];
                }
                if (lane == 0) s_fail = f;
            }
            __syncthreads();
            if (s_fail >= 0) {
                fail = c + s_fail;
                break;
            }
            if (last) break;
            const int nt = (kw + 15) >> 4, c1 = (c >> 4) + 1;
            {
                // the block's rows below inside L11: X = L(rows, c:c+16) inv(L_cc)^T
                const int ti = c1 + wave;
                if (ti < nt) {
                    v4 x = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int kk = 0; kk < 16; kk += 4)
                        x = Mma<T>::mma(D[(ti * 16 + lr) * LDD + c + kk + lk], Mb[(kk + lk) * L16 + lr], x);
#pragma unroll
                    for (int r = 0; r < 4; ++r) D[(ti * 16 + Mma<T>::drow(lane, r)) * LDD + c + lr] = x[r];
                }
            }
            __syncthreads();
            {
                // the rest of L11: tile (i, j), c1 <= j <= i < nt, minus X_i X_j^T
                int pair = 0;
                for (int i = c1; i < nt; ++i)
                    for (int j = c1; j <= i; ++j, ++pair) {
                        if ((pair & 3) != wave) continue;
                        v4 a4;
#pragma unroll
                        for (int r = 0; r < 4; ++r) a4[r] = D[(i * 16 + Mma<T>::drow(lane, r)) * LDD + j * 16 + lr];
#pragma unroll
                        for (int kk = 0; kk < 16; kk += 4)
                            a4 = Mma<T>::mma(-D[(i * 16 + lr) * LDD + c + kk + lk],
                                             D[(j * 16 + lr) * LDD + c + kk + lk], a4);
#pragma unroll
                        for (int r = 0; r < 4; ++r) D[(i * 16 + Mma<T>::drow(lane, r)) * LDD + j * 16 + lr] = a4[r];
                    }
            }
            __syncthreads();
        }
        // L11's lower triangle to A
        for (int idx = t; idx < W * W; idx += BNT) {
            const int i = idx & (W - 1), c = idx / W;
            if (i < kw && c < kw && i >= c) Ap[i + (size_t)c * lda] = D[i * LDD + c];
        }
        if (fail >= 0) {
            // NonHPDMatrixException at global column K0 + fail. A then holds the finished steps
            // of L11, the rest of the panel as updated by the earlier panels.
            if (below && K0 > 0) {
#pragma unroll
                for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                    for (int u = 1; u < MTH; ++u)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = (wave + 4 * u) * 16 + lr;
                            if (row < m) Ap[row + (size_t)(ct * 16 + Mma<T>::drow(lane, r)) * lda] = acc[ct][u][r];
                        }
            }
            if (t == 0) *info = K0 + fail + 1;
            __syncthreads();
            return;
        }
        if (below) {
            // (3) inv(L11): wave J forms block column J from the block inverses,
            // X_JJ = inv(L_JJ), X_IJ = -inv(L_II) sum_{K=J}^{I-1} L_IK X_KJ; X[I] holds the
            // entries (16I + drow(lane, r), 16J + lr), the operand layout of the next product
            static_for<0, CTW>([&](auto jc) {
                constexpr int J = decltype(jc)::value;
                if (wave != J) return;
                v4 X[CTW];
#pragma unroll
                for (int r = 0; r < 4; ++r) X[J][r] = M16[J * 16 * L16 + lr * L16 + Mma<T>::drow(lane, r)];
                static_for<J + 1, CTW>([&](auto ic) {
                    constexpr int I = decltype(ic)::value;
                    v4 s = v4{T(0), T(0), T(0), T(0)};
                    static_for<J, I>([&](auto kc) {
                        constexpr int K = decltype(kc)::value;
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            s = Mma<T>::mma(D[(16 * I + lr) * LDD + 16 * K + Mma<T>::drow(lane, r)], X[K][r], s);
                    });
                    v4 x = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        x = Mma<T>::mma(-M16[I * 16 * L16 + Mma<T>::drow(lane, r) * L16 + lr], s[r], x);
                    X[I] = x;
                });
                static_for<0, CTW>([&](auto ic) {
                    constexpr int I = decltype(ic)::value;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        T v = T(0);
                        if constexpr (I >= J) v = X[I][r];
                        Minv[(16 * J + lr) * LDD + 16 * I + Mma<T>::drow(lane, r)] = v;
                    }
                });
            });
            __syncthreads();
            // (4) rows below: X^T = inv(L11) P21^T, the panel's transposed tiles as the second
            // operand; first the last pass's rows, from the registers
#pragma unroll
            for (int u = 1; u < MTH; ++u) {
                const int rt = wave + 4 * u;
                if (rt * 16 >= m) continue;
                v4 x[CTW];
#pragma unroll
                for (int ct = 0; ct < CTW; ++ct) x[ct] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int kt = 0; kt < CTW; ++kt)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int ct = 0; ct < CTW; ++ct)
                            x[ct] = Mma<T>::mma(Minv[(kt * 16 + Mma<T>::drow(lane, r)) * LDD + ct * 16 + lr],
                                                acc[kt][u][r], x[ct]);
                const int row = rt * 16 + lr;
                if (row < m)
#pragma unroll
                    for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            Ap[row + (size_t)(ct * 16 + Mma<T>::drow(lane, r)) * lda] = x[ct][r];
            }
            // then the earlier pass's rows (n > 256 + K0), from A
            for (int rt = H / 16 + wave; rt * 16 < m; rt += 4) {
                const int row = rt * 16 + lr;
                T bv[CTW][4];
#pragma unroll
                for (int kt = 0; kt < CTW; ++kt)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        bv[kt][r] = (row < m) ? Ap[row + (size_t)(kt * 16 + Mma<T>::drow(lane, r)) * lda] : T(0);
                v4 x[CTW];
#pragma unroll
                for (int ct = 0; ct < CTW; ++ct) x[ct] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int kt = 0; kt < CTW; ++kt)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int ct = 0; ct < CTW; ++ct)
                            x[ct] = Mma<T>::mma(Minv[(kt * 16 + Mma<T>::drow(lane, r)) * LDD + ct * 16 + lr],
                                                bv[kt][r], x[ct]);
                if (row < m)
#pragma unroll
                    for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            Ap[row + (size_t)(ct * 16 + Mma<T>::drow(lane, r)) * lda] = x[ct][r];
            }
        }
        __syncthreads();
    }
}

// Workgroups per CU a bin's kernel is compiled for (registers); LDS may allow fewer: the
// n <= 512 bin fits two with 16-column blocks (72 KB each in FP64), one with 32.
template <int NM, int NB>
constexpr int bin_occ_max() {
    return NM <= 128 ? 3 : (NM <= 256 ? 2 : (NB == 16 ? 2 : 1));
}

// Persistent workgroups of one size bin: items[begin, end) (orders <= NM), pulled through the
// bin's queue word.
template <typename T, int NM, int NB>
__global__ __launch_bounds__(BNT, (bin_occ_max<NM, NB>())) void potrf_vbatched_kernel(
    const BatchItem* __restrict__ items, int begin, int end, T* const* __restrict__ As,
    int32_t* __restrict__ info, int* queue) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);  // NM x (NB+1)
    T* Minv = P + NM * (NB + 1);            // NB x (NB+1)
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = begin + atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= end) break;
        const BatchItem it = items[i];
        if (threadIdx.x == 0) info[it.index] = 0;
        if (it.n > 0) potrf_one<T, NM, NB>(it.n, (gmem<T>*)As[it.index], it.lda, info + it.index, P, Minv);
    }
}

// Register-direct persistent workgroups of one size bin (variant kind 1), compiled for two
// workgroups per CU.
constexpr int kRdOcc = 2;
template <typename T, int NM, int NB>
__global__ __launch_bounds__(BNT, kRdOcc) void potrf_rd_vbatched_kernel(
    const BatchItem* __restrict__ items, int begin, int end, T* const* __restrict__ As,
    int32_t* __restrict__ info, int* queue) {
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = begin + atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= end) break;
        const BatchItem it = items[i];
        if (threadIdx.x == 0) info[it.index] = 0;
        if (it.n > 0) potrf_rd_one<T, NM, NB>(it.n, (gmem<T>*)As[it.index], it.lda, info + it.index);
    }
}

// Wide-panel persistent workgroups of one size bin (variant kind 2), two per CU.
constexpr int kWideOcc = 2;
template <typename T, int NM>
__global__ __launch_bounds__(BNT, kWideOcc) void potrf_wide_vbatched_kernel(
    const BatchItem* __restrict__ items, int begin, int end, T* const* __restrict__ As,
    int32_t* __restrict__ info, int* queue) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = begin + atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= end) break;
        const BatchItem it = items[i];
        if (threadIdx.x == 0) info[it.index] = 0;
        if (it.n > 0) potrf_wide_one<T, NM>(it.n, (gmem<T>*)As[it.index], it.lda, info + it.index, sm);
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One workgroup: L L^T X = B for one matrix (L from potrf), up to NR right-hand sides at a time
// in LDS, blocked by SB = 32 rows. Forward (right-looking): wave 0 solves the diagonal block
// with its rows in registers, then every row below subtracts the block's products (a row per
// thread). Backward: each block column's dot products with the solved part below (a wave per
// column, lanes down the column), then wave 0 solves the transposed diagonal block with its
// columns in registers.
constexpr int NR = 8, SB = 32, SLD = SB + 1;
template <typename T>
__device__ __forceinline__ void potrs_one(int n, const gmem<T>* L, int lda, gmem<T>* B, int ldb,
                                          int nrhs, T* Y, T* Part, T* Lb) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int nblk = (n + SB - 1) / SB;
    auto stage_l11 = [&](int k0, int kb) {
        for (int idx = t; idx < kb * kb; idx += BNT) {
            const int i = idx % kb, j = idx / kb;
            Lb[i * SLD + j] = (i >= j) ? L[(k0 + i) + (size_t)(k0 + j) * lda] : T(0);
        }
    };
    for (int c0 = 0; c0 < nrhs; c0 += NR) {
        const int nr = min(NR, nrhs - c0);
        for (int idx = t; idx < n * nr; idx += BNT) {
            const int i = idx % n, c = idx / n;
            Y[c * BNMAX + i] = B[i + (size_t)(c0 + c) * ldb];
        }
        __syncthreads();
        // forward: L y = b (dtrsm Left Lower NoTrans NonUnit), right-looking by 32-row blocks:
        // the diagonal block's solve, then its products subtracted from every row below (a
        // row per thread: the block's 32 columns of L are independent loads, coalesced across
        // the threads)
        for (int b = 0; b < nblk; ++b) {
            const int k0 = b * SB, kb = min(SB, n - k0);
            stage_l11(k0, kb);
            __syncthreads();
            if (wave == 0) {
                T lrow[SB], yv[NR];
#pragma unroll
                for (int j = 0; j < SB; ++j) lrow[j] = (lane < kb) ? Lb[lane * SLD + j] : T(0);
#pragma unroll
                for (int q = 0; q < NR; ++q) yv[q] = (lane < kb && q < nr) ? Y[q * BNMAX + k0 + lane] : T(0);
                // reciprocals of the diagonal, one division per lane, broadcast per step
                const T dinv = (lane < kb) ? T(1) / Lb[lane * SLD + lane] : T(0);
#pragma unroll
                for (int k = 0; k < SB; ++k) {
                    if (k < kb) {
                        const T dk = bcast(dinv, k);
                        // every column slot, branch-free (slots >= nr hold zeros): a guard per
                        // slot would split the step chain into small branch blocks
#pragma unroll
                        for (int q = 0; q < NR; ++q) {
                            if (lane == k) yv[q] = yv[q] * dk;
                            const T yk = bcast(yv[q], k);
                            if (lane > k) yv[q] = __builtin_fma(-lrow[k], yk, yv[q]);
                        }
                    }
                }
                if (lane < kb)
#pragma unroll
                    for (int q = 0; q < NR; ++q)
                        if (q < nr) Y[q * BNMAX + k0 + lane] = yv[q];
            }
            __syncthreads();
            for (int i = k0 + kb + t; i < n; i += BNT) {
                T acc[NR];
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = (q < nr) ? Y[q * BNMAX + i] : T(0);
                const gmem<T>* li = L + i + (size_t)k0 * lda;
#pragma unroll 8
                for (int k = 0; k < kb; ++k) {
                    const T l = li[(size_t)k * lda];
#pragma unroll
                    for (int q = 0; q < NR; ++q)
                        if (q < nr) acc[q] = __builtin_fma(-l, Y[q * BNMAX + k0 + k], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nr) Y[q * BNMAX + i] = acc[q];
            }
            __syncthreads();
        }
        // backward: L^T x = y (dtrsm Left Lower Trans NonUnit), blocks from the bottom
        for (int b = nblk - 1; b >= 0; --b) {
            const int k0 = b * SB, kb = min(SB, n - k0), r0 = k0 + kb;
            for (int j = wave; j < kb; j += 4) {
                T acc[NR];
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = T(0);
                const gmem<T>* col = L + (size_t)(k0 + j) * lda;
#pragma unroll 2
                for (int i = r0 + lane; i < n; i += 64) {
                    const T l = col[i];
#pragma unroll
                    for (int q = 0; q < NR; ++q)
                        if (q < nr) acc[q] = __builtin_fma(l, Y[q * BNMAX + i], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = wave_sum(acc[q]);
                if (lane == 0)
#pragma unroll
                    for (int q = 0; q < NR; ++q) Part[j * NR + q] = acc[q];
            }
            stage_l11(k0, kb);
            __syncthreads();
            if (wave == 0) {
                // lane k holds column k of L11: L(i, k) for every i
                T lcol[SB], xv[NR];
#pragma unroll
                for (int i = 0; i < SB; ++i) lcol[i] = (lane < kb) ? Lb[i * SLD + lane] : T(0);
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    xv[q] = (lane < kb && q < nr) ? Y[q * BNMAX + k0 + lane] - Part[lane * NR + q] : T(0);
                const T dinv = (lane < kb) ? T(1) / Lb[lane * SLD + lane] : T(0);
#pragma unroll
                for (int i = SB - 1; i >= 0; --i) {
                    if (i < kb) {
                        const T di = bcast(dinv, i);
#pragma unroll
                        for (int q = 0; q < NR; ++q) {
                            if (lane == i) xv[q] = xv[q] * di;
                            const T xi = bcast(xv[q], i);
                            if (lane < i) xv[q] = __builtin_fma(-lcol[i], xi, xv[q]);
                        }
                    }
                }
                if (lane < kb)
#pragma unroll
                    for (int q = 0; q < NR; ++q)
                        if (q < nr) Y[q * BNMAX + k0 + lane] = xv[q];
            }
            __syncthreads();
        }
        for (int idx = t; idx < n * nr; idx += BNT) {
            const int i = idx % n, c = idx / n;
            B[i + (size_t)(c0 + c) * ldb] = Y[c * BNMAX + i];
        }
        __syncthreads();
    }
}

// info (may be null): matrices whose factorization failed are left unsolved.
template <typename T>
__global__ __launch_bounds__(BNT, 2) void potrs_vbatched_kernel(
    const BatchItem* __restrict__ items, int count, const T* const* __restrict__ Ls,
    T* const* __restrict__ Bs, const int32_t* __restrict__ info, int* queue) {
    __shared__ T Y[NR * BNMAX];
    __shared__ T Part[SB * NR];
    __shared__ T Lb[SB * SLD];
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= count) break;
        const BatchItem it = items[i];
        if (info && info[it.index] != 0) continue;
        if (it.n > 0 && it.nrhs > 0)
            potrs_one<T>(it.n, (const gmem<T>*)Ls[it.index], it.lda, (gmem<T>*)Bs[it.index], it.ldb,
                         it.nrhs, Y, Part, Lb);
    }
}

__global__ void zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------------- host side
// A schedule: the valid items in decreasing order of n (ties by index) on the device, and
// the item ranges of the size bins.
struct Schedule {
    std::vector<int32_t> key;
    BatchItem* dev = nullptr;
    BatchItem* host = nullptr;  // pinned staging (cached schedules)
    size_t cap = 0;
    int valid = 0;
    int bin_end[NBIN] = {0, 0, 0};  // bin_end[b]: number of items of order > kBinMax[b-1]
    hipEvent_t copied = nullptr;
    double flops = 0.0, bytes = 0.0;
};

// Items of bin b (n in (kBinMax[b-1], kBinMax[b]]) in the descending order: [lo, hi).
void bin_range(const Schedule& sc, int b, int* lo, int* hi) {
    // bins in the sorted order: 512, 256, 128
    *lo = (b == NBIN - 1) ? 0 : sc.bin_end[b + 1];
    *hi = sc.bin_end[b];
}

// Per-device cached schedules (factor, solve) and work queue, kept between calls: a call with
// the same sizes and leading dimensions as the previous one of its kind reuses the device
// schedule without sorting or copying. The host staging buffer is pinned, and its previous
// copy is waited for before it is refilled.
struct BatchArea {
    Schedule sched[2];  // 0: factor, 1: solve
    int* queue = nullptr;
};
std::mutex g_bmu;
std::map<int, BatchArea> g_batch;

// Variant overrides set by elx_batched_set_variant (the autotuner); nb 0 = the table's.
elx::tuned::BatchedVariant g_override[2][NBIN] = {};

elx_status_t batch_area(BatchArea** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_bmu);
    BatchArea& a = g_batch[dev];
    if (!a.queue) ELX_HIP(hipMalloc(&a.queue, sizeof(int) * kQueueWords));
    *out = &a;
    return ELX_SUCCESS;
}

void free_schedule(Schedule& sc) {
    if (sc.copied) (void)hipEventSynchronize(sc.copied);
    if (sc.dev) (void)hipFree(sc.dev);
    if (sc.host) (void)hipHostFree(sc.host);
    if (sc.copied) (void)hipEventDestroy(sc.copied);
    sc = Schedule();
}

}  // namespace

namespace elx {
// Frees every batch schedule (device and pinned host copies, their events) and work queue.
void batched_release() {
    std::lock_guard<std::mutex> lk(g_bmu);
    for (auto& kv : g_batch) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(kv.first);
        BatchArea& a = kv.second;
        for (Schedule& sc : a.sched) free_schedule(sc);
        if (a.queue) (void)hipFree(a.queue);
        a.queue = nullptr;
        (void)hipSetDevice(cur);
    }
    g_batch.clear();
}
}  // namespace elx

namespace {

// Sorts the valid items, fills sc (device copy through the pinned buffer when `cached`,
// else a synchronous copy from a temporary host vector).
elx_status_t schedule(hipStream_t s, Schedule& sc, int64_t count, const int32_t* n,
                      const int32_t* lda, const int32_t* nrhs, const int32_t* ldb, size_t tsize,
                      bool cached) {
    std::vector<int32_t> key;
    key.reserve((size_t)count * 4);
    for (int64_t i = 0; i < count; ++i) {
        ELX_ARG(n[i] >= 0 && n[i] <= BNMAX && lda[i] >= std::max(n[i], 1));
        ELX_ARG(!nrhs || (nrhs[i] >= 0 && ldb[i] >= std::max(n[i], 1)));
        key.push_back(n[i]);
        key.push_back(lda[i]);
        key.push_back(nrhs ? nrhs[i] : 0);
        key.push_back(ldb ? ldb[i] : 0);
    }
    if (cached && sc.dev && key == sc.key) return ELX_SUCCESS;
    if (sc.copied) ELX_HIP(hipEventSynchronize(sc.copied));
    if (sc.cap < (size_t)count) {
        if (sc.dev) {
            ELX_HIP(hipDeviceSynchronize());
            ELX_HIP(hipFree(sc.dev));
            if (sc.host) ELX_HIP(hipHostFree(sc.host));
            sc.dev = nullptr;
            sc.host = nullptr;
        }
        const size_t cap = std::max<size_t>((size_t)count, cached ? 1024 : 1);
        ELX_HIP(hipMalloc(&sc.dev, sizeof(BatchItem) * cap));
        if (cached)
            ELX_HIP(hipHostMalloc(reinterpret_cast<void**>(&sc.host), sizeof(BatchItem) * cap, 0));
        sc.cap = cap;
    }
    std::vector<int32_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return n[x] > n[y]; });
    std::vector<BatchItem> tmp;
    BatchItem* h = sc.host;
    if (!cached) {
        tmp.resize((size_t)count);
        h = tmp.data();
    }
    sc.flops = 0.0;
    sc.bytes = 0.0;
    for (int b = 0; b < NBIN; ++b) sc.bin_end[b] = 0;
    for (int64_t k = 0; k < count; ++k) {
        const int32_t i = order[(size_t)k];
        BatchItem& it = h[k];
        it.n = n[i];
        it.lda = lda[i];
        it.index = i;
        it.nrhs = nrhs ? nrhs[i] : 0;
        it.ldb = ldb ? ldb[i] : 0;
        it.pad = 0;
        const double nn = (double)n[i];
        sc.flops += nn * nn * nn / 3.0;
        sc.bytes += nn * nn * tsize + 2.0 * nn * (double)it.nrhs * tsize;
        // in the descending order, the items of order > kBinMax[b-1] are a prefix
        for (int b = 0; b < NBIN; ++b)
            if (n[i] > (b == 0 ? -1 : kBinMax[b - 1])) sc.bin_end[b] = (int)(k + 1);
    }
    if (cached) {
        if (!sc.copied) ELX_HIP(hipEventCreateWithFlags(&sc.copied, hipEventDisableTiming));
        ELX_HIP(hipMemcpyAsync(sc.dev, sc.host, sizeof(BatchItem) * (size_t)count,
                               hipMemcpyHostToDevice, s));
        ELX_HIP(hipEventRecord(sc.copied, s));
    } else if (count > 0) {
        ELX_HIP(hipMemcpy(sc.dev, h, sizeof(BatchItem) * (size_t)count, hipMemcpyHostToDevice));
    }
    sc.key.swap(key);
    sc.valid = (int)count;
    return ELX_SUCCESS;
}

template <typename T>
elx::tuned::BatchedVariant variant_of(int bin) {
    const int d = sizeof(T) == 8 ? 0 : 1;
    const elx::tuned::BatchedVariant o = g_override[d][bin];
    return o.nb ? o : elx::tuned::kBatchedPotrf[d][bin];
}

template <typename T, int NM, int NB>
elx_status_t launch_bin(hipStream_t s, const BatchItem* items, int lo, int hi, T* const* dA,
                        int32_t* dinfo, int* queue, int occ) {
    const void* fn = reinterpret_cast<const void*>(&potrf_vbatched_kernel<T, NM, NB>);
    const size_t lds = sizeof(T) * ((size_t)NM * (NB + 1) + (size_t)NB * (NB + 1));
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
        attr = true;
    }
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BNT, lds));
    per_cu = std::max(1, std::min(per_cu, occ));
    const elx::DeviceInfo& di = elx::device_info();
    const int G = (int)std::min<int64_t>(hi - lo, (int64_t)per_cu * di.num_cus);
    hipLaunchKernelGGL((potrf_vbatched_kernel<T, NM, NB>), dim3(G), dim3(BNT), lds, s, items, lo,
                       hi, dA, dinfo, queue);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T, int NM, int NB>
elx_status_t launch_rd_bin(hipStream_t s, const BatchItem* items, int lo, int hi, T* const* dA,
                           int32_t* dinfo, int* queue, int occ) {
    const void* fn = reinterpret_cast<const void*>(&potrf_rd_vbatched_kernel<T, NM, NB>);
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BNT, 0));
    per_cu = std::max(1, std::min(per_cu, occ));
    const elx::DeviceInfo& di = elx::device_info();
    const int G = (int)std::min<int64_t>(hi - lo, (int64_t)per_cu * di.num_cus);
    hipLaunchKernelGGL((potrf_rd_vbatched_kernel<T, NM, NB>), dim3(G), dim3(BNT), 0, s, items, lo,
                       hi, dA, dinfo, queue);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T, int NM>
elx_status_t launch_wide_bin(hipStream_t s, const BatchItem* items, int lo, int hi, T* const* dA,
                             int32_t* dinfo, int* queue, int occ) {
    const void* fn = reinterpret_cast<const void*>(&potrf_wide_vbatched_kernel<T, NM>);
    const size_t lds = Wide<T>::lds;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
        attr = true;
    }
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BNT, lds));
    per_cu = std::max(1, std::min(per_cu, occ));
    const elx::DeviceInfo& di = elx::device_info();
    const int G = (int)std::min<int64_t>(hi - lo, (int64_t)per_cu * di.num_cus);
    hipLaunchKernelGGL((potrf_wide_vbatched_kernel<T, NM>), dim3(G), dim3(BNT), lds, s, items, lo,
                       hi, dA, dinfo, queue);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t launch_factor(hipStream_t s, const Schedule& sc, T* const* dA, int32_t* dinfo,
                           int* queue) {
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, queue + kQueueFactor, NBIN);
    ELX_HIP(hipGetLastError());
    elx::ProfScope prof_(s, elx::kKBatched, sc.flops);
    for (int b = NBIN - 1; b >= 0; --b) {  // largest orders first
        int lo = 0, hi = 0;
        bin_range(sc, b, &lo, &hi);
        if (hi <= lo) continue;
        const elx::tuned::BatchedVariant v = variant_of<T>(b);
        int* q = queue + kQueueFactor + b;
        const BatchItem* it = sc.dev;
        if (b == 0) {
            if (v.nb == 16) ELX_CHECK((launch_bin<T, 128, 16>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else ELX_CHECK((launch_bin<T, 128, 32>(s, it, lo, hi, dA, dinfo, q, v.occ)));
        } else if (b == 1) {
            if (v.kind == 2)
                ELX_CHECK((launch_wide_bin<T, 256>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else if (v.kind == 1 && v.nb == 16)
                ELX_CHECK((launch_rd_bin<T, 256, 16>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else if (v.kind == 1)
                ELX_CHECK((launch_rd_bin<T, 256, 32>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else if (v.nb == 16)
                ELX_CHECK((launch_bin<T, 256, 16>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else
                ELX_CHECK((launch_bin<T, 256, 32>(s, it, lo, hi, dA, dinfo, q, v.occ)));
        } else {
            if (v.kind == 2)
                ELX_CHECK((launch_wide_bin<T, 512>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else if (v.kind == 1 && v.nb == 16)
                ELX_CHECK((launch_rd_bin<T, 512, 16>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else if (v.kind == 1)
                ELX_CHECK((launch_rd_bin<T, 512, 32>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else if (v.nb == 16)
                ELX_CHECK((launch_bin<T, 512, 16>(s, it, lo, hi, dA, dinfo, q, v.occ)));
            else
                ELX_CHECK((launch_bin<T, 512, 32>(s, it, lo, hi, dA, dinfo, q, v.occ)));
        }
    }
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t launch_solve(hipStream_t s, const Schedule& sc, const T* const* dA, T* const* dB,
                          const int32_t* dinfo, int* queue) {
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, queue + kQueueSolve, 1);
    ELX_HIP(hipGetLastError());
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&potrs_vbatched_kernel<T>), BNT, 0));
    const elx::DeviceInfo& di = elx::device_info();
    const int G = (int)std::min<int64_t>(sc.valid, (int64_t)std::max(per_cu, 1) * di.num_cus);
    elx::ProfScope prof_(s, elx::kKBatchedSolve, sc.bytes);
    hipLaunchKernelGGL(potrs_vbatched_kernel<T>, dim3(G), dim3(BNT), 0, s, sc.dev, sc.valid, dA, dB,
                       dinfo, queue + kQueueSolve);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t potrf_vbatched(hipStream_t s, int uplo, int64_t count, const int32_t* n,
                            T* const* dA, const int32_t* lda, int32_t* dinfo) {
    ELX_ARG(uplo == elx::kLower);
    ELX_ARG(count >= 0 && count < (1ll << 30));
    if (count == 0) return ELX_SUCCESS;
    ELX_ARG(n && lda && dA && dinfo);
    BatchArea* a = nullptr;
    ELX_CHECK(batch_area(&a));
    ELX_CHECK(schedule(s, a->sched[0], count, n, lda, nullptr, nullptr, sizeof(T), true));
    return launch_factor<T>(s, a->sched[0], dA, dinfo, a->queue);
}

template <typename T>
elx_status_t potrs_vbatched(hipStream_t s, int uplo, int64_t count, const int32_t* n,
                            const int32_t* nrhs, const T* const* dA, const int32_t* lda,
                            T* const* dB, const int32_t* ldb) {
    ELX_ARG(uplo == elx::kLower);
    ELX_ARG(count >= 0 && count < (1ll << 30));
    if (count == 0) return ELX_SUCCESS;
    ELX_ARG(n && nrhs && lda && ldb && dA && dB);
    BatchArea* a = nullptr;
    ELX_CHECK(batch_area(&a));
    ELX_CHECK(schedule(s, a->sched[1], count, n, lda, nrhs, ldb, sizeof(T), true));
    return launch_solve<T>(s, a->sched[1], dA, dB, nullptr, a->queue);
}

}  // namespace

// A captured factor + solve: its own schedules and queue words, so other batched calls can
// run between replays.
struct elx_graph_st {
    Schedule sched;
    int* queue = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
};

namespace {

template <typename T>
elx_status_t posv_graph(hipStream_t s, int uplo, int64_t count, const int32_t* n,
                        const int32_t* nrhs, T* const* dA, const int32_t* lda, T* const* dB,
                        const int32_t* ldb, int32_t* dinfo, elx_graph_t* out) {
    ELX_ARG(out != nullptr);
    *out = nullptr;
    ELX_ARG(uplo == elx::kLower);
    ELX_ARG(count >= 1 && count < (1ll << 30));
    ELX_ARG(n && nrhs && lda && ldb && dA && dB && dinfo);
    elx_graph_st* g = new elx_graph_st();
    auto fail = [&](elx_status_t st) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        free_schedule(g->sched);
        if (g->queue) (void)hipFree(g->queue);
        delete g;
        return st;
    };
    if (hipGetDevice(&g->device) != hipSuccess) return fail(ELX_ERROR_HIP);
    // every allocation and copy happens before the capture
    elx_status_t st = schedule(s, g->sched, count, n, lda, nrhs, ldb, sizeof(T), false);
    if (st != ELX_SUCCESS) return fail(st);
    if (hipMalloc(&g->queue, sizeof(int) * kQueueWords) != hipSuccess) return fail(ELX_ERROR_ALLOC);
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return fail(ELX_ERROR_HIP);
    hipGraph_t graph = nullptr;
    st = ELX_SUCCESS;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) {
        st = ELX_ERROR_HIP;
    } else {
        st = launch_factor<T>(cs, g->sched, dA, dinfo, g->queue);
        if (st == ELX_SUCCESS)
            st = launch_solve<T>(cs, g->sched, const_cast<const T* const*>(dA), dB, dinfo, g->queue);
        if (hipStreamEndCapture(cs, &graph) != hipSuccess && st == ELX_SUCCESS) st = ELX_ERROR_HIP;
    }
    if (st == ELX_SUCCESS && hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) != hipSuccess)
        st = ELX_ERROR_HIP;
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    if (st != ELX_SUCCESS) return fail(st);
    *out = g;
    return ELX_SUCCESS;
}

}  // namespace

extern "C" {

elx_status_t elx_dpotrf_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 double* const* dA_array, const int32_t* lda, int32_t* dinfo) {
    return potrf_vbatched<double>(static_cast<hipStream_t>(stream), uplo, count, n, dA_array, lda,
                                  dinfo);
}

elx_status_t elx_spotrf_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 float* const* dA_array, const int32_t* lda, int32_t* dinfo) {
    return potrf_vbatched<float>(static_cast<hipStream_t>(stream), uplo, count, n, dA_array, lda,
                                 dinfo);
}

elx_status_t elx_dpotrs_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 const int32_t* nrhs, const double* const* dA_array,
                                 const int32_t* lda, double* const* dB_array, const int32_t* ldb) {
    return potrs_vbatched<double>(static_cast<hipStream_t>(stream), uplo, count, n, nrhs, dA_array,
                                  lda, dB_array, ldb);
}

elx_status_t elx_spotrs_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 const int32_t* nrhs, const float* const* dA_array,
                                 const int32_t* lda, float* const* dB_array, const int32_t* ldb) {
    return potrs_vbatched<float>(static_cast<hipStream_t>(stream), uplo, count, n, nrhs, dA_array,
                                 lda, dB_array, ldb);
}

elx_status_t elx_dposv_vbatched_graph(void* stream, int uplo, int64_t count, const int32_t* n,
                                      const int32_t* nrhs, double* const* dA_array,
                                      const int32_t* lda, double* const* dB_array,
                                      const int32_t* ldb, int32_t* dinfo, elx_graph_t* graph) {
    return posv_graph<double>(static_cast<hipStream_t>(stream), uplo, count, n, nrhs, dA_array, lda,
                              dB_array, ldb, dinfo, graph);
}

elx_status_t elx_sposv_vbatched_graph(void* stream, int uplo, int64_t count, const int32_t* n,
                                      const int32_t* nrhs, float* const* dA_array,
                                      const int32_t* lda, float* const* dB_array,
                                      const int32_t* ldb, int32_t* dinfo, elx_graph_t* graph) {
    return posv_graph<float>(static_cast<hipStream_t>(stream), uplo, count, n, nrhs, dA_array, lda,
                             dB_array, ldb, dinfo, graph);
}

elx_status_t elx_graph_launch(elx_graph_t graph, void* stream) {
    ELX_ARG(graph != nullptr && graph->exec != nullptr);
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    ELX_ARG(dev == graph->device);
    ELX_HIP(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
    return ELX_SUCCESS;
}

elx_status_t elx_graph_destroy(elx_graph_t graph) {
    if (!graph) return ELX_SUCCESS;
    ELX_HIP(hipDeviceSynchronize());
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    free_schedule(graph->sched);
    if (graph->queue) (void)hipFree(graph->queue);
    delete graph;
    return ELX_SUCCESS;
}

elx_status_t elx_batched_set_variant(int dtype, int bin, int nb, int occ, int kind) {
    ELX_ARG(dtype == 0 || dtype == 1);
    ELX_ARG(bin >= 0 && bin < NBIN);
    ELX_ARG(nb == 0 || nb == 16 || nb == 32 || nb == 64);
    ELX_ARG(nb == 0 || (occ >= 1 && occ <= 8));
    ELX_ARG(nb == 0 || (kind == 0 && nb != 64) || (kind == 1 && bin >= 1 && nb != 64) ||
            (kind == 2 && bin >= 1 && nb == 64));
    std::lock_guard<std::mutex> lk(g_bmu);
    g_override[dtype][bin] =
        elx::tuned::BatchedVariant{nb, nb ? occ : 0, nb ? kind : 0};
    return ELX_SUCCESS;
}

elx_status_t elx_batched_get_variant(int dtype, int bin, int* nb, int* occ, int* kind) {
    ELX_ARG(dtype == 0 || dtype == 1);
    ELX_ARG(bin >= 0 && bin < NBIN && nb && occ && kind);
    const elx::tuned::BatchedVariant v =
        dtype == 0 ? variant_of<double>(bin) : variant_of<float>(bin);
    *nb = v.nb;
    *occ = v.occ;
    *kind = v.kind;
    return ELX_SUCCESS;
}

}  // extern "C"
