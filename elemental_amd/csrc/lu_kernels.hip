// Row permutation, small triangular solve and unblocked Cholesky kernels.
#include <algorithm>

#include "elx_common.hpp"

namespace {

// ---------------------------------------------------------------------------------------
// Row moves: the net effect of one panel's swap sequence on the columns outside the panel
// (PB.PermuteRows(AB0) / PB.PermuteRows(AB2), src/lapack_like/factor/LU.cpp:136-137).
// One wave per column. All sources of a column are read (and staged in LDS) before any
// destination is written, so overlapping source/destination rows are safe. A move whose
// source or destination lies outside the `rows` rows below r0 is skipped, so a list that a
// panel left unwritten can never reach past the column.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int2 move_in_range(int2 p, i64 rows) {
    return (p.x >= 0 && p.y >= 0 && p.x < rows && p.y < rows) ? p : int2{-1, -1};
}

template <typename T>
__global__ __launch_bounds__(256) void row_moves_kernel(const int2* __restrict__ pairs,
                                                        int npairs, T* A, i64 lda, i64 r0,
                                                        i64 rows, i64 c0, i64 c1, i64 c2,
                                                        i64 c3, const int32_t* __restrict__ stop) {
    __shared__ T stage[4][256];
    __shared__ int2 pr[256];
    if (stop && *stop != 0) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < 256; i += 256)
        pr[i] = (i < npairs) ? move_in_range(pairs[i], rows) : int2{-1, -1};
    __syncthreads();
    const i64 n1 = c1 - c0, total = n1 + (c3 - c2);
    const i64 cidx = (i64)blockIdx.x * 4 + wave;
    const bool active = cidx < total;
    T* col = nullptr;
    if (active) {
        const i64 j = (cidx < n1) ? (c0 + cidx) : (c2 + (cidx - n1));
        col = A + r0 + j * lda;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = lane + 64 * q;
            if (pr[i].x >= 0) stage[wave][i] = col[pr[i].y];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = lane + 64 * q;
            if (pr[i].x >= 0) col[pr[i].x] = stage[wave][i];
        }
    }
}

// The same moves for an fp32 matrix stored row by row (row i contiguous at A + i*lda), as the
// mixed-precision factorization keeps its working copy: each move is a contiguous segment
// copy. A workgroup owns a 64-column slice of every moved row; wave w reads the slice of moves
// 32w .. 32w+31 (a float per lane, 256 contiguous bytes per move), and all writes follow one
// barrier, so overlapping sources and destinations are safe. Moves outside the `rows` rows
// below r0 are skipped.
__global__ __launch_bounds__(512) void row_moves_rowmajor_kernel(const int2* __restrict__ pairs,
                                                                 int npairs, float* A, i64 lda,
                                                                 i64 r0, i64 rows, i64 c0, i64 c1,
                                                                 i64 c2, i64 c3) {
    __shared__ int2 pr[256];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < 256) pr[t] = (t < npairs) ? move_in_range(pairs[t], rows) : int2{-1, -1};
    __syncthreads();
    const i64 s1 = (c1 - c0 + 63) / 64;  // slices of the first range
    const i64 sl = blockIdx.x;
    const i64 col = (sl < s1) ? (c0 + sl * 64 + lane) : (c2 + (sl - s1) * 64 + lane);
    const bool in = (sl < s1) ? (col < c1) : (col < c3);
    float v[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int2 p = pr[wave * 32 + q];
        v[q] = (in && p.x >= 0) ? A[(r0 + p.y) * lda + col] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int2 p = pr[wave * 32 + q];
        if (in && p.x >= 0) A[(r0 + p.x) * lda + col] = v[q];
    }
}

// ---------------------------------------------------------------------------------------
// Explicit permutation of a swap sequence (Permutation::ExplicitVector,
// src/lapack_like/perm/Permutation.cpp:731-745). The swaps are sequential by nature; one
// thread replays them on an LDS copy of the index vector when it fits, else in global
// memory.
// ---------------------------------------------------------------------------------------
constexpr int kSwapChunk = 1024;  // swap destinations staged in LDS per round
__global__ __launch_bounds__(256) void swaps_to_perm_kernel(const int32_t* __restrict__ dests,
                                                            i64 nswaps, i64 size, int32_t* perm,
                                                            int inverse, int in_lds) {
    extern __shared__ int32_t sp[];
    __shared__ int32_t dq[kSwapChunk];
    const int t = threadIdx.x;
    int32_t* p = in_lds ? sp : perm;
    for (i64 i = t; i < size; i += 256) p[i] = (int32_t)i;
    // The swaps are applied one after another by one thread; the whole workgroup first stages
    // each chunk of destinations (in application order) in LDS, so the dependent chain waits
    // on no global load.
    for (i64 c0 = 0; c0 < nswaps; c0 += kSwapChunk) {
        const int cn = (int)((nswaps - c0) < kSwapChunk ? (nswaps - c0) : kSwapChunk);
        __syncthreads();  // p initialised, or the previous chunk applied
        for (int u = t; u < cn; u += 256) dq[u] = dests[inverse ? (nswaps - 1 - (c0 + u)) : (c0 + u)];
        __syncthreads();
        if (t == 0) {
#pragma unroll 4
            for (int u = 0; u < cn; ++u) {
                const i64 j = inverse ? (nswaps - 1 - (c0 + u)) : (c0 + u);
                const i64 d = dq[u];
                if (d != j && d >= 0 && d < size) {
                    const int32_t a = p[j];
                    p[j] = p[d];
                    p[d] = a;
                }
            }
        }
    }
    __syncthreads();
    if (in_lds)
        for (i64 i = t; i < size; i += 256) perm[i] = p[i];
}

template <typename T>
__global__ void gather_rows_kernel(const int32_t* __restrict__ perm, i64 m, i64 n,
                                   const T* __restrict__ Bin, i64 ldin, T* __restrict__ Bout,
                                   i64 ldout) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i >= m || j >= n) return;
    Bout[i + j * ldout] = Bin[(i64)perm[i] + j * ldin];
}

// ---------------------------------------------------------------------------------------
// Triangular solves M x = b for many independent vectors with an order <= 128 triangle.
// Left side: the vectors are columns of B and M = op(A). Right side (X op(A) = B): the
// vectors are rows of B and M = op(A)^T. The vectors are staged in LDS; each vector is worked
// on by the TP lanes of one 16-lane quarter-wave, which split the rows of every update step.
// The update order is the reference BLAS dtrsm's: in forward and column-oriented backward
// substitution x_i -= M(i,j) x_j is applied for j in turn, one FMA each; for the transposed
// backward forms the reference BLAS accumulates each x_i as a dot product in ascending k,
// which one lane of the quarter-wave does alone. Division by the diagonal on the left side,
// multiplication by its reciprocal on the right side, as the reference BLAS does.
// The lanes of a vector live in the same wave and LDS operations of a wave complete in
// order, so a step reads what the previous step wrote without a workgroup barrier.
// ---------------------------------------------------------------------------------------
// Left, lower, no transpose, unit diagonal, order m <= 128, many right-hand sides: the LU's
// U12 = L11^{-1} A12 (Trsm(LEFT,LOWER,NORMAL,UNIT), src/lapack_like/factor/LU.cpp:139) and the
// row-echelon form's B update. 64 columns per workgroup in LDS, 32-row blocks: per block the
// block's column panel of L (its rows r0.., 32 columns) is staged, wave 0 solves the diagonal
// block (a column per lane, 32 values in registers), and all waves subtract its contribution
// from the rows below. Every entry still receives its updates one FMA at a time in increasing
// k -- b_i = fma(-x_k, l_ik, b_i), the column-oriented substitution's order -- so the result
// is bitwise that of trsm_small_kernel and the restated BLAS; the chain is 4 blocks of 32
// steps instead of 128 dependent steps.
// One 32-row block of trsm_llnu_kernel: wave 0 solves the diagonal block, then every wave
// subtracts its contribution from the rows below, four (then two, then one) rows per pass.
// FULL (rb == 32) drops the per-row guards, so the fully unrolled loops carry no run-time
// branch and their LDS reads are issued in batches rather than one round trip per FMA.
constexpr int kLlnuXL = 65, kLlnuPL = 33;
template <typename T, bool FULL>
__device__ __forceinline__ void llnu_block(int r0, int rb, int rows, int lane, int wave, T* Xs,
                                           const T* Ps) {
    constexpr int XL = kLlnuXL, PL = kLlnuPL;
    if (wave == 0) {
        T x[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) x[i] = (FULL || i < rb) ? Xs[(r0 + i) * XL + lane] : T(0);
        if (FULL) {
            // column k + 1's entries are read while step k's FMAs run (one wave per SIMD here:
            // nothing else hides an LDS round trip)
            T cur[32], nxt[32];
#pragma unroll
            for (int i = 1; i < 32; ++i) cur[i] = Ps[i * PL];
#pragma unroll
            for (int k = 0; k < 32; ++k) {
#pragma unroll
                for (int i = k + 2; i < 32; ++i) nxt[i] = Ps[i * PL + k + 1];
#pragma unroll
                for (int i = k + 1; i < 32; ++i) x[i] = __builtin_fma(-x[k], cur[i], x[i]);
#pragma unroll
                for (int i = k + 2; i < 32; ++i) cur[i] = nxt[i];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 32; ++k) {
#pragma unroll
                for (int i = k + 1; i < 32; ++i)
                    if (i < rb) x[i] = __builtin_fma(-x[k], Ps[i * PL + k], x[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (FULL || i < rb) Xs[(r0 + i) * XL + lane] = x[i];
    }
    __syncthreads();
    if (rb < rows) {
        T xb[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) xb[k] = (FULL || k < rb) ? Xs[(r0 + k) * XL + lane] : T(0);
        int i = rb + wave;
        for (; i + 12 < rows; i += 16) {
            T a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = Xs[(r0 + i + 4 * u) * XL + lane];
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (FULL || k < rb) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        a[u] = __builtin_fma(-xb[k], Ps[(i + 4 * u) * PL + k], a[u]);
                }
#pragma unroll
            for (int u = 0; u < 4; ++u) Xs[(r0 + i + 4 * u) * XL + lane] = a[u];
        }
        for (; i + 4 < rows; i += 8) {
            T a0 = Xs[(r0 + i) * XL + lane], a1 = Xs[(r0 + i + 4) * XL + lane];
            const T* p0 = Ps + i * PL;
            const T* p1 = p0 + 4 * PL;
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (FULL || k < rb) {
                    a0 = __builtin_fma(-xb[k], p0[k], a0);
                    a1 = __builtin_fma(-xb[k], p1[k], a1);
                }
            Xs[(r0 + i) * XL + lane] = a0;
            Xs[(r0 + i + 4) * XL + lane] = a1;
        }
        if (i < rows) {
            T acc = Xs[(r0 + i) * XL + lane];
            const T* pi = Ps + i * PL;
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (FULL || k < rb) acc = __builtin_fma(-xb[k], pi[k], acc);
            Xs[(r0 + i) * XL + lane] = acc;
        }
    }
}

// Left, lower, no transpose, unit diagonal, order m <= 128, many right-hand sides: the LU's
// U12 = L11^{-1} A12 (Trsm(LEFT,LOWER,NORMAL,UNIT), src/lapack_like/factor/LU.cpp:139) and the
// row-echelon form's B update. 64 columns per workgroup in LDS, 32-row blocks: per block the
// block's column panel of L (its rows r0.., 32 columns) is staged, wave 0 solves the diagonal
// block (a column per lane, 32 values in registers), and all waves subtract its contribution
// from the rows below. Every entry still receives its updates one FMA at a time in increasing
// k -- b_i = fma(-x_k, l_ik, b_i), the column-oriented substitution's order -- so the result
// is bitwise that of trsm_small_kernel and the restated BLAS; the chain is 4 blocks of 32
// steps instead of 128 dependent steps.
template <typename T>
__global__ __launch_bounds__(256) void trsm_llnu_kernel(int m, i64 ncols, const T* __restrict__ A,
                                                        i64 lda, T* B, i64 ldb,
                                                        const int32_t* __restrict__ stop) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int XL = kLlnuXL, PL = kLlnuPL;
    // Ps first: its entries' LDS addresses (< 34 KB) fit the 16-bit immediate offset of
    // ds_read, so the unrolled loops address them from one register instead of materialising
    // an address per entry
    T* Ps = reinterpret_cast<T*>(smem_raw);  // 128 x 33: Ps[i*PL + kk] = A(r0 + i, r0 + kk)
    T* Xs = Ps + 128 * PL;                   // 128 x 65: Xs[i*XL + c]
    if (stop && *stop != 0) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const i64 c0 = (i64)blockIdx.x * 64;
    const int nc = (int)((ncols - c0) < 64 ? (ncols - c0) : 64);
    // Staging loads are issued in unrolled batches before their LDS stores: with one wave per
    // SIMD nothing else hides a global load's latency, so a load-store-load loop would pay it
    // once per element.
    if (m == 128) {
        T v[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int idx = t + 256 * u, i = idx & 127, c = idx >> 7;  // down the columns
            v[u] = (c < nc) ? B[i + (c0 + c) * ldb] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int idx = t + 256 * u;
            Xs[(idx & 127) * XL + (idx >> 7)] = v[u];
        }
    } else {
        for (int idx = t; idx < m * 64; idx += 256) {
            const int i = idx % m, c = idx / m;  // down the columns: coalesced
            Xs[i * XL + c] = (c < nc) ? B[i + (c0 + c) * ldb] : T(0);
        }
    }
    for (int r0 = 0; r0 < m; r0 += 32) {
        const int rb = min(32, m - r0), rows = m - r0;
        __syncthreads();  // Xs staged, or the previous block's updates done with Ps
        {
            // thread t stages row t % 128 of the block's column panel, every other column
            const int i = t & 127, kh = t >> 7;
            if (i < rows) {
                T v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int kk = kh + 2 * u;
                    v[u] = (kk < rb) ? A[(r0 + i) + (i64)(r0 + kk) * lda] : T(0);
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int kk = kh + 2 * u;
                    if (kk < rb) Ps[i * PL + kk] = v[u];
                }
            }
        }
        __syncthreads();
        if (rb == 32) llnu_block<T, true>(r0, rb, rows, lane, wave, Xs, Ps);
        else llnu_block<T, false>(r0, rb, rows, lane, wave, Xs, Ps);
    }
    __syncthreads();
    if (m == 128) {
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int idx = t + 256 * u, i = idx & 127, c = idx >> 7;
            if (c < nc) B[i + (c0 + c) * ldb] = Xs[i * XL + c];
        }
    } else {
        for (int idx = t; idx < m * 64; idx += 256) {
            const int i = idx % m, c = idx / m;
            if (c < nc) B[i + (c0 + c) * ldb] = Xs[i * XL + c];
        }
    }
}

constexpr int TP = 16;           // lanes per vector
constexpr int TNT = 256;         // threads per workgroup
constexpr int TV = TNT / TP;     // vectors per workgroup

// maxbits (may be null): the largest |entry| written is folded into *maxbits (the bits of a
// non-negative float, atomic max), for the exponent of a low-precision product that follows;
// one atomic per workgroup (called by every thread of the TNT-thread workgroup).
__device__ __forceinline__ void fold_absmax(float mx, unsigned* maxbits) {
    __shared__ float wmx[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) wmx[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(maxbits, __float_as_uint(fmaxf(fmaxf(wmx[0], wmx[1]), fmaxf(wmx[2], wmx[3]))));
}

template <typename T>
__global__ __launch_bounds__(TNT) void trsm_small_kernel(int m, i64 nvec, const T* __restrict__ A,
                                                         i64 lda, T* B, i64 ldb, int right,
                                                         int lower_eff, int mtrans, int unit,
                                                         const int32_t* __restrict__ stop,
                                                         unsigned* maxbits) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if (stop && *stop != 0) return;
    // Both LDS arrays have a leading dimension of m + 1, so transposed stagings (consecutive
    // threads walking a column of one and a row of the other) hit distinct banks.
    const int xm = m + 1;
    T* x = reinterpret_cast<T*>(smem_raw);  // x[v * xm + i]: a vector's rows are adjacent, so
                                            // the 16 lanes of a vector hit distinct banks
    T* Ms = x + (size_t)xm * TV;            // the triangle, Ms[i + j*xm] = op(A)(i, j)
    const int t = threadIdx.x;
    const int v = t / TP, q = t % TP;
    const i64 v0 = (i64)blockIdx.x * TV;
    // The triangle is staged in LDS once, so each of the m dependent steps below waits on
    // LDS rather than on the cache hierarchy. At order 128 the loads are issued 32 at a time
    // before their LDS stores (a load-store loop would wait out each load's latency in turn).
    if (m == 128) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            T v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + TNT * (32 * h + u), i = idx & 127, j = idx >> 7;
                v[u] = A[i + (i64)j * lda];
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + TNT * (32 * h + u), i = idx & 127, j = idx >> 7;
                Ms[mtrans ? (j + i * xm) : (i + j * xm)] = v[u];
            }
        }
    } else {
        for (int idx = t; idx < m * m; idx += TNT) {
            const int i = idx % m, j = idx / m;  // walk A's columns: coalesced reads
            Ms[mtrans ? (j + i * xm) : (i + j * xm)] = A[i + (i64)j * lda];
        }
    }
    // Coalesced staging: left side walks down each column, right side across rows.
#pragma unroll 8
    for (int idx = t; idx < m * TV; idx += TNT) {
        int i, vv;
        if (!right) {
            i = idx % m;
            vv = idx / m;
        } else {
            vv = idx % TV;
            i = idx / TV;
        }
        const i64 gv = v0 + vv;
        T val = T(0);
        if (gv < nvec) val = right ? B[gv + (i64)i * ldb] : B[i + gv * ldb];
        x[vv * xm + i] = val;
    }
    __syncthreads();
    auto M = [&](int i, int j) -> T { return Ms[i + j * xm]; };
    // Column-oriented substitutions keep each vector in the registers of its 16 lanes (lane q
    // holds entries q, q+16, ...); x_j reaches the other lanes by one shuffle per step. Every
    // entry still receives its updates one FMA at a time in step order.
    constexpr int KR = 128 / TP;  // register slots per lane (order <= 128)
    if (lower_eff || !mtrans) {
        T xr[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int i = q + TP * k;
            xr[k] = (i < m) ? x[v * xm + i] : T(0);
        }
        if (lower_eff) {
#pragma unroll
            for (int kb = 0; kb < KR; ++kb) {
                for (int jj = 0; jj < TP; ++jj) {
                    const int j = kb * TP + jj;
                    if (j >= m) break;
                    T xj = __shfl(xr[kb], jj, TP);
                    if (!unit) {
                        xj = right ? xj * (T(1) / M(j, j)) : xj / M(j, j);
                        if (q == jj) xr[kb] = xj;
                    }
#pragma unroll
                    for (int k = kb; k < KR; ++k) {
                        const int i = q + TP * k;
                        if (i > j && i < m) xr[k] = __builtin_fma(-M(i, j), xj, xr[k]);
                    }
                }
            }
        } else {
#pragma unroll
            for (int kb = KR - 1; kb >= 0; --kb) {
                for (int jj = TP - 1; jj >= 0; --jj) {
                    const int j = kb * TP + jj;
                    if (j >= m) continue;
                    T xj = __shfl(xr[kb], jj, TP);
                    if (!unit) {
                        xj = right ? xj * (T(1) / M(j, j)) : xj / M(j, j);
                        if (q == jj) xr[kb] = xj;
                    }
#pragma unroll
                    for (int k = 0; k <= kb; ++k) {
                        const int i = q + TP * k;
                        if (i < j) xr[k] = __builtin_fma(-M(i, j), xj, xr[k]);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int i = q + TP * k;
            if (i < m) x[v * xm + i] = xr[k];
        }
    } else if (q == 0) {
        for (int i = m - 1; i >= 0; --i) {
            T temp = x[v * xm + i];
            for (int k = i + 1; k < m; ++k) temp = __builtin_fma(-M(i, k), x[v * xm + k], temp);
            if (!unit) temp = right ? temp * (T(1) / M(i, i)) : temp / M(i, i);
            x[v * xm + i] = temp;
        }
    }
    __syncthreads();
    float mx = 0.0f;
    for (int idx = t; idx < m * TV; idx += TNT) {
        int i, vv;
        if (!right) {
            i = idx % m;
            vv = idx / m;
        } else {
            vv = idx % TV;
            i = idx / TV;
        }
        const i64 gv = v0 + vv;
        if (gv < nvec) {
            if (right) B[gv + (i64)i * ldb] = x[vv * xm + i];
            else B[i + gv * ldb] = x[vv * xm + i];
            if (maxbits) mx = fmaxf(mx, (float)fabs(x[vv * xm + i]));
        }
    }
    if (maxbits) fold_absmax(mx, maxbits);
}

// ---------------------------------------------------------------------------------------
// v of lane `src` (uniform) for every lane
__device__ __forceinline__ float bcast_lane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double bcast_lane(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Unblocked Cholesky of an n x n (n <= 128) diagonal block in one workgroup, in LDS.
// Lower: cholesky::LowerVariant3Unblocked (src/lapack_like/factor/Cholesky/LowerVariant3.hpp:
// 15-41): alpha = sqrt(a_jj) (a_jj <= 0 -> NonHPD), a21 *= 1/alpha, dsyr('L', -1, a21).
// Upper: UpperVariant3Unblocked (UpperVariant3.hpp:15-46) on the row instead of the column.
// The dsyr update is a(i,l) = fma(x_i, -x_l, a(i,l)), skipped when x_l == 0, as in the
// reference BLAS.
//
// Schedule: every entry receives exactly those FMAs in column order, so the result is the
// unblocked algorithm's bit for bit, but the work is blocked by 32 columns: (1) all threads
// bring a 32-column block column up to date with the columns to its left (one sequential FMA
// chain per entry), (2) wave 0 factors the block column in registers (lane q holds rows q and
// q+64 of it; pivots and multipliers reach the other lanes by readlane), so only a few
// workgroup barriers are needed per 32 columns instead of two per column. The upper case is
// factored as its transpose; fma(x, -y, a) and fma(y, -x, a) are the same number, so only the
// skip rule differs (the row's own entry for the upper case). On a non-positive pivot the
// block columns to the right receive the updates of the finished columns, which leaves the
// block as the unblocked algorithm does when it stops.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void potf2_kernel(int n, T* A, i64 lda, int upper, i64 goff,
                                                    int32_t* info) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* S = reinterpret_cast<T*>(smem_raw);  // lower triangle L(r, c) at S[r + c*LD]
    __shared__ int s_fail;
    // an earlier block already failed: the reference has thrown, nothing more changes
    if (*info != 0) return;
    const int LD = n + 1, t = threadIdx.x, wave = t >> 6, lane = t & 63;
    for (int idx = t; idx < n * n; idx += 256) {
        const int i = idx % n, j = idx / n;
        // (r, c) = (i, j) of the lower triangle, read from (j, i) of the upper one
        if (i >= j) S[i + j * LD] = upper ? A[j + (i64)i * lda] : A[i + (i64)j * lda];
    }
    if (t == 0) s_fail = -1;
    __syncthreads();
    // (1) entries (r, c) of block column J, c in [c0, c1), r >= c: FMAs of columns [0, jmax)
    auto left_update = [&](int c0, int c1, int jmax) {
        const int h = n - c0, w = c1 - c0;
        for (int idx = t; idx < h * w; idx += 256) {
            const int r = c0 + idx % h, c = c0 + idx / h;
            if (r < c) continue;
            T acc = S[r + c * LD];
#pragma unroll 4
            for (int j = 0; j < jmax; ++j) {
                const T x = S[r + j * LD], y = S[c + j * LD];
                if ((upper ? x : y) != T(0)) acc = __builtin_fma(x, -y, acc);
            }
            S[r + c * LD] = acc;
        }
    };
    int fail = -1;
    for (int c0 = 0; c0 < n; c0 += 32) {
        const int c1 = min(n, c0 + 32);
        if (c0 > 0) {
            left_update(c0, c1, c0);
            __syncthreads();
        }
        // (2) the block column in wave 0's registers: rows c0 + q and c0 + 64 + q of lane q
        if (wave == 0) {
            T rw[2][32];
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int r = c0 + 64 * sl + lane;
#pragma unroll
                for (int cc = 0; cc < 32; ++cc)
                    rw[sl][cc] = (r < n && c0 + cc < c1 && r >= c0 + cc) ? S[r + (c0 + cc) * LD] : T(0);
            }
            int f = -1;
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                const int cj = c0 + jj;
                if (cj < c1 && f < 0) {
                    T alpha = bcast_lane(rw[0][jj], jj);
                    if (alpha <= T(0)) {
                        f = cj;
                    } else {
                        alpha = sqrt(alpha);
                        const T inv = T(1) / alpha;
#pragma unroll
                        for (int sl = 0; sl < 2; ++sl) {
                            const int r = c0 + 64 * sl + lane;
                            if (r == cj) rw[sl][jj] = alpha;
                            else if (r > cj) rw[sl][jj] = inv * rw[sl][jj];
                        }
#pragma unroll
                        for (int cc = jj + 1; cc < 32; ++cc) {
                            if (c0 + cc < c1) {
                                const T y = bcast_lane(rw[0][jj], cc);  // L(c0+cc, cj)
#pragma unroll
                                for (int sl = 0; sl < 2; ++sl) {
                                    const int r = c0 + 64 * sl + lane;
                                    const T x = rw[sl][jj];
                                    if (r >= c0 + cc && (upper ? x : y) != T(0))
                                        rw[sl][cc] = __builtin_fma(x, -y, rw[sl][cc]);
                                }
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int r = c0 + 64 * sl + lane;
#pragma unroll
                for (int cc = 0; cc < 32; ++cc)
                    if (r < n && c0 + cc < c1 && r >= c0 + cc) S[r + (c0 + cc) * LD] = rw[sl][cc];
            }
            if (lane == 0) s_fail = f;
        }
        __syncthreads();
        fail = s_fail;
        if (fail >= 0) {
            // the columns to the right, as the unblocked algorithm leaves them at the stop
            for (int d0 = c1; d0 < n; d0 += 32) left_update(d0, min(n, d0 + 32), fail);
            __syncthreads();
            break;
        }
    }
    if (fail >= 0 && t == 0) atomicCAS(info, 0, (int32_t)(goff + fail + 1));
    for (int idx = t; idx < n * n; idx += 256) {
        const int i = idx % n, j = idx / n;
        if (i >= j) {
            if (upper) A[j + (i64)i * lda] = S[i + j * LD];
            else A[i + (i64)j * lda] = S[i + j * LD];
        }
    }
}


// ---------------------------------------------------------------------------------------
// LU without pivoting of an n x n (n <= 128) diagonal block in one workgroup, in LDS:
// lu::Unb (src/lapack_like/factor/LU/Local.hpp:44-61). Per column j: alpha = a(j,j) (exactly
// zero -> SingularMatrixException in the reference: info = goff+j+1 and the block is left as
// it stands), a21 *= 1/alpha (blas::Scal), then Geru(-1): a(i,k) = fma(a(i,j), -a(j,k), a(i,k))
// for every k with a(j,k) != 0, the reference BLAS dger's skip rule.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void getf2_nopiv_kernel(int n, T* A, i64 lda, i64 goff,
                                                          int32_t* info) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* S = reinterpret_cast<T*>(smem_raw);  // S[i + j*LD]
    if (*info != 0) return;
    const int LD = n + 1, t = threadIdx.x;
    for (int idx = t; idx < n * n; idx += 256) {
        const int i = idx % n, j = idx / n;
        S[i + j * LD] = A[i + (i64)j * lda];
    }
    __syncthreads();
    int fail = -1;
    for (int j = 0; j < n; ++j) {
        const T alpha = S[j + j * LD];
        if (alpha == T(0)) {
            fail = j;
            break;
        }
        const T inv = T(1) / alpha;
        for (int i = j + 1 + t; i < n; i += 256) S[i + j * LD] = inv * S[i + j * LD];
        __syncthreads();
        const int h = n - (j + 1);
        for (int idx = t; idx < h * h; idx += 256) {
            const int i = j + 1 + idx % h, k = j + 1 + idx / h;
            const T y = S[j + k * LD];
            if (y != T(0)) S[i + k * LD] = __builtin_fma(S[i + j * LD], -y, S[i + k * LD]);
        }
        __syncthreads();
    }
    if (fail >= 0 && t == 0) atomicCAS(info, 0, (int32_t)(goff + fail + 1));
    for (int idx = t; idx < n * n; idx += 256) {
        const int i = idx % n, j = idx / n;
        A[i + (i64)j * lda] = S[i + j * LD];
    }
}

}  // namespace

namespace elx {

template <typename T>
static elx_status_t row_moves_impl(hipStream_t s, const int2* pairs, int npairs, T* A, i64 lda,
                                   i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3,
                                   const int32_t* stop) {
    ELX_ARG(npairs <= 256 && rows >= 0);
    const i64 total = std::max<i64>(c1 - c0, 0) + std::max<i64>(c3 - c2, 0);
    if (total <= 0 || npairs <= 0) return ELX_SUCCESS;
    if (c1 < c0) c1 = c0;
    if (c3 < c2) c3 = c2;
    ProfScope prof_(s, kKRowMoves, 2.0 * npairs * (double)total * sizeof(T));
    hipLaunchKernelGGL(row_moves_kernel<T>, dim3(cdiv(total, 4)), dim3(256), 0, s, pairs, npairs,
                       A, lda, r0, rows, c0, c1, c2, c3, stop);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_row_moves_f64(hipStream_t s, const int2* pairs, int npairs, double* A,
                                  i64 lda, i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3) {
    return row_moves_impl<double>(s, pairs, npairs, A, lda, r0, rows, c0, c1, c2, c3, nullptr);
}

elx_status_t launch_row_moves(hipStream_t s, const int2* pairs, int npairs, double* A, i64 lda,
                              i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3,
                              const int32_t* stop) {
    return row_moves_impl<double>(s, pairs, npairs, A, lda, r0, rows, c0, c1, c2, c3, stop);
}

elx_status_t launch_row_moves(hipStream_t s, const int2* pairs, int npairs, float* A, i64 lda,
                              i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3,
                              const int32_t* stop) {
    return row_moves_impl<float>(s, pairs, npairs, A, lda, r0, rows, c0, c1, c2, c3, stop);
}

elx_status_t launch_row_moves_f32_rowmajor(hipStream_t s, const int2* pairs, int npairs,
                                           float* A, i64 lda, i64 r0, i64 rows, i64 c0, i64 c1,
                                           i64 c2, i64 c3) {
    ELX_ARG(npairs <= 256 && rows >= 0);
    if (c1 < c0) c1 = c0;
    if (c3 < c2) c3 = c2;
    const i64 total = (c1 - c0) + (c3 - c2);
    if (total <= 0 || npairs <= 0) return ELX_SUCCESS;
    const i64 slices = cdiv(c1 - c0, 64) + cdiv(c3 - c2, 64);
    ProfScope prof_(s, kKRowMoves, 2.0 * npairs * (double)total * sizeof(float));
    hipLaunchKernelGGL(row_moves_rowmajor_kernel, dim3((unsigned)slices), dim3(512), 0, s, pairs,
                       npairs, A, lda, r0, rows, c0, c1, c2, c3);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_swaps_to_perm(hipStream_t s, const int32_t* dests, i64 nswaps, i64 size,
                                  int32_t* perm, int inverse) {
    if (size <= 0) return ELX_SUCCESS;
    const DeviceInfo& di = device_info();
    const size_t need = sizeof(int32_t) * (size_t)size;
    const size_t dyn_max = (size_t)di.max_lds_per_block - sizeof(int32_t) * kSwapChunk;
    const int in_lds = need <= dyn_max ? 1 : 0;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&swaps_to_perm_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_max);
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL(swaps_to_perm_kernel, dim3(1), dim3(256), in_lds ? need : 0, s, dests,
                       nswaps, size, perm, inverse, in_lds);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T>
static elx_status_t gather_rows_impl(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                     const T* Bin, i64 ldin, T* Bout, i64 ldout) {
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    ELX_ARG(n <= 65535);
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(cdiv(m, 256), (unsigned)n), dim3(256), 0, s,
                       perm, m, n, Bin, ldin, Bout, ldout);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_gather_rows_f64(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                    const double* Bin, i64 ldin, double* Bout, i64 ldout) {
    return gather_rows_impl<double>(s, perm, m, n, Bin, ldin, Bout, ldout);
}

elx_status_t launch_gather_rows(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                const double* Bin, i64 ldin, double* Bout, i64 ldout) {
    return gather_rows_impl<double>(s, perm, m, n, Bin, ldin, Bout, ldout);
}

elx_status_t launch_gather_rows(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                const float* Bin, i64 ldin, float* Bout, i64 ldout) {
    return gather_rows_impl<float>(s, perm, m, n, Bin, ldin, Bout, ldout);
}

// One right-hand side, left side, column-oriented substitution (M = op(A) lower, forward;
// or M = A upper, backward): lu::SolveAfter's diagonal blocks (src/lapack_like/solve/LU.cpp
// Trsm(LEFT,LOWER,NORMAL,UNIT) and Trsm(LEFT,UPPER,NORMAL,NON_UNIT)) at one column. The
// block is staged in LDS by the whole workgroup; then one wave solves, lane q holding rows q
// and q + 64: x_j reaches every lane by readlane (no cross-lane shuffle through LDS), and each
// lane reads its two entries of M's column j, which do not depend on x and are issued ahead.
// Every entry receives the same FMAs in the same order as trsm_small_kernel's
// column-oriented path (b_i = fma(-m_ij, x_j, b_i) in step order, x_j = b_j / m_jj for a
// non-unit diagonal), so the two agree bit for bit.
template <typename T>
__global__ __launch_bounds__(256) void trsv_col_kernel(int m, const T* __restrict__ A, i64 lda,
                                                       T* B, int lower_eff, int mtrans, int unit,
                                                       const int32_t* __restrict__ stop) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    if (stop && *stop != 0) return;
    T* Ms = reinterpret_cast<T*>(smem_raw);  // Ms[i + j*128] = op(A)(i, j)
    const int t = threadIdx.x;
    if (m == 128) {
        // loads issued 32 at a time before their LDS stores (the block is rarely in cache)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            T v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + 256 * (32 * h + u), i = idx & 127, j = idx >> 7;
                v[u] = A[i + (i64)j * lda];
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + 256 * (32 * h + u), i = idx & 127, j = idx >> 7;
                Ms[mtrans ? (j + i * 128) : (i + j * 128)] = v[u];
            }
        }
    } else {
        for (int idx = t; idx < m * m; idx += 256) {
            const int i = idx % m, j = idx / m;  // walk A's columns: coalesced reads
            Ms[mtrans ? (j + i * 128) : (i + j * 128)] = A[i + (i64)j * lda];
        }
    }
    __syncthreads();
    if (t >= 64) return;
    const int q = t;
    const bool h0 = q < m, h1 = q + 64 < m;
    T x0 = h0 ? B[q] : T(0), x1 = h1 ? B[q + 64] : T(0);
    if (lower_eff) {
        for (int j = 0; j < m; ++j) {
            const T* col = Ms + j * 128;
            const T m0 = col[q], m1 = col[q + 64];
            T xj = bcast_lane(j < 64 ? x0 : x1, j & 63);
            if (!unit) {
                xj = xj / col[j];
                if (q == (j & 63)) {
                    if (j < 64) x0 = xj;
                    else x1 = xj;
                }
            }
            if (h0 && q > j) x0 = __builtin_fma(-m0, xj, x0);
            if (h1 && q + 64 > j) x1 = __builtin_fma(-m1, xj, x1);
        }
    } else {
        for (int j = m - 1; j >= 0; --j) {
            const T* col = Ms + j * 128;
            const T m0 = col[q], m1 = col[q + 64];
            T xj = bcast_lane(j < 64 ? x0 : x1, j & 63);
            if (!unit) {
                xj = xj / col[j];
                if (q == (j & 63)) {
                    if (j < 64) x0 = xj;
                    else x1 = xj;
                }
            }
            if (q < j) x0 = __builtin_fma(-m0, xj, x0);
            if (h1 && q + 64 < j) x1 = __builtin_fma(-m1, xj, x1);
        }
    }
    if (h0) B[q] = x0;
    if (h1) B[q + 64] = x1;
}

template <typename T>
static elx_status_t trsm_small_impl(hipStream_t s, int side, int uplo, int orient, int diag,
                                    i64 m, i64 n, const T* A, i64 lda, T* B, i64 ldb,
                                    const int32_t* stop, unsigned* maxbits = nullptr) {
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    const bool right = (side == kRight);
    const i64 order = right ? n : m;
    const i64 nvec = right ? m : n;
    ELX_ARG(order <= 128);
    const bool trans = (orient != kNormal);
    const bool lower = (uplo == kLower);
    // Left: M = op(A). Right: M = op(A)^T.
    const bool mtrans = right ? !trans : trans;
    const bool lower_eff = lower != mtrans;
    if (!right && lower && !trans && diag == kUnit && nvec >= 256 && !maxbits) {
        // many right-hand sides: the blocked form (same FMA order per entry)
        const size_t lds2 = sizeof(T) * (128 * 65 + 128 * 33);
        ProfScope prof2_(s, kKTrsm, (double)order * (double)order * (double)nvec);
        static bool attr2 = false;
        if (!attr2) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&trsm_llnu_kernel<T>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
            (void)hipGetLastError();
            attr2 = true;
        }
        hipLaunchKernelGGL(trsm_llnu_kernel<T>, dim3(cdiv(nvec, 64)), dim3(256), lds2, s,
                           (int)order, nvec, A, lda, B, ldb, stop);
        ELX_HIP(hipGetLastError());
        return ELX_SUCCESS;
    }
    if (!right && nvec == 1 && (lower_eff || !mtrans) && !maxbits) {
        // one right-hand side (the solve's diagonal blocks): same FMA order per entry
        const size_t lds1 = sizeof(T) * 128 * 128;
        ProfScope prof1_(s, kKTrsm, (double)order * (double)order);
        static bool attr1 = false;
        if (!attr1) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&trsv_col_kernel<T>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
            (void)hipGetLastError();
            attr1 = true;
        }
        hipLaunchKernelGGL(trsv_col_kernel<T>, dim3(1), dim3(256), lds1, s, (int)order, A, lda,
                           B, lower_eff ? 1 : 0, mtrans ? 1 : 0, diag == kUnit ? 1 : 0, stop);
        ELX_HIP(hipGetLastError());
        return ELX_SUCCESS;
    }
    const size_t lds = sizeof(T) * ((size_t)(order + 1) * TV + (size_t)order * (order + 1));
    ProfScope prof_(s, kKTrsm, (double)order * (double)order * (double)nvec);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&trsm_small_kernel<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(T) * (129 * TV + 128 * 129)));
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL(trsm_small_kernel<T>, dim3(cdiv(nvec, TV)), dim3(TNT), lds, s, (int)order,
                       nvec, A, lda, B, ldb, right ? 1 : 0, lower_eff ? 1 : 0, mtrans ? 1 : 0,
                       diag == kUnit ? 1 : 0, stop, maxbits);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_trsm_small(hipStream_t s, int side, int uplo, int orient, int diag, i64 m,
                               i64 n, const double* A, i64 lda, double* B, i64 ldb,
                               const int32_t* stop) {
    return trsm_small_impl<double>(s, side, uplo, orient, diag, m, n, A, lda, B, ldb, stop);
}

elx_status_t launch_trsm_small(hipStream_t s, int side, int uplo, int orient, int diag, i64 m,
                               i64 n, const float* A, i64 lda, float* B, i64 ldb,
                               const int32_t* stop) {
    return trsm_small_impl<float>(s, side, uplo, orient, diag, m, n, A, lda, B, ldb, stop);
}

elx_status_t launch_trsm_small_f32(hipStream_t s, int side, int uplo, int orient, int diag,
                                   i64 m, i64 n, const float* A, i64 lda, float* B, i64 ldb,
                                   unsigned* maxbits) {
    return trsm_small_impl<float>(s, side, uplo, orient, diag, m, n, A, lda, B, ldb, nullptr,
                                  maxbits);
}

elx_status_t launch_potf2_f64(hipStream_t s, int uplo, i64 n, double* A, i64 lda, i64 goff,
                              int32_t* info) {
    if (n <= 0) return ELX_SUCCESS;
    ELX_ARG(n <= 128);
    const size_t lds = sizeof(double) * (size_t)n * (n + 1);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&potf2_kernel<double>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(double) * 128 * 129));
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL(potf2_kernel<double>, dim3(1), dim3(256), lds, s, (int)n, A, lda,
                       uplo == kUpper ? 1 : 0, goff, info);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_potf2(hipStream_t s, int uplo, i64 n, double* A, i64 lda, i64 goff,
                          int32_t* info) {
    return launch_potf2_f64(s, uplo, n, A, lda, goff, info);
}

elx_status_t launch_potf2(hipStream_t s, int uplo, i64 n, float* A, i64 lda, i64 goff,
                          int32_t* info) {
    return launch_potf2_f32(s, uplo, n, A, lda, goff, info);
}

template <typename T>
static elx_status_t getf2_nopiv_impl(hipStream_t s, i64 n, T* A, i64 lda, i64 goff,
                                     int32_t* info) {
    if (n <= 0) return ELX_SUCCESS;
    ELX_ARG(n <= 128);
    const size_t lds = sizeof(T) * (size_t)n * (n + 1);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&getf2_nopiv_kernel<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(T) * 128 * 129));
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL(getf2_nopiv_kernel<T>, dim3(1), dim3(256), lds, s, (int)n, A, lda, goff,
                       info);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_getf2_nopiv(hipStream_t s, i64 n, double* A, i64 lda, i64 goff,
                                int32_t* info) {
    return getf2_nopiv_impl<double>(s, n, A, lda, goff, info);
}

elx_status_t launch_getf2_nopiv(hipStream_t s, i64 n, float* A, i64 lda, i64 goff,
                                int32_t* info) {
    return getf2_nopiv_impl<float>(s, n, A, lda, goff, info);
}

elx_status_t launch_potf2_f32(hipStream_t s, int uplo, i64 n, float* A, i64 lda, i64 goff,
                              int32_t* info) {
    if (n <= 0) return ELX_SUCCESS;
    ELX_ARG(n <= 128);
    const size_t lds = sizeof(float) * (size_t)n * (n + 1);
    hipLaunchKernelGGL(potf2_kernel<float>, dim3(1), dim3(256), lds, s, (int)n, A, lda,
                       uplo == kUpper ? 1 : 0, goff, info);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

}  // namespace elx
