// Batched mixed-precision LU solves on the device: thousands of independent nonsymmetric
// systems A_i X_i = B_i (n_i <= 2048), each factored and refined by one persistent workgroup,
// with no host round trip between the launches.
//
// Per system (the semantics of elx_dgesv_mixed with ELX_REFINE_AUTO, which restates
// El::LU / lu::SolveAfter, src/lapack_like/factor/LU.cpp:104-142, LU/SolveAfter.hpp:66-92, as
// the preconditioner of the iterative refinement of include/El/lapack_like/solve/Refined.hpp
// and FGMRES.hpp):
//   1. W = fp32(A), stored row by row (row i contiguous), padded to ldw = round_up(n, 32) rows
//      and columns with zeros.
//   2. Blocked right-looking LU with partial pivoting of W (32-column panels held in
//      registers, the first largest |a| wins a column's pivot, the pivot row scaled by its
//      reciprocal), the trailing matrix updated by FP16 MFMA (v_mfma_f32_16x16x32_f16, FP32
//      accumulation) with the U operand scaled by a power of two below 2^15.
//   3. Refinement in FP64: x = M^-1 b, then r = b - A x with A read in FP64; plain IR while a
//      step cuts the backward error at least 3x (a candidate kept only when the residual
//      decreased, Refined.hpp:96-105), then right-preconditioned FGMRES(30) (modified
//      Gram-Schmidt applied twice, Givens rotations), until ||r||inf / (||A||inf ||x||inf +
//      ||b||inf) <= tol.
//   4. A system whose refinement stalls (or whose low-precision factorization meets an exactly
//      zero pivot) falls back to the FP64 LU of A (the same kernel on FP64 storage with FP64
//      MFMA updates, 16-column panels) and the same refinement with FP64 factors.
// The preconditioner applies the row permutation, then solves with the unit-lower and upper
// factors by 32-row blocks: the products with the solved part by the whole workgroup, each
// 32 x 32 diagonal block by one wave (a row per lane, the solved entries broadcast by
// shuffles).
//
// Scheduling: systems are sorted by order, largest first, and split into two size bins with
// their own workgroup size (256 threads for n <= 512, 1024 for n <= 2048; two panel rows per
// thread), each bin pulling systems from a work queue. Per-system state (stage, iteration
// counts, backward error) lives on the device; the host reads it once at the end.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "elx_common.hpp"

namespace {

constexpr int IB = 32;       // diagonal blocks of the preconditioner (inverted)
constexpr int NRC = 4;       // right-hand sides refined together
constexpr int RESTART = 30;  // FGMRES restart length
constexpr int MAX_CYCLES = 3;
constexpr int MAX_IR = 30;

struct LuItem {
    const double* A;
    const double* B;
    double* X;
    int32_t n, lda, ldb, ldx, nrhs, index, ldw, pad;
    int64_t woff;  // working factor (ldw x ldw, row by row) in the pool of its element type
    int64_t poff;  // pivots (ldw int32), then the explicit row permutation (ldw int32)
};

// stage: 0 low-precision tier, 1 converged, 2 FP64 tier, 3 failed
struct SysState {
    int32_t stage, outer, inner, fell_back, factor_info, low_failed;
    double eta;
    double anorm;
};

template <typename T>
struct LuCfg;
template <>
struct LuCfg<float> {
    static constexpr int PB = 32;
};
template <>
struct LuCfg<double> {
    static constexpr int PB = 16;
};

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

// f(integral_constant<int, I>) for I = B .. E-1, each call with a compile-time index.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// ------------------------------------------------------------------ block reductions
template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    constexpr int NW = NT / 64;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double m = red[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) m = fmax(m, red[w]);
    return m;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    constexpr int NW = NT / 64;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w];
    return s;
}

// ------------------------------------------------------------------ factorization
// Shared memory of the factor kernel (dynamic): the U chunk of the trailing update, reused for
// the row-permutation replay at the end.
template <typename T>
struct FactorLds;
template <>
struct FactorLds<float> {
    static constexpr int NC = 1024;        // columns per chunk
    static constexpr int LDU = 40;         // halves per staged column (32 + 8 padding)
    static constexpr size_t bytes = (size_t)NC * LDU * 2;
};
template <>
struct FactorLds<double> {
    static constexpr int NC = 512;
    static constexpr size_t bytes = (size_t)16 * NC * 8;
};

// One system: W (ldw x ldw, row by row) = LU of fp(A) with partial pivoting. Returns the
// 1-based column of an exactly zero (or NaN) pivot, 0 on success.
template <typename T, int NT, int NMAX>
__device__ __forceinline__ int factor_one(const LuItem& it, T* W, int32_t* piv, unsigned char* dyn) {
    constexpr int PB = LuCfg<T>::PB;
    constexpr int RPT = NMAX / NT;  // panel rows per thread
    constexpr int NW = NT / 64;
    __shared__ T xrow[2][PB];
    __shared__ T l11[PB][PB + 1];
    __shared__ double red[NW];
    __shared__ int redi[NW];
    __shared__ int pivl[PB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = it.n, ldw = it.ldw;
    // (1) W = fp(A), row by row, zero padding: 32 x 32 tiles through LDS (the U chunk area)
    {
        T* tile = reinterpret_cast<T*>(dyn);  // NW tiles of 32 x 33
        T* my = tile + wave * 32 * 33;
        const int tiles = ldw / 32;
        for (int tt = wave; tt < tiles * tiles; tt += NW) {
            const int ti = tt % tiles, tj = tt / tiles;
            const int i0 = ti * 32, j0 = tj * 32;
            // lane: row i0 + (lane & 31), columns j0 + (lane >> 5) + 2s
#pragma unroll 4
            for (int s = 0; s < 16; ++s) {
                const int i = i0 + (lane & 31), j = j0 + (lane >> 5) + 2 * s;
                my[(j - j0) * 33 + (i - i0)] = (i < n && j < n) ? (T)it.A[i + (size_t)j * it.lda] : T(0);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll 4
            for (int s = 0; s < 16; ++s) {
                const int j = j0 + (lane & 31), i = i0 + (lane >> 5) + 2 * s;
                W[(size_t)i * ldw + j] = my[(j - j0) * 33 + (i - i0)];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
    }
    __syncthreads();
    int sing = 0;
    for (int k0 = 0; k0 < n; k0 += PB) {
        const int w = min(PB, n - k0), m = n - k0;
        // (2) the panel: rows k0.., columns k0..k0+PB-1, in registers (row r = t + q NT)
        T pr[RPT][PB];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = t + q * NT;
#pragma unroll
            for (int j = 0; j < PB; ++j) pr[q][j] = (r < m) ? W[(size_t)(k0 + r) * ldw + k0 + j] : T(0);
        }
        // every column a compile-time index into the register panel (no dynamic indexing)
        static_for<0, PB>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if (c >= w || sing) return;
            // the first row of largest |a(r, c)|, r >= c
            double bv = -1.0;
            int br = 0x7fffffff;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = t + q * NT;
                const double v = fabs((double)pr[q][c]);
                if (r >= c && r < m && (v > bv || (v == bv && r < br))) {
                    bv = v;
                    br = r;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(bv, off);
                const int orr = __shfl_xor(br, off);
                if (ov > bv || (ov == bv && orr < br)) {
                    bv = ov;
                    br = orr;
                }
            }
            if (lane == 0) {
                red[wave] = bv;
                redi[wave] = br;
            }
            __syncthreads();
            double gv = red[0];
            int p = redi[0];
#pragma unroll
            for (int ww = 1; ww < NW; ++ww)
                if (red[ww] > gv || (red[ww] == gv && redi[ww] < p)) {
                    gv = red[ww];
                    p = redi[ww];
                }
            if (!(gv > 0.0)) {  // exactly zero, or every candidate NaN
                sing = k0 + c + 1;
                return;
            }
            if (t == 0) pivl[c] = p;
            // exchange rows c and p of the panel through LDS
            const int tc = c % NT, qc = c / NT, tp = p % NT, qp = p / NT;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (t == tp && q == qp)
#pragma unroll
                    for (int j = 0; j < PB; ++j) xrow[0][j] = pr[q][j];
                if (t == tc && q == qc)
#pragma unroll
                    for (int j = 0; j < PB; ++j) xrow[1][j] = pr[q][j];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (t == tc && q == qc)
#pragma unroll
                    for (int j = 0; j < PB; ++j) pr[q][j] = xrow[0][j];
                if (t == tp && q == qp && p != c)
#pragma unroll
                    for (int j = 0; j < PB; ++j) pr[q][j] = xrow[1][j];
            }
            const T inv = T(1) / xrow[0][c];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int r = t + q * NT;
                if (r > c && r < m) {
                    const T l = pr[q][c] * inv;
                    pr[q][c] = l;
#pragma unroll
                    for (int j = c + 1; j < PB; ++j) pr[q][j] = __builtin_fma(-l, xrow[0][j], pr[q][j]);
                }
            }
            __syncthreads();
        });
        if (sing) break;
        // (3) the panel back to W; L11 to LDS; the pivots
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = t + q * NT;
            if (r < m)
#pragma unroll
                for (int j = 0; j < PB; ++j) W[(size_t)(k0 + r) * ldw + k0 + j] = pr[q][j];
            if (r < PB)
#pragma unroll
                for (int j = 0; j < PB; ++j) l11[r][j] = (r < m) ? pr[q][j] : T(0);
        }
        if (t < w) piv[k0 + t] = k0 + pivl[t];
        __syncthreads();
        // (4) the panel's interchanges on the other columns, in order
        for (int j = t; j < ldw; j += NT) {
            if (j >= k0 && j < k0 + PB) continue;
            for (int c = 0; c < w; ++c) {
                const int p = pivl[c];
                if (p != c) {
                    T* a = W + (size_t)(k0 + c) * ldw + j;
                    T* b = W + (size_t)(k0 + p) * ldw + j;
                    const T x = *a;
                    *a = *b;
                    *b = x;
                }
            }
        }
        __syncthreads();
        const int j1 = k0 + PB;  // first trailing column (padding beyond n is zero)
        if (j1 >= ldw) break;
        // (5) U12 = L11^-1 A12, one column per thread; its largest magnitude
        double umax = 0.0;
        for (int j = j1 + t; j < ldw; j += NT) {
            // keep the reads of L11 inside the loop (hoisted, they would not fit in registers)
            asm volatile("" ::: "memory");
            T u[PB];
#pragma unroll
            for (int i = 0; i < PB; ++i) u[i] = W[(size_t)(k0 + i) * ldw + j];
#pragma unroll
            for (int kk = 0; kk < PB; ++kk)
#pragma unroll
                for (int i = kk + 1; i < PB; ++i) u[i] = __builtin_fma(-u[kk], l11[i][kk], u[i]);
#pragma unroll
            for (int i = 0; i < PB; ++i) {
                W[(size_t)(k0 + i) * ldw + j] = u[i];
                umax = fmax(umax, fabs((double)u[i]));
            }
        }
        // (6) A22 -= L21 U12 in column chunks staged in LDS
        const int i1 = k0 + PB;  // first trailing row
        const int nstrip = (ldw - i1) / 16;
        if constexpr (sizeof(T) == 4) {
            umax = block_max<NT>(umax, red);
            int e = 0;
            if (umax > 0.0 && isfinite(umax)) {
                int p2;
                (void)frexp(umax, &p2);  // umax < 2^p2
                e = max(-100, min(100, 15 - p2));
            }
            const float scale = __builtin_ldexpf(1.0f, e), alpha = -__builtin_ldexpf(1.0f, -e);
            typedef FactorLds<float> FL;
            _Float16* Ub = reinterpret_cast<_Float16*>(dyn);
            for (int jc = j1; jc < ldw; jc += FL::NC) {
                const int nc = min(FL::NC, ldw - jc);
                for (int idx = t; idx < nc * PB; idx += NT) {
                    const int jj = idx % nc, k = idx / nc;
                    Ub[jj * FL::LDU + k] = (_Float16)(scale * (float)W[(size_t)(k0 + k) * ldw + jc + jj]);
                }
                __syncthreads();
                const int lr = lane & 15, kq = lane >> 4;
                for (int s = wave; s < nstrip; s += NW) {
                    const int i0 = i1 + 16 * s;
                    // L21 rows i0..i0+15, k = 8 kq .. 8 kq + 7, rounded to fp16
                    const float* lrow = reinterpret_cast<const float*>(W) + (size_t)(i0 + lr) * ldw + k0 + 8 * kq;
                    const f4 la = *reinterpret_cast<const f4*>(lrow);
                    const f4 lb = *reinterpret_cast<const f4*>(lrow + 4);
                    h8 lf;
                    lf[0] = (_Float16)la[0];
                    lf[1] = (_Float16)la[1];
                    lf[2] = (_Float16)la[2];
                    lf[3] = (_Float16)la[3];
                    lf[4] = (_Float16)lb[0];
                    lf[5] = (_Float16)lb[1];
                    lf[6] = (_Float16)lb[2];
                    lf[7] = (_Float16)lb[3];
                    for (int tj = 0; tj < nc; tj += 16) {
                        const h8 uf = *reinterpret_cast<const h8*>(Ub + (tj + lr) * FL::LDU + 8 * kq);
                        const f4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(uf, lf, f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                        // lane holds C(i0 + lr, jc + tj + 4 kq + r), r = 0..3
                        float* cp = reinterpret_cast<float*>(W) + (size_t)(i0 + lr) * ldw + jc + tj + 4 * kq;
                        f4 cv = *reinterpret_cast<f4*>(cp);
#pragma unroll
                        for (int r = 0; r < 4; ++r) cv[r] = __builtin_fmaf(alpha, d[r], cv[r]);
                        *reinterpret_cast<f4*>(cp) = cv;
                    }
                }
                __syncthreads();
            }
        } else {
            (void)umax;
            typedef FactorLds<double> FL;
            double* Ud = reinterpret_cast<double*>(dyn);  // [PB][NC]
            __syncthreads();
            for (int jc = j1; jc < ldw; jc += FL::NC) {
                const int nc = min(FL::NC, ldw - jc);
                for (int idx = t; idx < nc * PB; idx += NT) {
                    const int jj = idx % nc, k = idx / nc;
                    Ud[k * FL::NC + jj] = (double)W[(size_t)(k0 + k) * ldw + jc + jj];
                }
                __syncthreads();
                const int lr = lane & 15, kq = lane >> 4;
                for (int s = wave; s < nstrip; s += NW) {
                    const int i0 = i1 + 16 * s;
                    double a[4];
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks)
                        a[ks] = (double)W[(size_t)(i0 + lr) * ldw + k0 + 4 * ks + kq];
                    for (int tj = 0; tj < nc; tj += 16) {
                        d4 acc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int ks = 0; ks < 4; ++ks)
                            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], Ud[(4 * ks + kq) * FL::NC + tj + lr], acc, 0, 0, 0);
                        // lane holds C(i0 + kq + 4 r, jc + tj + lr)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            T* cp = W + (size_t)(i0 + kq + 4 * r) * ldw + jc + tj + lr;
                            *cp = *cp - (T)acc[r];
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    return sing;
}

// The explicit row permutation (the rows of P A in order), replayed from the pivots.
template <int NT>
__device__ void finish_one(const LuItem& it, const int32_t* piv, int32_t* perm, unsigned char* dyn) {
    const int t = threadIdx.x;
    const int n = it.n, ldw = it.ldw;
    int32_t* pl = reinterpret_cast<int32_t*>(dyn);
    for (int i = t; i < ldw; i += NT) pl[i] = i;
    __syncthreads();
    if (t == 0)
        for (int c = 0; c < n; ++c) {
            const int p = piv[c];
            if (p != c) {
                const int x = pl[c];
                pl[c] = pl[p];
                pl[p] = x;
            }
        }
    __syncthreads();
    for (int i = t; i < ldw; i += NT) perm[i] = pl[i];
    __syncthreads();
}

template <typename T, int NT, int NMAX>
__global__ __launch_bounds__(NT) void lu_factor_kernel(const LuItem* __restrict__ items, int begin,
                                                       int end, int* queue, SysState* st, T* pool,
                                                       int32_t* ipool, int tier) {
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = begin + atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= end) break;
        const LuItem it = items[i];
        SysState* S = st + it.index;
        const int stage = S->stage;
        __syncthreads();
        if (stage != (tier == 0 ? 0 : 2)) continue;
        T* W = pool + it.woff;
        int32_t* piv = ipool + it.poff;
        const int sing = factor_one<T, NT, NMAX>(it, W, piv, dyn);
        if (sing) {
            if (threadIdx.x == 0) {
                if (tier == 0) {
                    S->low_failed = sing;
                    S->factor_info = sing;
                    S->stage = 2;
                    S->fell_back = 1;
                } else {
                    S->factor_info = sing;
                    S->stage = 3;
                }
            }
            __syncthreads();
            continue;
        }
        finish_one<NT>(it, piv, piv + it.ldw, dyn);
    }
}

// ------------------------------------------------------------------ refinement
// Work space of one refining workgroup (FP64, NMAX entries per vector): XP, R, D (NRC vectors
// each), V (RESTART + 1), Z (RESTART).
// Lane src's value in every lane (src uniform), through the scalar unit: the diagonal-block
// solves' broadcast, one readlane (two for FP64) instead of a cross-lane LDS permute.
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

template <int NMAX>
struct RefWs {
    static constexpr size_t per_slot = (size_t)(3 * NRC + 2 * RESTART + 1) * NMAX;
};

template <typename T, int NT, int NMAX>
struct Refiner {
    static constexpr int NW = NT / 64;
    const T* W;
    const int32_t* perm;
    T* yl;  // LDS: NRC x NMAX
    T* tv;  // LDS: NRC x IB
    int n, ldw;

    // yl(k0:k0+IB, q) = T_bb^-1 tv(:, q) by wave 0, lane r holding row r of the diagonal block
    // (unit lower L_bb forward, or upper U_bb backward; rows >= n of the padding are identity).
    template <bool UPPER>
    __device__ void diag_solve(int k0) {
        const int lane = threadIdx.x & 63;
        if ((threadIdx.x >> 6) != 0) return;
        const int r = lane & 31;
        T row[IB], yv[NRC];
        const T* wr = W + (size_t)(k0 + r) * ldw + k0;
#pragma unroll
        for (int j = 0; j < IB; ++j) row[j] = wr[j];
        const bool live = (k0 + r < n);
#pragma unroll
        for (int q = 0; q < NRC; ++q) yv[q] = tv[q * IB + r];
        if constexpr (!UPPER) {
            static_for<0, IB>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
#pragma unroll
                for (int q = 0; q < NRC; ++q) {
                    const T yk = lane_bcast(yv[q], k);
                    if (r > k) yv[q] = __builtin_fma(-row[k], yk, yv[q]);
                }
            });
        } else {
            static_for<0, IB>([&](auto kk) {
                constexpr int k = IB - 1 - decltype(kk)::value;
#pragma unroll
                for (int q = 0; q < NRC; ++q) {
                    if (r == k) yv[q] = live ? yv[q] / row[k] : yv[q];
                    const T yk = lane_bcast(yv[q], k);
                    if (r < k) yv[q] = __builtin_fma(-row[k], yk, yv[q]);
                }
            });
        }
        if (lane < 32)
#pragma unroll
            for (int q = 0; q < NRC; ++q) yl[q * NMAX + k0 + r] = yv[q];
    }

    // dst(:, q) = M^-1 src(:, q) for q < nq; src/dst columns at stride ssd / dsd (FP64).
    __device__ __forceinline__ void precond(const double* src, int ssd, double* dst, int dsd, int nq) {
        const int t = threadIdx.x;
        for (int idx = t; idx < NRC * ldw; idx += NT) {
            const int q = idx / ldw, i = idx % ldw;
            yl[q * NMAX + i] = (q < nq && i < n) ? (T)src[(size_t)q * ssd + perm[i]] : T(0);
        }
        __syncthreads();
        const int g = t & 31, r0 = t >> 5;  // 32 lanes along a row, NT/32 rows at a time
        constexpr int RSTEP = NT / 32;
        const int nblk = ldw / IB;
        // forward: L y = P b, by 32-row blocks
        for (int b = 0; b < nblk; ++b) {
            const int k0 = b * IB;
            for (int r = r0; r < IB; r += RSTEP) {
                double acc[NRC] = {0.0, 0.0, 0.0, 0.0};
                const T* row = W + (size_t)(k0 + r) * ldw;
                for (int j = g; j < k0; j += 32) {
                    const double l = (double)row[j];
#pragma unroll
                    for (int q = 0; q < NRC; ++q) acc[q] = __builtin_fma(l, (double)yl[q * NMAX + j], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q) {
#pragma unroll
                    for (int off = 16; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
                    if (g == 0) tv[q * IB + r] = (T)((double)yl[q * NMAX + k0 + r] - acc[q]);
                }
            }
            __syncthreads();
            diag_solve<false>(k0);
            __syncthreads();
        }
        // backward: U z = y
        for (int b = nblk - 1; b >= 0; --b) {
            const int k0 = b * IB;
            for (int r = r0; r < IB; r += RSTEP) {
                double acc[NRC] = {0.0, 0.0, 0.0, 0.0};
                const T* row = W + (size_t)(k0 + r) * ldw;
                for (int j = k0 + IB + g; j < ldw; j += 32) {
                    const double u = (double)row[j];
#pragma unroll
                    for (int q = 0; q < NRC; ++q) acc[q] = __builtin_fma(u, (double)yl[q * NMAX + j], acc[q]);
                }
#pragma unroll
                for (int q = 0; q < NRC; ++q) {
#pragma unroll
                    for (int off = 16; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
                    if (g == 0) tv[q * IB + r] = (T)((double)yl[q * NMAX + k0 + r] - acc[q]);
                }
            }
            __syncthreads();
            diag_solve<true>(k0);
            __syncthreads();
        }
        for (int idx = t; idx < nq * n; idx += NT) {
            const int q = idx / n, i = idx % n;
            dst[(size_t)q * dsd + i] = (double)yl[q * NMAX + i];
        }
        __syncthreads();
    }
};

// R(:, q) = B(:, q) - A X(:, q) for q < nq (FP64); rowsum (optional) receives sum_j |a_ij|.
template <int NT>
__device__ void residual(int n, const double* A, int lda, const double* B, int ldb,
                         const double* X, int ldx, double* R, int ldr, int nq, double* rowmax_out,
                         double* red) {
    const int t = threadIdx.x;
    double rmax = 0.0;
    for (int i = t; i < n; i += NT) {
        double acc[NRC], rs = 0.0;
#pragma unroll
        for (int q = 0; q < NRC; ++q) acc[q] = (q < nq) ? B[i + (size_t)q * ldb] : 0.0;
        for (int j = 0; j < n; ++j) {
            const double a = A[i + (size_t)j * lda];
            rs += fabs(a);
#pragma unroll
            for (int q = 0; q < NRC; ++q)
                if (q < nq) acc[q] = __builtin_fma(-a, X[j + (size_t)q * ldx], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < NRC; ++q)
            if (q < nq) R[i + (size_t)q * ldr] = acc[q];
        rmax = fmax(rmax, rs);
    }
    if (rowmax_out) {
        const double m = block_max<NT>(rmax, red);
        if (t == 0) *rowmax_out = m;
    }
    __syncthreads();
}

// max |v(:, q)| for q < nq into out[q] (every thread gets the values).
template <int NT>
__device__ void col_absmax(int n, const double* V, int ldv, int nq, double* out, double* red) {
    for (int q = 0; q < nq; ++q) {
        double m = 0.0;
        for (int i = threadIdx.x; i < n; i += NT) m = fmax(m, fabs(V[i + (size_t)q * ldv]));
        out[q] = block_max<NT>(m, red);
    }
}

template <typename T, int NT, int NMAX>
__global__ __launch_bounds__(NT) void lu_refine_kernel(const LuItem* __restrict__ items, int begin,
                                                       int end, int* queue, SysState* st,
                                                       const T* pool, const int32_t* ipool,
                                                       double* wsp, int tier, double tol_in) {
    constexpr int NW = NT / 64;
    __shared__ T yl[NRC * NMAX];
    __shared__ T tv[NRC * IB];
    __shared__ double red[NW];
    __shared__ double H[(RESTART + 1) * RESTART];
    __shared__ double gcs[RESTART], gsn[RESTART], gt[RESTART + 1], hcol[RESTART + 1];
    __shared__ int s_i;
    double* ws = wsp + (size_t)blockIdx.x * RefWs<NMAX>::per_slot;
    double* XP = ws;
    double* R = XP + (size_t)NRC * NMAX;
    double* D = R + (size_t)NRC * NMAX;
    double* V = D + (size_t)NRC * NMAX;
    double* Z = V + (size_t)(RESTART + 1) * NMAX;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) s_i = begin + atomicAdd(queue, 1);
        __syncthreads();
        const int ii = s_i;
        __syncthreads();
        if (ii >= end) break;
        const LuItem it = items[ii];
        SysState* S = st + it.index;
        const int stage = S->stage;
        __syncthreads();
        if (stage != (tier == 0 ? 0 : 2)) continue;
        const int n = it.n;
        const double tol = tol_in > 0.0 ? tol_in : (double)n * 0x1p-53;
        Refiner<T, NT, NMAX> M{pool + it.woff, ipool + it.poff + it.ldw, yl, tv, n, it.ldw};
        double anorm = -1.0;
        double worst = 0.0;
        bool all_ok = true;
        int outer = 0, inner = 0;
        for (int c0 = 0; c0 < it.nrhs; c0 += NRC) {
            const int nq = min(NRC, it.nrhs - c0);
            const double* B = it.B + (size_t)c0 * it.ldb;
            double* X = it.X + (size_t)c0 * it.ldx;
            M.precond(B, it.ldb, X, it.ldx, nq);
            double eta[NRC], eta_prev[NRC], rn_prev[NRC], bn[NRC], rn[NRC], xn[NRC];
            bool done[NRC], gm[NRC];
            for (int q = 0; q < NRC; ++q) {
                done[q] = q >= nq;
                gm[q] = false;
                eta_prev[q] = INFINITY;
                rn_prev[q] = INFINITY;
                eta[q] = INFINITY;
            }
            col_absmax<NT>(n, B, it.ldb, nq, bn, red);
            // plain IR on the block of right-hand sides
            for (int step = 0;; ++step) {
                residual<NT>(n, it.A, it.lda, B, it.ldb, X, it.ldx, R, NMAX, nq,
                             anorm < 0.0 ? &red[0] : nullptr, red);
                if (anorm < 0.0) {
                    anorm = red[0];
                    __syncthreads();
                }
                col_absmax<NT>(n, R, NMAX, nq, rn, red);
                col_absmax<NT>(n, X, it.ldx, nq, xn, red);
                bool any = false;
                for (int q = 0; q < nq; ++q) {
                    if (done[q] || gm[q]) continue;
                    const double den = anorm * xn[q] + bn[q];
                    eta[q] = den > 0.0 ? rn[q] / den : (rn[q] > 0.0 ? INFINITY : 0.0);
                    if (eta[q] <= tol) {
                        done[q] = true;
                        continue;
                    }
                    if (step > 0 && !(rn[q] < rn_prev[q])) {
                        // Refined.hpp:96-105: back to the previous iterate, on to FGMRES
                        for (int i = t; i < n; i += NT) X[i + (size_t)q * it.ldx] = XP[i + (size_t)q * NMAX];
                        eta[q] = eta_prev[q];
                        gm[q] = true;
                        continue;
                    }
                    if (step >= MAX_IR || (step > 0 && !(eta[q] * 3.0 <= eta_prev[q]))) {
                        gm[q] = true;
                        continue;
                    }
                    eta_prev[q] = eta[q];
                    rn_prev[q] = rn[q];
                    any = true;
                }
                __syncthreads();
                if (!any) break;
                ++outer;
                M.precond(R, NMAX, D, NMAX, nq);
                for (int q = 0; q < nq; ++q) {
                    if (done[q] || gm[q]) continue;
                    for (int i = t; i < n; i += NT) {
                        const double x = X[i + (size_t)q * it.ldx];
                        XP[i + (size_t)q * NMAX] = x;
                        X[i + (size_t)q * it.ldx] = x + D[i + (size_t)q * NMAX];
                    }
                }
                __syncthreads();
            }
            // FGMRES-IR for the columns plain IR could not finish
            for (int q = 0; q < nq; ++q) {
                if (done[q]) {
                    worst = fmax(worst, eta[q]);
                    continue;
                }
                const double* b = B + (size_t)q * it.ldb;
                double* x = X + (size_t)q * it.ldx;
                double* r = R;
                bool conv = false;
                for (int cyc = 0; cyc < MAX_CYCLES && !conv; ++cyc) {
                    residual<NT>(n, it.A, it.lda, b, it.ldb, x, it.ldx, r, NMAX, 1, nullptr, red);
                    double rq[1], xq[1];
                    col_absmax<NT>(n, r, NMAX, 1, rq, red);
                    col_absmax<NT>(n, x, it.ldx, 1, xq, red);
                    const double den = anorm * xq[0] + bn[q];
                    eta[q] = den > 0.0 ? rq[0] / den : (rq[0] > 0.0 ? INFINITY : 0.0);
                    if (eta[q] <= tol) {
                        conv = true;
                        break;
                    }
                    // ||r||_2 and v_0
                    double s = 0.0;
                    for (int i = t; i < n; i += NT) s += r[i] * r[i];
                    const double beta = sqrt(block_sum<NT>(s, red));
                    if (!(beta > 0.0) || !isfinite(beta)) break;
                    for (int i = t; i < n; i += NT) V[i] = r[i] / beta;
                    if (t == 0) gt[0] = beta;
                    __syncthreads();
                    // the new residual has to reach half the tolerance's residual norm (2-norm
                    // bound on the inf-norm target)
                    const double relTol = 0.5 * tol * den / beta;
                    int jend = 0;
                    for (int j = 0; j < RESTART; ++j) {
                        double* vj = V + (size_t)j * NMAX;
                        double* zj = Z + (size_t)j * NMAX;
                        M.precond(vj, NMAX, zj, NMAX, 1);
                        double* w = V + (size_t)(j + 1) * NMAX;
                        // w = A z_j (FP64)
                        for (int i = t; i < n; i += NT) {
                            double acc = 0.0;
                            for (int kk = 0; kk < n; ++kk)
                                acc = __builtin_fma(it.A[i + (size_t)kk * it.lda], zj[kk], acc);
                            w[i] = acc;
                        }
                        __syncthreads();
                        for (int i2 = 0; i2 <= j; ++i2) hcol[i2] = 0.0;
                        __syncthreads();
                        for (int pass = 0; pass < 2; ++pass) {
                            for (int i2 = 0; i2 <= j; ++i2) {
                                const double* vi = V + (size_t)i2 * NMAX;
                                double d = 0.0;
                                for (int i = t; i < n; i += NT) d = __builtin_fma(vi[i], w[i], d);
                                const double h = block_sum<NT>(d, red);
                                for (int i = t; i < n; i += NT) w[i] = __builtin_fma(-h, vi[i], w[i]);
                                if (t == 0) hcol[i2] += h;
                                __syncthreads();
                            }
                        }
                        double s2 = 0.0;
                        for (int i = t; i < n; i += NT) s2 += w[i] * w[i];
                        const double delta = sqrt(block_sum<NT>(s2, red));
                        if (delta > 0.0)
                            for (int i = t; i < n; i += NT) w[i] = w[i] / delta;
                        if (t == 0) {
                            double* Hj = H + (size_t)j * (RESTART + 1);
                            for (int i2 = 0; i2 <= j; ++i2) Hj[i2] = hcol[i2];
                            Hj[j + 1] = delta;
                            for (int i2 = 0; i2 < j; ++i2) {
                                const double e0 = Hj[i2], e1 = Hj[i2 + 1];
                                Hj[i2] = gcs[i2] * e0 + gsn[i2] * e1;
                                Hj[i2 + 1] = -gsn[i2] * e0 + gcs[i2] * e1;
                            }
                            const double rho = hypot(Hj[j], Hj[j + 1]);
                            gcs[j] = rho > 0.0 ? Hj[j] / rho : 1.0;
                            gsn[j] = rho > 0.0 ? Hj[j + 1] / rho : 0.0;
                            Hj[j] = rho;
                            Hj[j + 1] = 0.0;
                            const double t0 = gt[j];
                            gt[j] = gcs[j] * t0;
                            gt[j + 1] = -gsn[j] * t0;
                        }
                        __syncthreads();
                        jend = j + 1;
                        ++inner;
                        if (!(H[(size_t)j * (RESTART + 1) + j] > 0.0)) break;
                        if (fabs(gt[j + 1]) <= relTol * beta || delta == 0.0) break;
                        if (jend >= 20 && fabs(gt[j + 1]) > 0.5 * beta) break;  // not worth finishing
                    }
                    // y = H^-1 t (upper triangular), x += Z y
                    if (t == 0) {
                        for (int i2 = jend - 1; i2 >= 0; --i2) {
                            double v = gt[i2];
                            for (int k = i2 + 1; k < jend; ++k) v -= H[(size_t)k * (RESTART + 1) + i2] * hcol[k];
                            const double hd = H[(size_t)i2 * (RESTART + 1) + i2];
                            hcol[i2] = hd != 0.0 ? v / hd : 0.0;
                        }
                    }
                    __syncthreads();
                    for (int i = t; i < n; i += NT) {
                        double acc = x[i];
                        for (int k = 0; k < jend; ++k) acc = __builtin_fma(hcol[k], Z[(size_t)k * NMAX + i], acc);
                        x[i] = acc;
                    }
                    __syncthreads();
                }
                if (!conv) {
                    residual<NT>(n, it.A, it.lda, b, it.ldb, x, it.ldx, r, NMAX, 1, nullptr, red);
                    double rq[1], xq[1];
                    col_absmax<NT>(n, r, NMAX, 1, rq, red);
                    col_absmax<NT>(n, x, it.ldx, 1, xq, red);
                    const double den = anorm * xq[0] + bn[q];
                    eta[q] = den > 0.0 ? rq[0] / den : (rq[0] > 0.0 ? INFINITY : 0.0);
                    conv = eta[q] <= tol;
                }
                worst = fmax(worst, eta[q]);
                all_ok = all_ok && conv;
                if (!all_ok && tier == 0) break;  // this system goes to the FP64 tier
            }
            if (!all_ok && tier == 0) break;
        }
        if (t == 0) {
            S->outer += outer;
            S->inner += inner;
            S->anorm = anorm;
            S->eta = worst;
            if (all_ok) {
                S->stage = 1;
            } else if (tier == 0) {
                S->stage = 2;
                S->fell_back = 1;
            } else {
                S->stage = 3;
            }
        }
        __syncthreads();
    }
}

__global__ void lu_state_init(SysState* st, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        SysState s;
        s.stage = 0;
        s.outer = s.inner = s.fell_back = s.factor_info = s.low_failed = 0;
        s.eta = 0.0;
        s.anorm = 0.0;
        st[i] = s;
    }
}

__global__ void zero_words(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ host side
constexpr int kBins = 2;
constexpr int kBinNmax[kBins] = {512, 2048};
constexpr int kBinThreads[kBins] = {256, 1024};

template <typename T, int NT, int NMAX>
size_t factor_lds() {
    constexpr int NW = NT / 64;
    size_t chunk;
    if constexpr (sizeof(T) == 4)
        chunk = (size_t)std::min(FactorLds<float>::NC, NMAX) * FactorLds<float>::LDU * 2;
    else
        chunk = (size_t)16 * std::min(FactorLds<double>::NC, NMAX) * 8;
    return std::max({chunk, (size_t)NW * 32 * 33 * sizeof(T), (size_t)NMAX * 4});
}

template <typename T, int NT, int NMAX>
elx_status_t launch_factor_bin(hipStream_t s, const LuItem* items, int lo, int hi, int* queue,
                               SysState* st, T* pool, int32_t* ipool, int tier) {
    const void* fn = reinterpret_cast<const void*>(&lu_factor_kernel<T, NT, NMAX>);
    const size_t lds = factor_lds<T, NT, NMAX>();
    static bool attr = false;
    if (!attr) {
        ELX_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds));
    ELX_ARG(per_cu >= 1);
    const int G = (int)std::min<int64_t>(hi - lo, (int64_t)per_cu * elx::device_info().num_cus);
    hipLaunchKernelGGL((lu_factor_kernel<T, NT, NMAX>), dim3(G), dim3(NT), lds, s, items, lo, hi,
                       queue, st, pool, ipool, tier);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T, int NT, int NMAX>
int refine_grid(int count) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&lu_refine_kernel<T, NT, NMAX>), NT, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return (int)std::min<int64_t>(count, (int64_t)per_cu * elx::device_info().num_cus);
}

template <typename T, int NT, int NMAX>
elx_status_t launch_refine_bin(hipStream_t s, const LuItem* items, int lo, int hi, int G,
                               int* queue, SysState* st, const T* pool, const int32_t* ipool,
                               double* ws, int tier, double tol) {
    hipLaunchKernelGGL((lu_refine_kernel<T, NT, NMAX>), dim3(G), dim3(NT), 0, s, items, lo, hi,
                       queue, st, pool, ipool, ws, tier, tol);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t launch_tier(hipStream_t s, const LuItem* items, const int* bin_lo, const int* bin_hi,
                         const int* Gr, int* queue, SysState* st, T* pool, int32_t* ipool,
                         double* const* ws, int tier, double tol) {
    int* qf = queue + (tier == 0 ? 0 : 4);
    int* qr = qf + 2;
    if (bin_hi[1] > bin_lo[1])
        ELX_CHECK((launch_factor_bin<T, 1024, 2048>(s, items, bin_lo[1], bin_hi[1], qf + 1, st,
                                                     pool, ipool, tier)));
    if (bin_hi[0] > bin_lo[0])
        ELX_CHECK((launch_factor_bin<T, 256, 512>(s, items, bin_lo[0], bin_hi[0], qf, st, pool,
                                                   ipool, tier)));
    if (bin_hi[1] > bin_lo[1])
        ELX_CHECK((launch_refine_bin<T, 1024, 2048>(s, items, bin_lo[1], bin_hi[1], Gr[1], qr + 1,
                                                     st, pool, ipool, ws[1], tier, tol)));
    if (bin_hi[0] > bin_lo[0])
        ELX_CHECK((launch_refine_bin<T, 256, 512>(s, items, bin_lo[0], bin_hi[0], Gr[0], qr, st,
                                                   pool, ipool, ws[0], tier, tol)));
    return ELX_SUCCESS;
}

// Device buffers of one batched call, grown on demand and kept per device.
struct LuBatchBuffers {
    void* items = nullptr;
    size_t items_bytes = 0;
    void* state = nullptr;
    size_t state_bytes = 0;
    void* f32 = nullptr;
    size_t f32_bytes = 0;
    void* f64 = nullptr;
    size_t f64_bytes = 0;
    void* ints = nullptr;
    size_t ints_bytes = 0;
    void* ws[kBins] = {nullptr, nullptr};
    size_t ws_bytes[kBins] = {0, 0};
    int* queue = nullptr;
};
std::mutex g_lumu;
std::map<int, LuBatchBuffers> g_lubuf;

elx_status_t grow(void** p, size_t* have, size_t need) {
    if (*have >= need) return ELX_SUCCESS;
    if (*p) {
        ELX_HIP(hipDeviceSynchronize());
        ELX_HIP(hipFree(*p));
        *p = nullptr;
        *have = 0;
    }
    if (hipMalloc(p, need) != hipSuccess) {
        *p = nullptr;
        return ELX_ERROR_ALLOC;
    }
    *have = need;
    return ELX_SUCCESS;
}

}  // namespace

namespace elx {

void batched_lu_release() {
    std::lock_guard<std::mutex> lk(g_lumu);
    for (auto& kv : g_lubuf) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(kv.first);
        LuBatchBuffers& b = kv.second;
        for (void* p : {b.items, b.state, b.f32, b.f64, b.ints, b.ws[0], b.ws[1]})
            if (p) (void)hipFree(p);
        if (b.queue) (void)hipFree(b.queue);
        (void)hipSetDevice(cur);
    }
    g_lubuf.clear();
}

// The batched mixed LU-IR of elx_dgesv_mixed_batched (FP16 factors, refinement as
// ELX_REFINE_AUTO, FP64 LU fallback) for systems of order <= 2048. Synchronous.
elx_status_t gesv_mixed_batched_device(hipStream_t s, int64_t count, const int32_t* n,
                                       const int32_t* nrhs, const double* const* A_ptrs,
                                       const int32_t* lda, const double* const* B_ptrs,
                                       const int32_t* ldb, double* const* X_ptrs,
                                       const int32_t* ldx, double tol, elx_ir_stats_t* stats) {
    ELX_ARG(count > 0 && count < (1 << 30));
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        ELX_ARG(n[i] >= 0 && n[i] <= kBinNmax[kBins - 1] && nrhs[i] >= 0);
        ELX_ARG(lda[i] >= std::max(n[i], 1) && ldb[i] >= std::max(n[i], 1) &&
                ldx[i] >= std::max(n[i], 1));
        order[(size_t)i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return n[a] > n[b]; });
    std::vector<LuItem> items;
    items.reserve((size_t)count);
    int64_t woff = 0, poff = 0;
    int bin_lo[kBins] = {0, 0}, bin_hi[kBins] = {0, 0};
    for (int64_t k : order) {
        if (n[k] == 0 || nrhs[k] == 0) continue;
        LuItem it;
        it.A = A_ptrs[k];
        it.B = B_ptrs[k];
        it.X = X_ptrs[k];
        it.n = n[k];
        it.lda = lda[k];
        it.ldb = ldb[k];
        it.ldx = ldx[k];
        it.nrhs = nrhs[k];
        it.index = (int32_t)k;
        it.ldw = (int32_t)((n[k] + 31) / 32 * 32);
        it.pad = 0;
        it.woff = woff;
        woff += (int64_t)it.ldw * it.ldw;  // a multiple of 64 elements (ldw is of 32)
        it.poff = poff;
        poff += 2 * (int64_t)it.ldw;
        items.push_back(it);
    }
    const int cnt = (int)items.size();
    // bins in the descending order: [0, c1) orders > 512, [c1, cnt) orders <= 512
    int c1 = 0;
    while (c1 < cnt && items[(size_t)c1].n > kBinNmax[0]) ++c1;
    bin_lo[1] = 0;
    bin_hi[1] = c1;
    bin_lo[0] = c1;
    bin_hi[0] = cnt;
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_lumu);
    LuBatchBuffers& bf = g_lubuf[dev];
    const int Gr[kBins] = {std::max(1, refine_grid<float, 256, 512>(bin_hi[0] - bin_lo[0])),
                           std::max(1, refine_grid<float, 1024, 2048>(bin_hi[1] - bin_lo[1]))};
    const int Gr64[kBins] = {std::max(1, refine_grid<double, 256, 512>(bin_hi[0] - bin_lo[0])),
                             std::max(1, refine_grid<double, 1024, 2048>(bin_hi[1] - bin_lo[1]))};
    ELX_CHECK(grow(&bf.items, &bf.items_bytes, sizeof(LuItem) * std::max(cnt, 1)));
    ELX_CHECK(grow(&bf.state, &bf.state_bytes, sizeof(SysState) * (size_t)count));
    ELX_CHECK(grow(&bf.f32, &bf.f32_bytes, sizeof(float) * (size_t)std::max<int64_t>(woff, 64)));
    ELX_CHECK(grow(&bf.f64, &bf.f64_bytes, sizeof(double) * (size_t)std::max<int64_t>(woff, 64)));
    ELX_CHECK(grow(&bf.ints, &bf.ints_bytes, sizeof(int32_t) * (size_t)std::max<int64_t>(poff, 64)));
    ELX_CHECK(grow(&bf.ws[0], &bf.ws_bytes[0],
                   sizeof(double) * RefWs<512>::per_slot * (size_t)std::max(Gr[0], Gr64[0])));
    ELX_CHECK(grow(&bf.ws[1], &bf.ws_bytes[1],
                   sizeof(double) * RefWs<2048>::per_slot * (size_t)std::max(Gr[1], Gr64[1])));
    if (!bf.queue) ELX_HIP(hipMalloc(&bf.queue, sizeof(int) * 16));
    LuItem* d_items = static_cast<LuItem*>(bf.items);
    SysState* d_st = static_cast<SysState*>(bf.state);
    if (cnt > 0)
        ELX_HIP(hipMemcpyAsync(d_items, items.data(), sizeof(LuItem) * (size_t)cnt,
                               hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(lu_state_init, dim3(cdiv(count, 256)), dim3(256), 0, s, d_st, (int)count);
    hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, s, bf.queue, 16);
    ELX_HIP(hipGetLastError());
    hipEvent_t ev[3];
    for (hipEvent_t& e : ev) ELX_HIP(hipEventCreate(&e));
    ELX_HIP(hipEventRecord(ev[0], s));
    double* ws[kBins] = {static_cast<double*>(bf.ws[0]), static_cast<double*>(bf.ws[1])};
    int32_t* ip = static_cast<int32_t*>(bf.ints);
    elx_status_t r = ELX_SUCCESS;
    if (cnt > 0) {
        r = launch_tier<float>(s, d_items, bin_lo, bin_hi, Gr, bf.queue, d_st,
                               static_cast<float*>(bf.f32), ip, ws, 0, tol);
        if (r == ELX_SUCCESS) r = (hipEventRecord(ev[1], s) == hipSuccess) ? ELX_SUCCESS : ELX_ERROR_HIP;
        if (r == ELX_SUCCESS)
            r = launch_tier<double>(s, d_items, bin_lo, bin_hi, Gr64, bf.queue, d_st,
                                    static_cast<double*>(bf.f64), ip, ws, 1, tol);
    } else {
        (void)hipEventRecord(ev[1], s);
    }
    (void)hipEventRecord(ev[2], s);
    std::vector<SysState> hs((size_t)count);
    if (r == ELX_SUCCESS &&
        hipMemcpyAsync(hs.data(), d_st, sizeof(SysState) * (size_t)count, hipMemcpyDeviceToHost,
                       s) != hipSuccess)
        r = ELX_ERROR_HIP;
    if (r == ELX_SUCCESS && hipStreamSynchronize(s) != hipSuccess) r = ELX_ERROR_HIP;
    float ms_low = 0.f, ms_all = 0.f;
    if (r == ELX_SUCCESS) {
        (void)hipEventElapsedTime(&ms_low, ev[0], ev[1]);
        (void)hipEventElapsedTime(&ms_all, ev[0], ev[2]);
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (r != ELX_SUCCESS) return r;
    bool all_ok = true;
    for (int64_t i = 0; i < count; ++i) {
        const SysState& S = hs[(size_t)i];
        const bool trivial = (n[i] == 0 || nrhs[i] == 0);
        const bool ok = trivial || S.stage == 1;
        all_ok = all_ok && ok;
        if (stats) {
            elx_ir_stats_t& o = stats[i];
            std::memset(&o, 0, sizeof(o));
            o.converged = ok ? 1 : 0;
            o.outer_iters = S.outer;
            o.inner_iters = S.inner;
            o.fell_back = S.fell_back;
            o.factor_info = trivial ? 0 : S.low_failed;
            o.backward_error = trivial ? 0.0 : S.eta;
            // device time of the whole batch: the low-precision tier, then the FP64 tier
            o.factor_seconds = ms_low * 1e-3;
            o.refine_seconds = (ms_all - ms_low) * 1e-3;
        }
    }
    return all_ok ? ELX_SUCCESS : ELX_ERROR_NUMERIC;
}

}  // namespace elx
