// Batched variable-size Cholesky factorization and solve (lower, FP64 and FP32).
//
// Each matrix runs the blocked lower Cholesky of El::Cholesky (variant 3,
// src/lapack_like/factor/Cholesky/LowerVariant3.hpp:15-68) with 32-column blocks, in the
// left-looking order: a block column first receives the updates of every earlier block column
// (one GEMM with K = the number of earlier columns, on MFMA), then its 32 x 32 diagonal block
// is factored unblocked (alpha = sqrt(a_jj), a_jj <= 0 -> not HPD, a21 *= 1/alpha, then the
// dsyr rank-1 update, as LowerVariant3Unblocked does)
// and the rows below are solved against it (X L11^T = A21, here as the MFMA product of A21
// with the transposed inverse of L11, which differs from the reference's substitution only
// in rounding). The solve follows
// cholesky::SolveAfter (src/lapack_like/factor/Cholesky/SolveAfter.hpp:15-43): L y = b, then
// L^T x = y, blocked by 32 rows.
//
// Design for MI355X: one persistent workgroup per CU slot pulls matrices from a work queue in
// decreasing size (largest first, so the tail of the batch is made of small matrices); the
// block column being factored is accumulated in MFMA registers and staged in LDS (at most
// 512 x 33 values), the earlier columns stream from memory once per block column. Matrices
// are column-major, n <= 512, addressed through a device array of pointers.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "elx_common.hpp"

namespace {

constexpr int BNB = 32;     // block columns
constexpr int BNT = 256;    // threads per workgroup
constexpr int BNMAX = 512;  // largest order
constexpr int BLD = BNB + 1;

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

// S[kk * SM + r] = A(k0 + r, c0 + kk) for r < m, kk < kc, zero for kc <= kk < kcp: threads
// walk the (row, column) pairs column by column (coalesced), 16 loads in flight per thread.
template <typename T>
__device__ __forceinline__ void stage_cols(const T* A, int lda, int k0, int m, int c0, int kc,
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

// One workgroup: the lower Cholesky factor of one matrix, in place.
template <typename T>
__device__ void potrf_one(int n, T* A, int lda, int32_t* info, T* P, T* Ls, T* Minv) {
    typedef typename Mma<T>::v4 v4;
    constexpr int MT = BNMAX / 16 / 4;  // row tiles per wave
    constexpr int PCAP = BNMAX * BLD;   // LDS values of P, reused as the staging area
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = lane >> 4;
    T* S = P;
    __shared__ int s_fail;
    for (int k0 = 0; k0 < n; k0 += BNB) {
        const int kb = min(BNB, n - k0), m = n - k0;
        const int nrt = (m + 15) >> 4;
        const int SM = nrt * 16;
        // chunk width: at most 32 staged values per thread (held in registers while the
        // previous chunk is multiplied), a multiple of 4 columns
        const int kcw = min(256, max(4, ((32 * BNT) / SM) & ~3));
        const int cq = BNT / m, cr = BNT - cq * m;  // BNT = cq*m + cr: the walk over (row, column)
        // (1) C = A(k0:n, k0:k0+kb) - L(k0:n, 0:k0) L(k0:k0+kb, 0:k0)^T: the block column and
        // then chunks of L to its left pass through LDS; wave w accumulates row tiles w, w+4, ...
        // (all of them at once, in MFMA registers). Rows >= m of a staged tile and columns
        // >= kb of the block only reach accumulator entries that are never used.
        stage_cols(A, lda, k0, m, k0, kb, BNB, SM, S);
        __syncthreads();
        v4 acc[MT][2];
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (wave + 4 * u) * 16 + Mma<T>::drow(lane, r);
                    const int col = ct * 16 + lr;
                    acc[u][ct][r] = (row < m) ? S[col * SM + row] : T(0);
                }
        // chunk pipeline: the loads of chunk i+1 are in flight while chunk i is multiplied
        T pv[32];
        auto load_chunk = [&](int kc, int w) {
            const int total = m * w;
            int kk = t / m, r = t - kk * m;
#pragma unroll
            for (int u = 0; u < 32; ++u) {
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
            for (int u = 0; u < 32; ++u) {
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
                const T b0 = Sk[lr], b1 = Sk[16 + lr];
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    const int rt = wave + 4 * u;
                    if (rt < nrt) {
                        const T a = -Sk[rt * 16 + lr];
                        acc[u][0] = Mma<T>::mma(a, b0, acc[u][0]);
                        acc[u][1] = Mma<T>::mma(a, b1, acc[u][1]);
                    }
                }
            }
            kc = kcn;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (wave + 4 * u) * 16 + Mma<T>::drow(lane, r);
                    const int col = ct * 16 + lr;
                    if (row < m) P[row * BLD + col] = acc[u][ct][r];
                }
        __syncthreads();
        // (2) the diagonal block, unblocked, by wave 0: lane i holds row i. Each step
        // broadcasts the pivot and the scaled column through the scalar unit; the dsyr update
        // is a select, not a branch.
        if (wave == 0) {
            T rw[BNB];
#pragma unroll
            for (int c = 0; c < BNB; ++c) rw[c] = (lane < kb) ? P[lane * BLD + c] : T(0);
            int fail = -1;
#pragma unroll
            for (int j = 0; j < BNB; ++j) {
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
                        for (int l = j + 1; l < BNB; ++l) rw[l] = __builtin_fma(xi, -bcast(rw[j], l), rw[l]);
                    }
                }
            }
            if (lane < kb)
#pragma unroll
                for (int c = 0; c < BNB; ++c) {
                    P[lane * BLD + c] = rw[c];
                    Ls[lane * BLD + c] = rw[c];
                }
            if (lane == 0) s_fail = fail;
        }
        __syncthreads();
        const int fail = s_fail;
        if (fail < 0 && m > BNB) {
            // inv(L11) (rows below exist, so kb == 32) by all threads: 8 lanes per column c,
            // lane g of the group holds rows g, g+8, g+16, g+24 of y = L11^{-1} e_c, formed in
            // the column order (y_k /= l_kk, then y_i -= l_ik y_k); stored as row c of Minv, so
            // Minv(k, j) = inv(L11)(j, k)
            const int c = t >> 3, g = t & 7;
            T y[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) y[q] = (g + 8 * q == c) ? T(1) : T(0);
#pragma unroll
            for (int k = 0; k < BNB; ++k) {
                const int qk = k >> 3, gk = k & 7;
                if (g == gk) y[qk] = y[qk] / Ls[k * BLD + k];
                const T yk = __shfl(y[qk], (lane & ~7) | gk);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = g + 8 * q;
                    if (i > k) y[q] = __builtin_fma(-Ls[i * BLD + k], yk, y[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) Minv[c * BLD + g + 8 * q] = y[q];
            __syncthreads();
        }
        if (fail >= 0) {
            // NonHPDMatrixException at global column k0+fail. The block column as the blocked
            // reference holds it when it throws: the diagonal block's lower triangle after the
            // finished steps, the rows below updated by the earlier block columns
            if (t == 0) *info = k0 + fail + 1;
            for (int idx = t; idx < m * kb; idx += BNT) {
                const int i = idx % m, c = idx / m;
                if (i >= c) A[(k0 + i) + (size_t)(k0 + c) * lda] = P[i * BLD + c];
            }
            __syncthreads();
            return;
        }
        // (3) rows below (present only when kb == 32): X = A21 L11^{-T} as one MFMA product
        // with the inverse, Minv(k, j) = inv(L11)(j, k); results go straight to A
        for (int rt = 2 + wave; rt < nrt; rt += 4) {
            v4 acc2[2];
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc2[ct] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kk = 0; kk < BNB; kk += 4) {
                const T a = P[(rt * 16 + lr) * BLD + kk + lk];
                const T b0 = Minv[(kk + lk) * BLD + lr];
                const T b1 = Minv[(kk + lk) * BLD + 16 + lr];
                acc2[0] = Mma<T>::mma(a, b0, acc2[0]);
                acc2[1] = Mma<T>::mma(a, b1, acc2[1]);
            }
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rt * 16 + Mma<T>::drow(lane, r);
                    if (row < m) A[(k0 + row) + (size_t)(k0 + ct * 16 + lr) * lda] = acc2[ct][r];
                }
        }
        // the diagonal block's lower triangle
        for (int idx = t; idx < kb * kb; idx += BNT) {
            const int i = idx % kb, c = idx / kb;
            if (i >= c) A[(k0 + i) + (size_t)(k0 + c) * lda] = P[i * BLD + c];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(BNT) void potrf_vbatched_kernel(const BatchItem* __restrict__ items,
                                                             int count, T* const* __restrict__ As,
                                                             int32_t* __restrict__ info,
                                                             int* queue) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);  // BNMAX x BLD
    T* Ls = P + BNMAX * BLD;                // BNB x BLD: L11
    T* Minv = Ls + BNB * BLD;               // BNB x BLD: inv(L11)^T
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= count) break;
        const BatchItem it = items[i];
        if (threadIdx.x == 0) info[it.index] = 0;
        if (it.n > 0) potrf_one<T>(it.n, As[it.index], it.lda, info + it.index, P, Ls, Minv);
    }
}

// One workgroup: L L^T X = B for one matrix (L from potrf), nrhs columns at a time in LDS.
template <typename T>
__device__ void potrs_one(int n, const T* L, int lda, T* B, int ldb, int nrhs, T* Y, T* Ls) {
    const int t = threadIdx.x;
    constexpr int NR = 8;
    for (int c0 = 0; c0 < nrhs; c0 += NR) {
        const int nr = min(NR, nrhs - c0);
        for (int idx = t; idx < n * nr; idx += BNT) {
            const int i = idx % n, c = idx / n;
            Y[c * BNMAX + i] = B[i + (size_t)(c0 + c) * ldb];
        }
        __syncthreads();
        // forward: L y = b (dtrsm Left Lower NoTrans NonUnit: y_k /= l_kk, y_i -= y_k l_ik)
        for (int k0 = 0; k0 < n; k0 += BNB) {
            const int kb = min(BNB, n - k0);
            for (int idx = t; idx < kb * kb; idx += BNT) {
                const int i = idx % kb, c = idx / kb;
                Ls[i * BLD + c] = L[(k0 + i) + (size_t)(k0 + c) * lda];
            }
            __syncthreads();
            if (t < nr) {
                T* y = Y + t * BNMAX + k0;
                for (int k = 0; k < kb; ++k) {
                    if (y[k] != T(0)) {
                        y[k] = y[k] / Ls[k * BLD + k];
                        for (int i = k + 1; i < kb; ++i) y[i] = __builtin_fma(-y[k], Ls[i * BLD + k], y[i]);
                    }
                }
            }
            __syncthreads();
            for (int i = k0 + kb + t; i < n; i += BNT) {
                T acc[NR];
#pragma unroll
                for (int c = 0; c < NR; ++c) acc[c] = (c < nr) ? Y[c * BNMAX + i] : T(0);
                for (int k = 0; k < kb; ++k) {
                    const T lik = L[i + (size_t)(k0 + k) * lda];
#pragma unroll
                    for (int c = 0; c < NR; ++c)
                        if (c < nr) acc[c] = __builtin_fma(-Y[c * BNMAX + k0 + k], lik, acc[c]);
                }
#pragma unroll
                for (int c = 0; c < NR; ++c)
                    if (c < nr) Y[c * BNMAX + i] = acc[c];
            }
            __syncthreads();
        }
        // backward: L^T x = y (dtrsm Left Lower Trans NonUnit: dot form, descending rows)
        const int nblk = (n + BNB - 1) / BNB;
        for (int b = nblk - 1; b >= 0; --b) {
            const int k0 = b * BNB, kb = min(BNB, n - k0);
            // y_blk(j) -= sum over rows below the block of l(i, k0+j) x(i)
            for (int pr = t; pr < kb * nr; pr += BNT) {
                const int j = pr % kb, c = pr / kb;
                T s = T(0);
                for (int i = k0 + kb; i < n; ++i)
                    s = __builtin_fma(L[i + (size_t)(k0 + j) * lda], Y[c * BNMAX + i], s);
                Y[c * BNMAX + k0 + j] -= s;
            }
            for (int idx = t; idx < kb * kb; idx += BNT) {
                const int i = idx % kb, c = idx / kb;
                Ls[i * BLD + c] = L[(k0 + i) + (size_t)(k0 + c) * lda];
            }
            __syncthreads();
            if (t < nr) {
                T* y = Y + t * BNMAX + k0;
                for (int i = kb - 1; i >= 0; --i) {
                    T v = y[i];
                    for (int k = i + 1; k < kb; ++k) v = __builtin_fma(-Ls[k * BLD + i], y[k], v);
                    y[i] = v / Ls[i * BLD + i];
                }
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

template <typename T>
__global__ __launch_bounds__(BNT) void potrs_vbatched_kernel(const BatchItem* __restrict__ items,
                                                             int count, const T* const* __restrict__ Ls,
                                                             T* const* __restrict__ Bs, int* queue) {
    __shared__ T Y[8 * BNMAX];
    __shared__ T Lsm[BNB * BLD];
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= count) break;
        const BatchItem it = items[i];
        if (it.n > 0 && it.nrhs > 0)
            potrs_one<T>(it.n, Ls[it.index], it.lda, Bs[it.index], it.ldb, it.nrhs, Y, Lsm);
    }
}

__global__ void zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Per-device schedules (sorted items) for the factor and the solve, kept between calls: a call
// with the same sizes and leading dimensions as the previous one of its kind reuses the
// device schedule without sorting or copying. The host staging buffer is pinned, and its
// previous copy is waited for before it is refilled.
struct Schedule {
    std::vector<int32_t> key;
    BatchItem* dev = nullptr;
    BatchItem* host = nullptr;  // pinned
    size_t cap = 0;
    int valid = 0;
    hipEvent_t copied = nullptr;
};
struct BatchArea {
    Schedule sched[2];  // 0: factor, 1: solve
    int* queue = nullptr;
};
std::mutex g_bmu;
std::map<int, BatchArea> g_batch;

elx_status_t batch_area(BatchArea** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_bmu);
    BatchArea& a = g_batch[dev];
    if (!a.queue) ELX_HIP(hipMalloc(&a.queue, sizeof(int) * 64));
    *out = &a;
    return ELX_SUCCESS;
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
        for (Schedule& sc : a.sched) {
            if (sc.copied) (void)hipEventSynchronize(sc.copied);
            if (sc.dev) (void)hipFree(sc.dev);
            if (sc.host) (void)hipHostFree(sc.host);
            if (sc.copied) (void)hipEventDestroy(sc.copied);
            sc = Schedule();
        }
        if (a.queue) (void)hipFree(a.queue);
        a.queue = nullptr;
        (void)hipSetDevice(cur);
    }
    g_batch.clear();
}
}  // namespace elx

namespace {

// Valid items in decreasing order of n (ties by index), on the device.
elx_status_t schedule(hipStream_t s, Schedule& sc, int64_t count, const int32_t* n,
                      const int32_t* lda, const int32_t* nrhs, const int32_t* ldb) {
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
    if (sc.dev && key == sc.key) return ELX_SUCCESS;
    if (sc.copied) ELX_HIP(hipEventSynchronize(sc.copied));
    if (sc.cap < (size_t)count) {
        if (sc.dev) {
            ELX_HIP(hipDeviceSynchronize());
            ELX_HIP(hipFree(sc.dev));
            ELX_HIP(hipHostFree(sc.host));
            sc.dev = nullptr;
            sc.host = nullptr;
        }
        const size_t cap = std::max<size_t>((size_t)count, 1024);
        ELX_HIP(hipMalloc(&sc.dev, sizeof(BatchItem) * cap));
        ELX_HIP(hipHostMalloc(reinterpret_cast<void**>(&sc.host), sizeof(BatchItem) * cap, 0));
        sc.cap = cap;
    }
    if (!sc.copied) ELX_HIP(hipEventCreateWithFlags(&sc.copied, hipEventDisableTiming));
    std::vector<int32_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return n[x] > n[y]; });
    for (int64_t k = 0; k < count; ++k) {
        const int32_t i = order[(size_t)k];
        BatchItem& it = sc.host[k];
        it.n = n[i];
        it.lda = lda[i];
        it.index = i;
        it.nrhs = nrhs ? nrhs[i] : 0;
        it.ldb = ldb ? ldb[i] : 0;
        it.pad = 0;
    }
    ELX_HIP(hipMemcpyAsync(sc.dev, sc.host, sizeof(BatchItem) * (size_t)count,
                           hipMemcpyHostToDevice, s));
    ELX_HIP(hipEventRecord(sc.copied, s));
    sc.key.swap(key);
    sc.valid = (int)count;
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
    ELX_CHECK(schedule(s, a->sched[0], count, n, lda, nullptr, nullptr));
    const int valid = a->sched[0].valid;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, a->queue, 1);
    ELX_HIP(hipGetLastError());
    const size_t lds = sizeof(T) * ((size_t)BNMAX * BLD + 2 * BNB * BLD);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&potrf_vbatched_kernel<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
        attr = true;
    }
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&potrf_vbatched_kernel<T>), BNT, lds));
    const elx::DeviceInfo& di = elx::device_info();
    const int G = (int)std::min<int64_t>(count, (int64_t)std::max(per_cu, 1) * di.num_cus);
    double flops = 0.0;
    for (int64_t i = 0; i < count; ++i) flops += (double)n[i] * n[i] * n[i] / 3.0;
    elx::ProfScope prof_(s, elx::kKBatched, flops);
    hipLaunchKernelGGL(potrf_vbatched_kernel<T>, dim3(G), dim3(BNT), lds, s, a->sched[0].dev, valid,
                       dA, dinfo, a->queue);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
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
    ELX_CHECK(schedule(s, a->sched[1], count, n, lda, nrhs, ldb));
    const int valid = a->sched[1].valid;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, a->queue + 1, 1);
    ELX_HIP(hipGetLastError());
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&potrs_vbatched_kernel<T>), BNT, 0));
    const elx::DeviceInfo& di = elx::device_info();
    const int G = (int)std::min<int64_t>(count, (int64_t)std::max(per_cu, 1) * di.num_cus);
    double bytes = 0.0;
    for (int64_t i = 0; i < count; ++i) bytes += (double)n[i] * n[i] * sizeof(T) + 2.0 * n[i] * nrhs[i] * sizeof(T);
    elx::ProfScope prof_(s, elx::kKBatchedSolve, bytes);
    hipLaunchKernelGGL(potrs_vbatched_kernel<T>, dim3(G), dim3(BNT), 0, s, a->sched[1].dev, valid, dA,
                       dB, a->queue + 1);
    ELX_HIP(hipGetLastError());
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

}  // extern "C"
