// Batched variable-size Cholesky factorization and solve (lower, FP64 and FP32).
//
// Each matrix runs the blocked lower Cholesky of El::Cholesky (variant 3,
// src/lapack_like/factor/Cholesky/LowerVariant3.hpp:15-68) with 32-column blocks, in the
// left-looking order: a block column first receives the updates of every earlier block column
// (one GEMM with K = the number of earlier columns, on MFMA), then its 32 x 32 diagonal block
// is factored unblocked (alpha = sqrt(a_jj), a_jj <= 0 -> not HPD, a21 *= 1/alpha, dsyr with
// the update skipped for a zero x_l, as LowerVariant3Unblocked and the reference BLAS dsyr do)
// and the rows below are solved against it (X L11^T = A21 in the column order of the
// reference BLAS dtrsm: x_k *= 1/l_kk, then b_j -= l_jk x_k). The solve follows
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

// One workgroup: the lower Cholesky factor of one matrix, in place.
template <typename T>
__device__ void potrf_one(int n, T* A, int lda, int32_t* info, T* P) {
    typedef typename Mma<T>::v4 v4;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < n; k0 += BNB) {
        const int kb = min(BNB, n - k0), m = n - k0;
        const int nrt = (m + 15) >> 4;
        // (1) C = A(k0:n, k0:k0+kb) - L(k0:n, 0:k0) L(k0:k0+kb, 0:k0)^T into P (row-major, BLD)
        for (int rt0 = wave; rt0 < nrt; rt0 += 8) {
            const int rt1 = rt0 + 4;
            const bool has1 = rt1 < nrt;
            v4 acc[2][2];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = (u ? rt1 : rt0) * 16 + Mma<T>::drow(lane, r);
                        const int col = ct * 16 + lr;
                        acc[u][ct][r] = (row < m && col < kb && (u == 0 || has1))
                                            ? A[(k0 + row) + (size_t)(k0 + col) * lda]
                                            : T(0);
                    }
            const int ra0 = rt0 * 16 + lr, ra1 = rt1 * 16 + lr;
            const bool va0 = ra0 < m, va1 = has1 && ra1 < m;
            const bool vb0 = lr < kb, vb1 = 16 + lr < kb;
            const T* pa0 = A + (k0 + (va0 ? ra0 : 0));
            const T* pa1 = A + (k0 + (va1 ? ra1 : 0));
            const T* pb0 = A + (k0 + (vb0 ? lr : 0));
            const T* pb1 = A + (k0 + (vb1 ? 16 + lr : 0));
#pragma unroll 2
            for (int k = 0; k < k0; k += 4) {
                const size_t off = (size_t)(k + lk) * lda;
                const T a0 = va0 ? -pa0[off] : T(0);
                const T a1 = va1 ? -pa1[off] : T(0);
                const T b0 = vb0 ? pb0[off] : T(0);
                const T b1 = vb1 ? pb1[off] : T(0);
                acc[0][0] = Mma<T>::mma(a0, b0, acc[0][0]);
                acc[0][1] = Mma<T>::mma(a0, b1, acc[0][1]);
                acc[1][0] = Mma<T>::mma(a1, b0, acc[1][0]);
                acc[1][1] = Mma<T>::mma(a1, b1, acc[1][1]);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int ct = 0; ct < 2; ++ct)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = (u ? rt1 : rt0) * 16 + Mma<T>::drow(lane, r);
                        const int col = ct * 16 + lr;
                        if (row < m && col < BNB && (u == 0 || has1)) P[row * BLD + col] = acc[u][ct][r];
                    }
        }
        __syncthreads();
        // (2) the diagonal block, unblocked, by wave 0: lane i holds row i
        __shared__ int s_fail;
        if (wave == 0) {
            T rw[BNB];
#pragma unroll
            for (int c = 0; c < BNB; ++c) rw[c] = (lane < kb) ? P[lane * BLD + c] : T(0);
            int fail = -1;
#pragma unroll
            for (int j = 0; j < BNB; ++j) {
                if (j < kb && fail < 0) {
                    const T ajj = __shfl(rw[j], j);
                    if (ajj <= T(0)) {
                        fail = j;
                    } else {
                        const T alpha = fsqrt(ajj);
                        const T inv = T(1) / alpha;
                        if (lane == j) rw[j] = alpha;
                        else if (lane > j) rw[j] = rw[j] * inv;
                        const T xi = rw[j];
#pragma unroll
                        for (int l = j + 1; l < BNB; ++l) {
                            const T xl = __shfl(rw[j], l);
                            if (lane >= l && xl != T(0)) rw[l] = __builtin_fma(xi, -xl, rw[l]);
                        }
                    }
                }
            }
            if (lane < kb)
#pragma unroll
                for (int c = 0; c < BNB; ++c) P[lane * BLD + c] = rw[c];
            if (lane == 0) s_fail = fail;
        }
        __syncthreads();
        const int fail = s_fail;
        if (fail >= 0) {
            // NonHPDMatrixException at global column k0+fail. The block column as the blocked reference holds it when it throws: the diagonal
            // block's lower triangle after the finished steps, the rows below updated by the
            // earlier block columns
            if (t == 0) *info = k0 + fail + 1;
            for (int idx = t; idx < m * kb; idx += BNT) {
                const int i = idx % m, c = idx / m;
                if (i >= c) A[(k0 + i) + (size_t)(k0 + c) * lda] = P[i * BLD + c];
            }
            __syncthreads();
            return;
        }
        // (3) rows below: X L11^T = A21, one row per thread, row in registers
        for (int row = kb + t; row < m; row += BNT) {
            T x[BNB];
#pragma unroll
            for (int c = 0; c < BNB; ++c) x[c] = (c < kb) ? P[row * BLD + c] : T(0);
#pragma unroll
            for (int k = 0; k < BNB; ++k) {
                if (k < kb) {
                    x[k] = x[k] * (T(1) / P[k * BLD + k]);
#pragma unroll
                    for (int j = k + 1; j < BNB; ++j) {
                        if (j < kb) {
                            const T ljk = P[j * BLD + k];
                            if (ljk != T(0)) x[j] = __builtin_fma(-ljk, x[k], x[j]);
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < BNB; ++c)
                if (c < kb) A[(k0 + row) + (size_t)(k0 + c) * lda] = x[c];
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
    __shared__ int s_i;
    for (;;) {
        if (threadIdx.x == 0) s_i = atomicAdd(queue, 1);
        __syncthreads();
        const int i = s_i;
        __syncthreads();
        if (i >= count) break;
        const BatchItem it = items[i];
        if (threadIdx.x == 0) info[it.index] = 0;
        if (it.n > 0) potrf_one<T>(it.n, As[it.index], it.lda, info + it.index, P);
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

// Per-device batch area: the schedule (sorted items) and the queue counters.
struct BatchArea {
    BatchItem* items = nullptr;
    int* queue = nullptr;
    size_t cap = 0;
};
std::mutex g_bmu;
std::map<int, BatchArea> g_batch;

elx_status_t batch_area(size_t count, BatchArea** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_bmu);
    BatchArea& a = g_batch[dev];
    if (a.cap < count) {
        if (a.items) {
            ELX_HIP(hipDeviceSynchronize());
            ELX_HIP(hipFree(a.items));
            a.items = nullptr;
        }
        const size_t cap = std::max<size_t>(count, 1024);
        ELX_HIP(hipMalloc(&a.items, sizeof(BatchItem) * cap));
        a.cap = cap;
    }
    if (!a.queue) ELX_HIP(hipMalloc(&a.queue, sizeof(int) * 64));
    *out = &a;
    return ELX_SUCCESS;
}

// Host schedule: valid items in decreasing order of n (ties by index), uploaded to the device.
elx_status_t upload_schedule(hipStream_t s, int64_t count, const int32_t* n, const int32_t* lda,
                             const int32_t* nrhs, const int32_t* ldb, BatchArea* a, int* valid) {
    std::vector<BatchItem> items;
    items.reserve((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        BatchItem it;
        std::memset(&it, 0, sizeof(it));
        it.n = n[i];
        it.lda = lda[i];
        it.index = (int32_t)i;
        it.nrhs = nrhs ? nrhs[i] : 0;
        it.ldb = ldb ? ldb[i] : 0;
        ELX_ARG(it.n >= 0 && it.n <= BNMAX && it.lda >= std::max(it.n, 1));
        ELX_ARG(!nrhs || (it.nrhs >= 0 && it.ldb >= std::max(it.n, 1)));
        items.push_back(it);
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const BatchItem& x, const BatchItem& y) { return x.n > y.n; });
    ELX_HIP(hipMemcpyAsync(a->items, items.data(), sizeof(BatchItem) * items.size(),
                           hipMemcpyHostToDevice, s));
    ELX_HIP(hipStreamSynchronize(s));  // the host vector is released on return
    *valid = (int)items.size();
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
    ELX_CHECK(batch_area((size_t)count, &a));
    int valid = 0;
    ELX_CHECK(upload_schedule(s, count, n, lda, nullptr, nullptr, a, &valid));
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, a->queue, 1);
    ELX_HIP(hipGetLastError());
    const size_t lds = sizeof(T) * (size_t)BNMAX * BLD;
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
    hipLaunchKernelGGL(potrf_vbatched_kernel<T>, dim3(G), dim3(BNT), lds, s, a->items, valid, dA,
                       dinfo, a->queue);
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
    ELX_CHECK(batch_area((size_t)count, &a));
    int valid = 0;
    ELX_CHECK(upload_schedule(s, count, n, lda, nrhs, ldb, a, &valid));
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
    hipLaunchKernelGGL(potrs_vbatched_kernel<T>, dim3(G), dim3(BNT), 0, s, a->items, valid, dA, dB,
                       a->queue + 1);
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
