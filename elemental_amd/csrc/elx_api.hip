// Host orchestration of the plain FP64 path and the C ABI entry points.
//
// Every driver enqueues the same blocked schedule as the reference's sequential code (same
// block size and block boundaries), with each BLAS call replaced by a CDNA4 kernel:
//   El::LU                 src/lapack_like/factor/LU.cpp:104-142
//   lin_solve::Overwrite   src/lapack_like/solve/Linear.cpp:30-76,176-183
//   El::Cholesky           src/lapack_like/factor/Cholesky.cpp:26-38 (variant 3)
//   lu::SolveAfter         src/lapack_like/factor/LU/SolveAfter.hpp:66-92
//   cholesky::SolveAfter   src/lapack_like/factor/Cholesky/SolveAfter.hpp:15-43
#include <algorithm>
#include <cstdlib>
#include <map>

#include "elx_common.hpp"

// Work space of elx_dgetrs_noalloc: one fixed array per GPU, held until the process ends. Its
// host-side name is not a device address: host code uses only the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[ELX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
unsigned char* g_na_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
elx_status_t noalloc_scratch(unsigned char** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    ELX_ARG(dev >= 0 && dev < kMaxDevices);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            std::fprintf(stderr, "elemental_amd: no device address for the no-alloc scratch: %s\n",
                         hipGetErrorString(e));
            return ELX_ERROR_HIP;
        }
        g_na_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_na_ptr[dev];
    return ELX_SUCCESS;
}
}  // namespace

namespace elx {

namespace {
std::mutex g_mu;
std::map<int, DeviceInfo> g_info;
std::map<std::pair<int, int>, Scratch> g_scratch;
std::map<int, PanelSync> g_psync;
std::map<int, PanelSync> g_xsync;  // one-XCD panel exchange areas (ordinary memory)
}  // namespace

const DeviceInfo& device_info() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_info.find(dev);
    if (it != g_info.end()) return it->second;
    DeviceInfo d;
    d.device = dev;
    (void)hipDeviceGetAttribute(&d.num_cus, hipDeviceAttributeMultiprocessorCount, dev);
    int lds = 0;
    (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    d.max_lds_per_block = lds;
    (void)hipDeviceGetAttribute(&d.coop, hipDeviceAttributeCooperativeLaunch, dev);
    return g_info.emplace(dev, d).first->second;
}

elx_status_t coop_fits(const void* fn, int block, size_t lds, i64 grid) {
    const DeviceInfo& di = device_info();
    if (!di.coop) return ELX_ERROR_NOT_SUPPORTED;
    hipFuncAttributes fa;
    ELX_HIP(hipFuncGetAttributes(&fa, fn));
    int per_cu = 0;
    ELX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds));
    if (grid > (i64)per_cu * di.num_cus) {
        std::fprintf(stderr,
                     "elemental_amd: cooperative launch of %lld workgroups exceeds %d per CU x %d "
                     "CUs\n",
                     (long long)grid, per_cu, di.num_cus);
        return ELX_ERROR_NOT_SUPPORTED;
    }
    return ELX_SUCCESS;
}

elx_status_t launch_resident(const void* fn, i64 grid, int block, void** args, size_t lds,
                             hipStream_t s) {
    ELX_CHECK(coop_fits(fn, block, lds, grid));
    hipError_t e = hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3((unsigned)block), args,
                                              (unsigned)lds, s);
    if (e == hipErrorIllegalState) {
        (void)hipGetLastError();
        e = hipLaunchKernel(fn, dim3((unsigned)grid), dim3((unsigned)block), args, lds, s);
    }
    if (e != hipSuccess) {
        std::fprintf(stderr, "elemental_amd: launch of a resident kernel failed: %s\n",
                     hipGetErrorString(e));
        return ELX_ERROR_HIP;
    }
    return ELX_SUCCESS;
}

elx_status_t scratch_get(int slot, size_t bytes, void** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    Scratch& sc = g_scratch[{dev, slot}];
    if (sc.bytes < bytes) {
        if (sc.ptr) {
            ELX_HIP(hipDeviceSynchronize());
            ELX_HIP(hipFree(sc.ptr));
            sc.ptr = nullptr;
            sc.bytes = 0;
        }
        const size_t want = std::max<size_t>(bytes, 1 << 20);
        if (hipMalloc(&sc.ptr, want) != hipSuccess) {
            sc.ptr = nullptr;
            return ELX_ERROR_ALLOC;
        }
        sc.bytes = want;
    }
    *out = sc.ptr;
    return ELX_SUCCESS;
}

// The exchange area of the cooperative panel kernel is uncached device memory (coherent
// across XCDs without cache maintenance). Uncached memory that a kernel has read is never
// freed before the process ends; a larger area replaces a smaller one and the old one is
// kept.
elx_status_t panel_sync_get(int max_groups, size_t slot_bytes, PanelSync** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    PanelSync& ps = g_psync[dev];
    if (ps.max_groups < max_groups || ps.slot_bytes < slot_bytes) {
        if (ps.flags) ELX_HIP(hipDeviceSynchronize());
        int64_t* flags = nullptr;
        void* slots = nullptr;
        ELX_HIP(hipExtMallocWithFlags(reinterpret_cast<void**>(&flags),
                                      sizeof(int64_t) * max_groups, hipDeviceMallocUncached));
        ELX_HIP(hipMemset(flags, 0, sizeof(int64_t) * max_groups));
        ELX_HIP(hipExtMallocWithFlags(&slots, slot_bytes, hipDeviceMallocUncached));
        ELX_HIP(hipDeviceSynchronize());
        ps.flags = flags;
        ps.slots = slots;
        ps.max_groups = max_groups;
        ps.slot_bytes = slot_bytes;
        ps.epoch = 0;
    }
    *out = &ps;
    return ELX_SUCCESS;
}

elx_status_t xcd_sync_get(int max_groups, size_t slot_bytes, PanelSync** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    PanelSync& ps = g_xsync[dev];
    if (ps.max_groups < max_groups || ps.slot_bytes < slot_bytes) {
        ELX_HIP(hipDeviceSynchronize());
        if (ps.slots) ELX_HIP(hipFree(ps.slots));
        void* slots = nullptr;
        ELX_HIP(hipMalloc(&slots, slot_bytes));
        ELX_HIP(hipMemset(slots, 0, slot_bytes));
        ELX_HIP(hipDeviceSynchronize());
        ps.slots = slots;
        ps.max_groups = max_groups;
        ps.slot_bytes = slot_bytes;
        ps.epoch = 0;
    }
    *out = &ps;
    return ELX_SUCCESS;
}

// ---------------------------------------------------------------- kernel-class timing
namespace {
struct ProfRec {
    int cls;
    hipEvent_t a, b;
    double work;
};
std::mutex g_pmu;
bool g_prof = false;
std::vector<hipEvent_t> g_evpool;
std::vector<ProfRec> g_pending;
ProfRec g_open[kKCount];
double g_sum_ms[kKCount], g_sum_work[kKCount];
int64_t g_count[kKCount];

hipEvent_t ev_get() {
    if (!g_evpool.empty()) {
        hipEvent_t e = g_evpool.back();
        g_evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void prof_drain() {
    (void)hipDeviceSynchronize();
    for (auto& r : g_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            g_sum_ms[r.cls] += ms;
            g_sum_work[r.cls] += r.work;
            g_count[r.cls] += 1;
        }
        g_evpool.push_back(r.a);
        g_evpool.push_back(r.b);
    }
    g_pending.clear();
}
}  // namespace

bool prof_on() { return g_prof; }

void prof_begin(hipStream_t s, int cls) {
    std::lock_guard<std::mutex> lk(g_pmu);
    ProfRec r{cls, ev_get(), nullptr, 0.0};
    (void)hipEventRecord(r.a, s);
    g_open[cls] = r;
}

void prof_end(hipStream_t s, int cls, double work) {
    std::lock_guard<std::mutex> lk(g_pmu);
    ProfRec r = g_open[cls];
    r.b = ev_get();
    r.work = work;
    (void)hipEventRecord(r.b, s);
    g_pending.push_back(r);
    if (g_pending.size() > 20000) prof_drain();
}

elx_status_t release_all() {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        (void)hipDeviceSynchronize();
        for (auto& kv : g_scratch)
            if (kv.second.ptr) (void)hipFree(kv.second.ptr);
        g_scratch.clear();
        for (auto& kv : g_xsync)
            if (kv.second.slots) (void)hipFree(kv.second.slots);
        g_xsync.clear();
    }
    library_gemm_release();
    batched_release();
    batched_lu_release();
    {
        std::lock_guard<std::mutex> lk(g_pmu);
        prof_drain();
        for (hipEvent_t e : g_evpool) (void)hipEventDestroy(e);
        g_evpool.clear();
    }
    // the panel exchange areas (uncached memory) stay allocated until the process ends
    return ELX_SUCCESS;
}

// No release at process exit: an exit handler of this library can run after the HIP runtime
// has begun its own teardown (the runtime registers its handlers when it initialises, after
// this library was loaded), and HIP calls made then hang or crash the process. Callers that
// want the memory back before exit call elx_finalize.

// Scratch slots
enum { kSlotPairs = 0, kSlotPerm = 1, kSlotRhs = 2, kSlotInfo = 3, kSlotGate = 4 };

elx_status_t gate_slot(double** out) {
    void* p = nullptr;
    ELX_CHECK(scratch_get(kSlotGate, 2 * sizeof(double), &p));
    *out = static_cast<double*>(p);
    return ELX_SUCCESS;
}

namespace {

// General triangular solve: small-kernel diagonal blocks + MFMA GEMM updates. `stop` as in
// the launchers: a recorded failure makes every step a no-op. allow_library false keeps the
// updates on the hand-written GEMM (no library call, so nothing is allocated).
template <typename T>
elx_status_t trsm_blocked(hipStream_t s, int side, int uplo, int orient, int diag, i64 m, i64 n,
                          T alpha, const T* A, i64 lda, T* B, i64 ldb,
                          const int32_t* stop = nullptr, bool allow_library = true) {
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    if (alpha != T(1))
        ELX_CHECK(launch_gemm(s, false, false, m, n, 0, T(0), (const T*)nullptr, 1,
                              (const T*)nullptr, 1, alpha, B, ldb, 0, stop, allow_library));
    const bool trans = (orient != kNormal);
    const i64 bs = 128;
    if (side == kLeft) {
        if (m <= bs) return launch_trsm_small(s, side, uplo, orient, diag, m, n, A, lda, B, ldb, stop);
        // M = op(A); forward when M is lower.
        const bool mlower = (uplo == kLower) != trans;
        const i64 nblk = (m + bs - 1) / bs;
        for (i64 q = 0; q < nblk; ++q) {
            const i64 b = mlower ? q * bs : (nblk - 1 - q) * bs;
            const i64 w = std::min(bs, m - b);
            ELX_CHECK(launch_trsm_small(s, side, uplo, orient, diag, w, n, A + b + b * lda, lda,
                                        B + b, ldb, stop));
            if (mlower) {
                const i64 r0 = b + w, rows = m - r0;
                if (rows <= 0) continue;
                // B[r0:,:] -= M[r0:, b:b+w] B[b:b+w,:]
                if (!trans)
                    ELX_CHECK(launch_gemm(s, false, false, rows, n, w, T(-1), A + r0 + b * lda, lda,
                                          B + b, ldb, T(1), B + r0, ldb, 0, stop, allow_library));
                else
                    ELX_CHECK(launch_gemm(s, true, false, rows, n, w, T(-1), A + b + r0 * lda, lda,
                                          B + b, ldb, T(1), B + r0, ldb, 0, stop, allow_library));
            } else {
                const i64 rows = b;
                if (rows <= 0) continue;
                // B[0:b,:] -= M[0:b, b:b+w] B[b:b+w,:]
                if (!trans)
                    ELX_CHECK(launch_gemm(s, false, false, rows, n, w, T(-1), A + b * lda, lda,
                                          B + b, ldb, T(1), B, ldb, 0, stop, allow_library));
                else
                    ELX_CHECK(launch_gemm(s, true, false, rows, n, w, T(-1), A + b, lda, B + b,
                                          ldb, T(1), B, ldb, 0, stop, allow_library));
            }
        }
        return ELX_SUCCESS;
    }
    // Right side: X op(A) = B, M = op(A) (n x n). Forward over column blocks when M is upper.
    if (n <= bs) return launch_trsm_small(s, side, uplo, orient, diag, m, n, A, lda, B, ldb, stop);
    const bool mupper = (uplo == kUpper) != trans;
    const i64 nblk = (n + bs - 1) / bs;
    for (i64 q = 0; q < nblk; ++q) {
        const i64 b = mupper ? q * bs : (nblk - 1 - q) * bs;
        const i64 w = std::min(bs, n - b);
        ELX_CHECK(launch_trsm_small(s, side, uplo, orient, diag, m, w, A + b + b * lda, lda,
                                    B + b * ldb, ldb, stop));
        if (mupper) {
            const i64 c0 = b + w, cols = n - c0;
            if (cols <= 0) continue;
            // B[:, c0:] -= X[:, b:b+w] M[b:b+w, c0:]
            if (!trans)
                ELX_CHECK(launch_gemm(s, false, false, m, cols, w, T(-1), B + b * ldb, ldb,
                                      A + b + c0 * lda, lda, T(1), B + c0 * ldb, ldb, 0, stop, allow_library));
            else
                ELX_CHECK(launch_gemm(s, false, true, m, cols, w, T(-1), B + b * ldb, ldb,
                                      A + c0 + b * lda, lda, T(1), B + c0 * ldb, ldb, 0, stop, allow_library));
        } else {
            const i64 cols = b;
            if (cols <= 0) continue;
            // B[:, 0:b] -= X[:, b:b+w] M[b:b+w, 0:b]
            if (!trans)
                ELX_CHECK(launch_gemm(s, false, false, m, cols, w, T(-1), B + b * ldb, ldb, A + b,
                                      lda, T(1), B, ldb, 0, stop, allow_library));
            else
                ELX_CHECK(launch_gemm(s, false, true, m, cols, w, T(-1), B + b * ldb, ldb,
                                      A + b * lda, lda, T(1), B, ldb, 0, stop, allow_library));
        }
    }
    return ELX_SUCCESS;
}

// El::LU with partial pivoting (LU.cpp:104-142), or, with B, RowEchelon on [A|B]
// (Linear.cpp:30-76). Every launch after the first takes `info` as its stop word: once a
// panel meets an exactly zero pivot, the rest of the sequence changes nothing, so A (and B)
// are left as the reference leaves them when it throws.
template <typename T>
elx_status_t getrf_impl(hipStream_t s, i64 m, i64 n, T* A, i64 lda, i64 nb, int32_t* dests,
                        int32_t* info, T* B, i64 nrhs, i64 ldb, bool echelon) {
    const i64 minDim = std::min(m, n);
    void* pw = nullptr;
    ELX_CHECK(scratch_get(kSlotPairs, sizeof(int2) * 256, &pw));
    int2* pairs = static_cast<int2*>(pw);
    ELX_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), s));
    for (i64 k = 0; k < minDim; k += nb) {
        const i64 w = std::min(nb, minDim - k);
        T* AB1 = A + k + k * lda;
        ELX_CHECK(launch_lu_panel(s, m - k, w, AB1, lda, dests + k, k, pairs, info));
        // PB.PermuteRows(AB0) (not in the row-echelon form) and PB.PermuteRows(AB2)
        if (!echelon)
            ELX_CHECK(launch_row_moves(s, pairs, (int)(2 * w), A, lda, k, m - k, 0, k, k + w, n, info));
        else
            ELX_CHECK(launch_row_moves(s, pairs, (int)(2 * w), A, lda, k, m - k, k + w, n, 0, 0, info));
        if (B)
            ELX_CHECK(launch_row_moves(s, pairs, (int)(2 * w), B, ldb, k, m - k, 0, nrhs, 0, 0, info));
        const i64 n2 = n - k - w, m2 = m - k - w;
        T* A11 = AB1;
        T* A12 = A + k + (k + w) * lda;
        T* A21 = A + (k + w) + k * lda;
        T* A22 = A + (k + w) + (k + w) * lda;
        ELX_CHECK(launch_trsm_small(s, kLeft, kLower, kNormal, kUnit, w, n2, A11, lda, A12, lda, info));
        if (B)
            ELX_CHECK(launch_trsm_small(s, kLeft, kLower, kNormal, kUnit, w, nrhs, A11, lda, B + k,
                                        ldb, info));
        ELX_CHECK(launch_gemm(s, false, false, m2, n2, w, T(-1), A21, lda, A12, lda, T(1), A22, lda,
                              0, info));
        if (B)
            ELX_CHECK(launch_gemm(s, false, false, m2, nrhs, w, T(-1), A21, lda, B + k, ldb, T(1),
                                  B + k + w, ldb, 0, info));
    }
    return ELX_SUCCESS;
}

// El::LU(Matrix<F>&) without pivoting, src/lapack_like/factor/LU.cpp:21-44: lu::Unb on the
// diagonal block, Trsm(RIGHT,UPPER,NORMAL,NON_UNIT) of A21, Trsm(LEFT,LOWER,NORMAL,UNIT) of
// A12, Gemm for A22; the info word stops the sequence at an exactly zero pivot.
template <typename T>
elx_status_t getrf_nopiv_impl(hipStream_t s, i64 m, i64 n, T* A, i64 lda, i64 nb,
                              int32_t* info) {
    const i64 minDim = std::min(m, n);
    ELX_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), s));
    for (i64 k = 0; k < minDim; k += nb) {
        const i64 w = std::min(nb, minDim - k);
        T* A11 = A + k + k * lda;
        T* A12 = A + k + (k + w) * lda;
        T* A21 = A + (k + w) + k * lda;
        T* A22 = A + (k + w) + (k + w) * lda;
        const i64 m2 = m - k - w, n2 = n - k - w;
        ELX_CHECK(launch_getf2_nopiv(s, w, A11, lda, k, info));
        if (m2 > 0)
            ELX_CHECK(launch_trsm_small(s, kRight, kUpper, kNormal, kNonUnit, m2, w, A11, lda, A21,
                                        lda, info));
        if (n2 > 0)
            ELX_CHECK(launch_trsm_small(s, kLeft, kLower, kNormal, kUnit, w, n2, A11, lda, A12, lda,
                                        info));
        ELX_CHECK(launch_gemm(s, false, false, m2, n2, w, T(-1), A21, lda, A12, lda, T(1), A22, lda,
                              0, info));
    }
    return ELX_SUCCESS;
}

// Row permutation by a swap sequence through work space perm (m int32) and W (m x n values).
template <typename T>
elx_status_t permute_rows_with(hipStream_t s, int inverse, const int32_t* dests, i64 nswaps,
                               i64 m, i64 n, T* B, i64 ldb, int32_t* perm, T* W) {
    if (m <= 0 || n <= 0 || nswaps <= 0) return ELX_SUCCESS;
    ELX_CHECK(launch_swaps_to_perm(s, dests, nswaps, m, perm, inverse));
    ELX_CHECK(launch_gather_rows(s, perm, m, n, B, ldb, W, m));
    ELX_HIP(hipMemcpy2DAsync(B, sizeof(T) * ldb, W, sizeof(T) * m, sizeof(T) * m, n,
                             hipMemcpyDeviceToDevice, s));
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t permute_rows_impl(hipStream_t s, int inverse, const int32_t* dests, i64 nswaps,
                               i64 m, i64 n, T* B, i64 ldb) {
    if (m <= 0 || n <= 0 || nswaps <= 0) return ELX_SUCCESS;
    void* pp = nullptr;
    void* pb = nullptr;
    ELX_CHECK(scratch_get(kSlotPerm, sizeof(int32_t) * (size_t)m, &pp));
    ELX_CHECK(scratch_get(kSlotRhs, sizeof(T) * (size_t)m * (size_t)n, &pb));
    int32_t* perm = static_cast<int32_t*>(pp);
    T* W = static_cast<T*>(pb);
    return permute_rows_with(s, inverse, dests, nswaps, m, n, B, ldb, perm, W);
}

// El::Cholesky, variant 3 (LowerVariant3.hpp:43-68, UpperVariant3.hpp:48-73), stopped by the
// info word at a non-positive pivot.
template <typename T>
elx_status_t potrf_impl(hipStream_t s, int uplo, i64 n, T* A, i64 lda, i64 nb, int32_t* info) {
    ELX_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), s));
    for (i64 k = 0; k < n; k += nb) {
        const i64 w = std::min(nb, n - k);
        const i64 n2 = n - k - w;
        T* A11 = A + k + k * lda;
        T* A22 = A + (k + w) + (k + w) * lda;
        ELX_CHECK(launch_potf2(s, uplo, w, A11, lda, k, info));
        if (n2 <= 0) continue;
        if (uplo == kLower) {
            T* A21 = A + (k + w) + k * lda;
            ELX_CHECK(launch_trsm_small(s, kRight, kLower, kAdjoint, kNonUnit, n2, w, A11, lda, A21,
                                        lda, info));
            ELX_CHECK(launch_gemm(s, false, true, n2, n2, w, T(-1), A21, lda, A21, lda, T(1), A22,
                                  lda, 1, info));
        } else {
            T* A12 = A + k + (k + w) * lda;
            ELX_CHECK(launch_trsm_small(s, kLeft, kUpper, kAdjoint, kNonUnit, w, n2, A11, lda, A12,
                                        lda, info));
            ELX_CHECK(launch_gemm(s, true, false, n2, n2, w, T(-1), A12, lda, A12, lda, T(1), A22,
                                  lda, 2, info));
        }
    }
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t potrs_impl(hipStream_t s, int uplo, i64 n, i64 nrhs, const T* A, i64 lda, T* B,
                        i64 ldb) {
    if (uplo == kLower) {
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, kNormal, kNonUnit, n, nrhs, T(1), A, lda, B, ldb));
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, kAdjoint, kNonUnit, n, nrhs, T(1), A, lda, B, ldb));
    } else {
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kAdjoint, kNonUnit, n, nrhs, T(1), A, lda, B, ldb));
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, T(1), A, lda, B, ldb));
    }
    return ELX_SUCCESS;
}

// ------------------------------------------------------------ typed bodies of the C ABI
template <typename T>
elx_status_t getrf_api(void* stream, int64_t m, int64_t n, T* dA, int64_t lda, int64_t nb,
                       int32_t* dDests, int32_t* dInfo) {
    ELX_ARG(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1) && nb >= 1);
    ELX_ARG(dInfo != nullptr);
    ELX_ARG(std::min(m, n) == 0 || (dA && dDests));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (std::min(m, n) == 0) {
        ELX_HIP(hipMemsetAsync(dInfo, 0, sizeof(int32_t), s));
        return ELX_SUCCESS;
    }
    // blocks wider than the panel's 128 columns are factored as 128-column panels with the
    // same pivots (the panel is column-by-column); only the trailing-update grouping differs
    return getrf_impl<T>(s, m, n, dA, lda, std::min<int64_t>(nb, 128), dDests, dInfo, nullptr, 0,
                         0, false);
}

template <typename T>
elx_status_t getrf_nopiv_api(void* stream, int64_t m, int64_t n, T* dA, int64_t lda, int64_t nb,
                             int32_t* dInfo) {
    ELX_ARG(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1) && nb >= 1 && dInfo != nullptr);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (std::min(m, n) == 0) {
        ELX_HIP(hipMemsetAsync(dInfo, 0, sizeof(int32_t), s));
        return ELX_SUCCESS;
    }
    return getrf_nopiv_impl<T>(s, m, n, dA, lda, std::min<int64_t>(nb, 128), dInfo);
}

// lu::SolveAfter's sequence. perm / W (m int32, n x nrhs values) are the row permutation's
// work space when given (the no-alloc entry point), else the library's scratch slots; with
// allow_library false every update stays on the hand-written GEMM.
template <typename T>
elx_status_t getrs_body(hipStream_t s, int orient, int64_t n, int64_t nrhs, const T* dA,
                        int64_t lda, const int32_t* dDests, T* dB, int64_t ldb, int32_t* perm,
                        T* W, bool allow_library) {
    auto permute = [&](int inverse) -> elx_status_t {
        if (!dDests) return ELX_SUCCESS;
        if (perm) return permute_rows_with(s, inverse, dDests, n, n, nrhs, dB, ldb, perm, W);
        return permute_rows_impl(s, inverse, dDests, n, n, nrhs, dB, ldb);
    };
    if (orient == kNormal) {
        ELX_CHECK(permute(0));
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, kNormal, kUnit, n, nrhs, T(1), dA, lda, dB, ldb,
                               nullptr, allow_library));
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, T(1), dA, lda, dB,
                               ldb, nullptr, allow_library));
    } else {
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, orient, kNonUnit, n, nrhs, T(1), dA, lda, dB, ldb,
                               nullptr, allow_library));
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, orient, kUnit, n, nrhs, T(1), dA, lda, dB, ldb,
                               nullptr, allow_library));
        ELX_CHECK(permute(1));
    }
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t getrs_api(void* stream, int orient, int64_t n, int64_t nrhs, const T* dA,
                       int64_t lda, const int32_t* dDests, T* dB, int64_t ldb) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    return getrs_body<T>(static_cast<hipStream_t>(stream), orient, n, nrhs, dA, lda, dDests, dB,
                         ldb, nullptr, nullptr, true);
}

template <typename T>
elx_status_t getrf_full_api(void* stream, int64_t m, int64_t n, T* dA, int64_t lda, int32_t* dP,
                            int32_t* dQ, int32_t* dInfo) {
    ELX_ARG(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1) && dInfo != nullptr);
    ELX_ARG(std::min(m, n) == 0 || (dA && dP && dQ));
    return getrf_full(static_cast<hipStream_t>(stream), m, n, dA, lda, dP, dQ, dInfo);
}

// lu::SolveAfter(orientation, A, P, Q, B), src/lapack_like/factor/LU/SolveAfter.hpp:123-151
template <typename T>
elx_status_t getrs_full_api(void* stream, int orient, int64_t n, int64_t nrhs, const T* dA,
                            int64_t lda, const int32_t* dP, const int32_t* dQ, T* dB,
                            int64_t ldb) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    ELX_ARG(dA && dP && dQ && dB);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (orient == kNormal) {
        ELX_CHECK(permute_rows_impl(s, 0, dP, n, n, nrhs, dB, ldb));
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, kNormal, kUnit, n, nrhs, T(1), dA, lda, dB, ldb));
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, T(1), dA, lda, dB, ldb));
        ELX_CHECK(permute_rows_impl(s, 1, dQ, n, n, nrhs, dB, ldb));
    } else {
        ELX_CHECK(permute_rows_impl(s, 0, dQ, n, n, nrhs, dB, ldb));
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, orient, kNonUnit, n, nrhs, T(1), dA, lda, dB, ldb));
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, orient, kUnit, n, nrhs, T(1), dA, lda, dB, ldb));
        ELX_CHECK(permute_rows_impl(s, 1, dP, n, n, nrhs, dB, ldb));
    }
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t gesv_api(void* stream, int64_t n, int64_t nrhs, T* dA, int64_t lda, T* dB,
                      int64_t ldb, int64_t nb, int32_t* dDests, int32_t* dInfo) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(nb >= 1 && dInfo != nullptr);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {
        ELX_HIP(hipMemsetAsync(dInfo, 0, sizeof(int32_t), s));
        return ELX_SUCCESS;
    }
    int32_t* dests = dDests;
    void* pd = nullptr;
    if (!dests) {
        ELX_CHECK(scratch_get(kSlotInfo, sizeof(int32_t) * (size_t)n, &pd));
        dests = static_cast<int32_t*>(pd);
    }
    ELX_CHECK(getrf_impl<T>(s, n, n, dA, lda, std::min<int64_t>(nb, 128), dests, dInfo,
                            nrhs > 0 ? dB : nullptr, nrhs, ldb, true));
    if (nrhs > 0)
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, T(1), (const T*)dA,
                               lda, dB, ldb, dInfo));
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t potrf_api(void* stream, int uplo, int64_t n, T* dA, int64_t lda, int64_t nb,
                       int32_t* dInfo) {
    ELX_ARG(n >= 0 && lda >= std::max<int64_t>(n, 1) && nb >= 1 && dInfo);
    ELX_ARG(uplo == kLower || uplo == kUpper);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {
        ELX_HIP(hipMemsetAsync(dInfo, 0, sizeof(int32_t), s));
        return ELX_SUCCESS;
    }
    return potrf_impl<T>(s, uplo, n, dA, lda, std::min<int64_t>(nb, 128), dInfo);
}

template <typename T>
elx_status_t potrs_api(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, const T* dA,
                       int64_t lda, T* dB, int64_t ldb) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG((uplo == kLower || uplo == kUpper) && orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    // TRANSPOSE conjugates B before and after: the identity for real data.
    return potrs_impl<T>(static_cast<hipStream_t>(stream), uplo, n, nrhs, dA, lda, dB, ldb);
}

template <typename T>
elx_status_t posv_api(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, T* dA,
                      int64_t lda, T* dB, int64_t ldb, int64_t nb, int32_t* dInfo) {
    ELX_CHECK(potrf_api<T>(stream, uplo, n, dA, lda, nb, dInfo));
    ELX_ARG(n >= 0 && nrhs >= 0 && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the solve is skipped on the device when the factorization failed (HPD.cpp throws first)
    if (uplo == kLower) {
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, kNormal, kNonUnit, n, nrhs, T(1), (const T*)dA,
                               lda, dB, ldb, dInfo));
        ELX_CHECK(trsm_blocked(s, kLeft, kLower, kAdjoint, kNonUnit, n, nrhs, T(1), (const T*)dA,
                               lda, dB, ldb, dInfo));
    } else {
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kAdjoint, kNonUnit, n, nrhs, T(1), (const T*)dA,
                               lda, dB, ldb, dInfo));
        ELX_CHECK(trsm_blocked(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, T(1), (const T*)dA,
                               lda, dB, ldb, dInfo));
    }
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t gemm_api(void* stream, int transA, int transB, int64_t m, int64_t n, int64_t k,
                      T alpha, const T* dA, int64_t lda, const T* dB, int64_t ldb, T beta, T* dC,
                      int64_t ldc) {
    ELX_ARG(m >= 0 && n >= 0 && k >= 0);
    ELX_ARG(ldc >= std::max<int64_t>(m, 1));
    ELX_ARG(lda >= std::max<int64_t>(transA == kNormal ? m : k, 1));
    ELX_ARG(ldb >= std::max<int64_t>(transB == kNormal ? k : n, 1));
    return launch_gemm(static_cast<hipStream_t>(stream), transA != kNormal, transB != kNormal, m, n,
                       k, alpha, dA, lda, dB, ldb, beta, dC, ldc, 0, nullptr);
}

template <typename T>
elx_status_t trsm_api(void* stream, int side, int uplo, int orient, int diag, int64_t m,
                      int64_t n, T alpha, const T* dA, int64_t lda, T* dB, int64_t ldb) {
    ELX_ARG(m >= 0 && n >= 0 && ldb >= std::max<int64_t>(m, 1));
    ELX_ARG(lda >= std::max<int64_t>(side == kLeft ? m : n, 1));
    ELX_ARG((side == kLeft || side == kRight) && (uplo == kLower || uplo == kUpper));
    ELX_ARG(orient >= 0 && orient <= 2 && (diag == kUnit || diag == kNonUnit));
    return trsm_blocked(static_cast<hipStream_t>(stream), side, uplo, orient, diag, m, n, alpha,
                        dA, lda, dB, ldb);
}

template <typename T>
elx_status_t syrk_api(void* stream, int uplo, int orient, int64_t n, int64_t k, T alpha,
                      const T* dA, int64_t lda, T beta, T* dC, int64_t ldc) {
    ELX_ARG(n >= 0 && k >= 0 && ldc >= std::max<int64_t>(n, 1));
    ELX_ARG(lda >= std::max<int64_t>(orient == kNormal ? n : k, 1));
    ELX_ARG(uplo == kLower || uplo == kUpper);
    const bool t = orient != kNormal;
    // NORMAL: C := alpha A A^T + beta C ; otherwise C := alpha A^T A + beta C
    return launch_gemm(static_cast<hipStream_t>(stream), t, !t, n, n, k, alpha, dA, lda, dA, lda,
                       beta, dC, ldc, uplo == kLower ? 1 : 2, nullptr);
}

template <typename T>
elx_status_t permute_rows_api(void* stream, int inverse, const int32_t* dDests, int64_t nswaps,
                              int64_t m, int64_t n, T* dA, int64_t lda) {
    ELX_ARG(m >= 0 && n >= 0 && nswaps >= 0 && nswaps <= m && lda >= std::max<int64_t>(m, 1));
    return permute_rows_impl<T>(static_cast<hipStream_t>(stream), inverse, dDests, nswaps, m, n,
                                dA, lda);
}

}  // namespace

elx_status_t dtrsm(hipStream_t s, int side, int uplo, int orient, int diag, i64 m, i64 n,
                   double alpha, const double* A, i64 lda, double* B, i64 ldb) {
    return trsm_blocked<double>(s, side, uplo, orient, diag, m, n, alpha, A, lda, B, ldb);
}

elx_status_t strsm(hipStream_t s, int side, int uplo, int orient, int diag, i64 m, i64 n,
                   float alpha, const float* A, i64 lda, float* B, i64 ldb) {
    return trsm_blocked<float>(s, side, uplo, orient, diag, m, n, alpha, A, lda, B, ldb);
}

}  // namespace elx

using namespace elx;

extern "C" {

const char* elx_version(void) { return "elemental_amd 0.1 (gfx950)"; }

int elx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

elx_status_t elx_finalize(void) { return release_all(); }

void elx_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(g_pmu);
    if (!on && g_prof) prof_drain();
    g_prof = (on != 0);
}

void elx_profile_reset(void) {
    std::lock_guard<std::mutex> lk(g_pmu);
    prof_drain();
    for (int c = 0; c < kKCount; ++c) {
        g_sum_ms[c] = 0.0;
        g_sum_work[c] = 0.0;
        g_count[c] = 0;
    }
}

elx_status_t elx_profile_query(int cls, int64_t* launches, double* total_ms, double* work) {
    ELX_ARG(cls >= 0 && cls < kKCount);
    std::lock_guard<std::mutex> lk(g_pmu);
    prof_drain();
    if (launches) *launches = g_count[cls];
    if (total_ms) *total_ms = g_sum_ms[cls];
    if (work) *work = g_sum_work[cls];
    return ELX_SUCCESS;
}

elx_status_t elx_dgetrf(void* stream, int64_t m, int64_t n, double* dA, int64_t lda, int64_t nb,
                        int32_t* dDests, int32_t* dInfo) {
    return getrf_api<double>(stream, m, n, dA, lda, nb, dDests, dInfo);
}
elx_status_t elx_sgetrf(void* stream, int64_t m, int64_t n, float* dA, int64_t lda, int64_t nb,
                        int32_t* dDests, int32_t* dInfo) {
    return getrf_api<float>(stream, m, n, dA, lda, nb, dDests, dInfo);
}

elx_status_t elx_dgetrf_nopiv(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                              int64_t nb, int32_t* dInfo) {
    return getrf_nopiv_api<double>(stream, m, n, dA, lda, nb, dInfo);
}
elx_status_t elx_sgetrf_nopiv(void* stream, int64_t m, int64_t n, float* dA, int64_t lda,
                              int64_t nb, int32_t* dInfo) {
    return getrf_nopiv_api<float>(stream, m, n, dA, lda, nb, dInfo);
}

elx_status_t elx_dgetrf_full(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                             int32_t* dP, int32_t* dQ, int32_t* dInfo) {
    return getrf_full_api<double>(stream, m, n, dA, lda, dP, dQ, dInfo);
}
elx_status_t elx_sgetrf_full(void* stream, int64_t m, int64_t n, float* dA, int64_t lda,
                             int32_t* dP, int32_t* dQ, int32_t* dInfo) {
    return getrf_full_api<float>(stream, m, n, dA, lda, dP, dQ, dInfo);
}
elx_status_t elx_dgetrs_full(void* stream, int orient, int64_t n, int64_t nrhs, const double* dA,
                             int64_t lda, const int32_t* dP, const int32_t* dQ, double* dB,
                             int64_t ldb) {
    return getrs_full_api<double>(stream, orient, n, nrhs, dA, lda, dP, dQ, dB, ldb);
}
elx_status_t elx_sgetrs_full(void* stream, int orient, int64_t n, int64_t nrhs, const float* dA,
                             int64_t lda, const int32_t* dP, const int32_t* dQ, float* dB,
                             int64_t ldb) {
    return getrs_full_api<float>(stream, orient, n, nrhs, dA, lda, dP, dQ, dB, ldb);
}

elx_status_t elx_dgetrs(void* stream, int orient, int64_t n, int64_t nrhs, const double* dA,
                        int64_t lda, const int32_t* dDests, double* dB, int64_t ldb) {
    return getrs_api<double>(stream, orient, n, nrhs, dA, lda, dDests, dB, ldb);
}
elx_status_t elx_sgetrs(void* stream, int orient, int64_t n, int64_t nrhs, const float* dA,
                        int64_t lda, const int32_t* dDests, float* dB, int64_t ldb) {
    return getrs_api<float>(stream, orient, n, nrhs, dA, lda, dDests, dB, ldb);
}

// lu::SolveAfter without allocation: the row permutation's work space is the fixed array
// `scratch` below, reached through its device address; every update runs on the hand-written
// kernels (no library call). Calls that overlap on one GPU share the array and race.
elx_status_t elx_dgetrs_noalloc(void* stream, int orient, int64_t n, int64_t nrhs,
                                const double* dA, int64_t lda, const int32_t* dDests, double* dB,
                                int64_t ldb) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    int32_t* perm = nullptr;
    double* W = nullptr;
    if (dDests) {
        const size_t perm_bytes = (sizeof(int32_t) * (size_t)n + 255) / 256 * 256;
        const double need = (double)perm_bytes + (double)sizeof(double) * (double)n * (double)nrhs;
        if (need > (double)ELX_NOALLOC_SCRATCH_BYTES) {
            std::fprintf(stderr,
                         "elemental_amd: elx_dgetrs_noalloc needs %.0f bytes of work space for "
                         "n=%lld, nrhs=%lld; its fixed scratch holds %lld bytes "
                         "(4n rounded up to 256, plus 8 n nrhs)\n",
                         need, (long long)n, (long long)nrhs,
                         (long long)ELX_NOALLOC_SCRATCH_BYTES);
            return ELX_ERROR_INVALID_ARG;
        }
        unsigned char* base = nullptr;
        ELX_CHECK(noalloc_scratch(&base));
        perm = reinterpret_cast<int32_t*>(base);
        W = reinterpret_cast<double*>(base + perm_bytes);
    }
    return getrs_body<double>(static_cast<hipStream_t>(stream), orient, n, nrhs, dA, lda, dDests,
                              dB, ldb, perm, W, false);
}

elx_status_t elx_dgesv(void* stream, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                       double* dB, int64_t ldb, int64_t nb, int32_t* dDests, int32_t* dInfo) {
    return gesv_api<double>(stream, n, nrhs, dA, lda, dB, ldb, nb, dDests, dInfo);
}
elx_status_t elx_sgesv(void* stream, int64_t n, int64_t nrhs, float* dA, int64_t lda, float* dB,
                       int64_t ldb, int64_t nb, int32_t* dDests, int32_t* dInfo) {
    return gesv_api<float>(stream, n, nrhs, dA, lda, dB, ldb, nb, dDests, dInfo);
}

elx_status_t elx_dpotrf(void* stream, int uplo, int64_t n, double* dA, int64_t lda, int64_t nb,
                        int32_t* dInfo) {
    return potrf_api<double>(stream, uplo, n, dA, lda, nb, dInfo);
}
elx_status_t elx_spotrf(void* stream, int uplo, int64_t n, float* dA, int64_t lda, int64_t nb,
                        int32_t* dInfo) {
    return potrf_api<float>(stream, uplo, n, dA, lda, nb, dInfo);
}

elx_status_t elx_dpotrs(void* stream, int uplo, int orient, int64_t n, int64_t nrhs,
                        const double* dA, int64_t lda, double* dB, int64_t ldb) {
    return potrs_api<double>(stream, uplo, orient, n, nrhs, dA, lda, dB, ldb);
}
elx_status_t elx_spotrs(void* stream, int uplo, int orient, int64_t n, int64_t nrhs,
                        const float* dA, int64_t lda, float* dB, int64_t ldb) {
    return potrs_api<float>(stream, uplo, orient, n, nrhs, dA, lda, dB, ldb);
}

elx_status_t elx_dposv(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, double* dA,
                       int64_t lda, double* dB, int64_t ldb, int64_t nb, int32_t* dInfo) {
    return posv_api<double>(stream, uplo, orient, n, nrhs, dA, lda, dB, ldb, nb, dInfo);
}
elx_status_t elx_sposv(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, float* dA,
                       int64_t lda, float* dB, int64_t ldb, int64_t nb, int32_t* dInfo) {
    return posv_api<float>(stream, uplo, orient, n, nrhs, dA, lda, dB, ldb, nb, dInfo);
}

elx_status_t elx_dgemm(void* stream, int transA, int transB, int64_t m, int64_t n, int64_t k,
                       double alpha, const double* dA, int64_t lda, const double* dB,
                       int64_t ldb, double beta, double* dC, int64_t ldc) {
    return gemm_api<double>(stream, transA, transB, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
}
elx_status_t elx_sgemm(void* stream, int transA, int transB, int64_t m, int64_t n, int64_t k,
                       float alpha, const float* dA, int64_t lda, const float* dB, int64_t ldb,
                       float beta, float* dC, int64_t ldc) {
    return gemm_api<float>(stream, transA, transB, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
}

elx_status_t elx_dtrsm(void* stream, int side, int uplo, int orient, int diag, int64_t m,
                       int64_t n, double alpha, const double* dA, int64_t lda, double* dB,
                       int64_t ldb) {
    return trsm_api<double>(stream, side, uplo, orient, diag, m, n, alpha, dA, lda, dB, ldb);
}
elx_status_t elx_strsm(void* stream, int side, int uplo, int orient, int diag, int64_t m,
                       int64_t n, float alpha, const float* dA, int64_t lda, float* dB,
                       int64_t ldb) {
    return trsm_api<float>(stream, side, uplo, orient, diag, m, n, alpha, dA, lda, dB, ldb);
}

elx_status_t elx_dsyrk(void* stream, int uplo, int orient, int64_t n, int64_t k, double alpha,
                       const double* dA, int64_t lda, double beta, double* dC, int64_t ldc) {
    return syrk_api<double>(stream, uplo, orient, n, k, alpha, dA, lda, beta, dC, ldc);
}
elx_status_t elx_ssyrk(void* stream, int uplo, int orient, int64_t n, int64_t k, float alpha,
                       const float* dA, int64_t lda, float beta, float* dC, int64_t ldc) {
    return syrk_api<float>(stream, uplo, orient, n, k, alpha, dA, lda, beta, dC, ldc);
}

elx_status_t elx_dpermute_rows(void* stream, int inverse, const int32_t* dDests, int64_t nswaps,
                               int64_t m, int64_t n, double* dA, int64_t lda) {
    return permute_rows_api<double>(stream, inverse, dDests, nswaps, m, n, dA, lda);
}
elx_status_t elx_spermute_rows(void* stream, int inverse, const int32_t* dDests, int64_t nswaps,
                               int64_t m, int64_t n, float* dA, int64_t lda) {
    return permute_rows_api<float>(stream, inverse, dDests, nswaps, m, n, dA, lda);
}

}  // extern "C"
