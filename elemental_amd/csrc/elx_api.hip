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
#include <map>

#include "elx_common.hpp"

namespace elx {

namespace {
std::mutex g_mu;
std::map<int, DeviceInfo> g_info;
std::map<std::pair<int, int>, Scratch> g_scratch;
std::map<int, PanelSync> g_psync;
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
    std::lock_guard<std::mutex> lk(g_mu);
    (void)hipDeviceSynchronize();
    for (auto& kv : g_scratch)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    g_scratch.clear();
    library_gemm_release();
    // the panel exchange areas (uncached memory) stay allocated until the process ends
    return ELX_SUCCESS;
}

// General triangular solve: small-kernel diagonal blocks + MFMA GEMM updates.
elx_status_t dtrsm(hipStream_t s, int side, int uplo, int orient, int diag, i64 m, i64 n,
                   double alpha, const double* A, i64 lda, double* B, i64 ldb) {
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    if (alpha != 1.0) ELX_CHECK(launch_dgemm(s, false, false, m, n, 0, 0.0, nullptr, 1, nullptr, 1, alpha, B, ldb, 0));
    const bool trans = (orient != kNormal);
    const i64 bs = 128;
    if (side == kLeft) {
        if (m <= bs) return launch_trsm_small(s, side, uplo, orient, diag, m, n, A, lda, B, ldb);
        // M = op(A); forward when M is lower.
        const bool mlower = (uplo == kLower) != trans;
        const i64 nblk = (m + bs - 1) / bs;
        for (i64 q = 0; q < nblk; ++q) {
            const i64 b = mlower ? q * bs : (nblk - 1 - q) * bs;
            const i64 w = std::min(bs, m - b);
            ELX_CHECK(launch_trsm_small(s, side, uplo, orient, diag, w, n, A + b + b * lda, lda,
                                        B + b, ldb));
            if (mlower) {
                const i64 r0 = b + w, rows = m - r0;
                if (rows <= 0) continue;
                // B[r0:,:] -= M[r0:, b:b+w] B[b:b+w,:]
                if (!trans)
                    ELX_CHECK(launch_dgemm(s, false, false, rows, n, w, -1.0, A + r0 + b * lda, lda,
                                           B + b, ldb, 1.0, B + r0, ldb, 0));
                else
                    ELX_CHECK(launch_dgemm(s, true, false, rows, n, w, -1.0, A + b + r0 * lda, lda,
                                           B + b, ldb, 1.0, B + r0, ldb, 0));
            } else {
                const i64 rows = b;
                if (rows <= 0) continue;
                // B[0:b,:] -= M[0:b, b:b+w] B[b:b+w,:]
                if (!trans)
                    ELX_CHECK(launch_dgemm(s, false, false, rows, n, w, -1.0, A + b * lda, lda, B + b,
                                           ldb, 1.0, B, ldb, 0));
                else
                    ELX_CHECK(launch_dgemm(s, true, false, rows, n, w, -1.0, A + b, lda, B + b, ldb,
                                           1.0, B, ldb, 0));
            }
        }
        return ELX_SUCCESS;
    }
    // Right side: X op(A) = B, M = op(A) (n x n). Forward over column blocks when M is upper.
    if (n <= bs) return launch_trsm_small(s, side, uplo, orient, diag, m, n, A, lda, B, ldb);
    const bool mupper = (uplo == kUpper) != trans;
    const i64 nblk = (n + bs - 1) / bs;
    for (i64 q = 0; q < nblk; ++q) {
        const i64 b = mupper ? q * bs : (nblk - 1 - q) * bs;
        const i64 w = std::min(bs, n - b);
        ELX_CHECK(launch_trsm_small(s, side, uplo, orient, diag, m, w, A + b + b * lda, lda,
                                    B + b * ldb, ldb));
        if (mupper) {
            const i64 c0 = b + w, cols = n - c0;
            if (cols <= 0) continue;
            // B[:, c0:] -= X[:, b:b+w] M[b:b+w, c0:]
            if (!trans)
                ELX_CHECK(launch_dgemm(s, false, false, m, cols, w, -1.0, B + b * ldb, ldb,
                                       A + b + c0 * lda, lda, 1.0, B + c0 * ldb, ldb, 0));
            else
                ELX_CHECK(launch_dgemm(s, false, true, m, cols, w, -1.0, B + b * ldb, ldb,
                                       A + c0 + b * lda, lda, 1.0, B + c0 * ldb, ldb, 0));
        } else {
            const i64 cols = b;
            if (cols <= 0) continue;
            // B[:, 0:b] -= X[:, b:b+w] M[b:b+w, 0:b]
            if (!trans)
                ELX_CHECK(launch_dgemm(s, false, false, m, cols, w, -1.0, B + b * ldb, ldb, A + b,
                                       lda, 1.0, B, ldb, 0));
            else
                ELX_CHECK(launch_dgemm(s, false, true, m, cols, w, -1.0, B + b * ldb, ldb,
                                       A + b * lda, lda, 1.0, B, ldb, 0));
        }
    }
    return ELX_SUCCESS;
}

// Scratch slots
enum { kSlotPairs = 0, kSlotPerm = 1, kSlotRhs = 2, kSlotInfo = 3 };

static elx_status_t getrf_impl(hipStream_t s, i64 m, i64 n, double* A, i64 lda, i64 nb,
                               int32_t* dests, int32_t* info, double* B, i64 nrhs, i64 ldb,
                               bool echelon) {
    const i64 minDim = std::min(m, n);
    void* pw = nullptr;
    ELX_CHECK(scratch_get(kSlotPairs, sizeof(int2) * 256, &pw));
    int2* pairs = static_cast<int2*>(pw);
    ELX_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), s));
    for (i64 k = 0; k < minDim; k += nb) {
        const i64 w = std::min(nb, minDim - k);
        double* AB1 = A + k + k * lda;
        ELX_CHECK(launch_lu_panel_f64(s, m - k, w, AB1, lda, dests + k, k, pairs, info));
        // PB.PermuteRows(AB0) (not in the row-echelon form) and PB.PermuteRows(AB2)
        if (!echelon)
            ELX_CHECK(launch_row_moves_f64(s, pairs, (int)(2 * w), A, lda, k, 0, k, k + w, n));
        else
            ELX_CHECK(launch_row_moves_f64(s, pairs, (int)(2 * w), A, lda, k, k + w, n, 0, 0));
        if (B) ELX_CHECK(launch_row_moves_f64(s, pairs, (int)(2 * w), B, ldb, k, 0, nrhs, 0, 0));
        const i64 n2 = n - k - w, m2 = m - k - w;
        double* A11 = AB1;
        double* A12 = A + k + (k + w) * lda;
        double* A21 = A + (k + w) + k * lda;
        double* A22 = A + (k + w) + (k + w) * lda;
        ELX_CHECK(launch_trsm_small(s, kLeft, kLower, kNormal, kUnit, w, n2, A11, lda, A12, lda));
        if (B) ELX_CHECK(launch_trsm_small(s, kLeft, kLower, kNormal, kUnit, w, nrhs, A11, lda, B + k, ldb));
        ELX_CHECK(launch_dgemm(s, false, false, m2, n2, w, -1.0, A21, lda, A12, lda, 1.0, A22, lda, 0));
        if (B)
            ELX_CHECK(launch_dgemm(s, false, false, m2, nrhs, w, -1.0, A21, lda, B + k, ldb, 1.0,
                                   B + k + w, ldb, 0));
    }
    return ELX_SUCCESS;
}

static elx_status_t permute_rows_impl(hipStream_t s, int inverse, const int32_t* dests,
                                      i64 nswaps, i64 m, i64 n, double* B, i64 ldb) {
    if (m <= 0 || n <= 0 || nswaps <= 0) return ELX_SUCCESS;
    void* pp = nullptr;
    void* pb = nullptr;
    ELX_CHECK(scratch_get(kSlotPerm, sizeof(int32_t) * (size_t)m, &pp));
    ELX_CHECK(scratch_get(kSlotRhs, sizeof(double) * (size_t)m * (size_t)n, &pb));
    int32_t* perm = static_cast<int32_t*>(pp);
    double* W = static_cast<double*>(pb);
    ELX_CHECK(launch_swaps_to_perm(s, dests, nswaps, m, perm, inverse));
    ELX_CHECK(launch_gather_rows_f64(s, perm, m, n, B, ldb, W, m));
    ELX_HIP(hipMemcpy2DAsync(B, sizeof(double) * ldb, W, sizeof(double) * m, sizeof(double) * m, n,
                             hipMemcpyDeviceToDevice, s));
    return ELX_SUCCESS;
}

static elx_status_t potrf_impl(hipStream_t s, int uplo, i64 n, double* A, i64 lda, i64 nb,
                               int32_t* info) {
    ELX_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), s));
    for (i64 k = 0; k < n; k += nb) {
        const i64 w = std::min(nb, n - k);
        const i64 n2 = n - k - w;
        double* A11 = A + k + k * lda;
        double* A22 = A + (k + w) + (k + w) * lda;
        ELX_CHECK(launch_potf2_f64(s, uplo, w, A11, lda, k, info));
        if (n2 <= 0) continue;
        if (uplo == kLower) {
            double* A21 = A + (k + w) + k * lda;
            ELX_CHECK(launch_trsm_small(s, kRight, kLower, kAdjoint, kNonUnit, n2, w, A11, lda, A21, lda));
            ELX_CHECK(launch_dgemm(s, false, true, n2, n2, w, -1.0, A21, lda, A21, lda, 1.0, A22, lda, 1));
        } else {
            double* A12 = A + k + (k + w) * lda;
            ELX_CHECK(launch_trsm_small(s, kLeft, kUpper, kAdjoint, kNonUnit, w, n2, A11, lda, A12, lda));
            ELX_CHECK(launch_dgemm(s, true, false, n2, n2, w, -1.0, A12, lda, A12, lda, 1.0, A22, lda, 2));
        }
    }
    return ELX_SUCCESS;
}

static elx_status_t potrs_impl(hipStream_t s, int uplo, i64 n, i64 nrhs, const double* A,
                               i64 lda, double* B, i64 ldb) {
    if (uplo == kLower) {
        ELX_CHECK(dtrsm(s, kLeft, kLower, kNormal, kNonUnit, n, nrhs, 1.0, A, lda, B, ldb));
        ELX_CHECK(dtrsm(s, kLeft, kLower, kAdjoint, kNonUnit, n, nrhs, 1.0, A, lda, B, ldb));
    } else {
        ELX_CHECK(dtrsm(s, kLeft, kUpper, kAdjoint, kNonUnit, n, nrhs, 1.0, A, lda, B, ldb));
        ELX_CHECK(dtrsm(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, 1.0, A, lda, B, ldb));
    }
    return ELX_SUCCESS;
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
    ELX_ARG(m >= 0 && n >= 0 && lda >= std::max<int64_t>(m, 1) && nb >= 1 && nb <= 128);
    ELX_ARG(dInfo != nullptr);
    ELX_ARG(std::min(m, n) == 0 || (dA && dDests));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (std::min(m, n) == 0) {
        ELX_HIP(hipMemsetAsync(dInfo, 0, sizeof(int32_t), s));
        return ELX_SUCCESS;
    }
    return getrf_impl(s, m, n, dA, lda, nb, dDests, dInfo, nullptr, 0, 0, false);
}

elx_status_t elx_dgetrs(void* stream, int orient, int64_t n, int64_t nrhs, const double* dA,
                        int64_t lda, const int32_t* dDests, double* dB, int64_t ldb) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (orient == kNormal) {
        ELX_CHECK(permute_rows_impl(s, 0, dDests, n, n, nrhs, dB, ldb));
        ELX_CHECK(dtrsm(s, kLeft, kLower, kNormal, kUnit, n, nrhs, 1.0, dA, lda, dB, ldb));
        ELX_CHECK(dtrsm(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, 1.0, dA, lda, dB, ldb));
    } else {
        ELX_CHECK(dtrsm(s, kLeft, kUpper, orient, kNonUnit, n, nrhs, 1.0, dA, lda, dB, ldb));
        ELX_CHECK(dtrsm(s, kLeft, kLower, orient, kUnit, n, nrhs, 1.0, dA, lda, dB, ldb));
        ELX_CHECK(permute_rows_impl(s, 1, dDests, n, n, nrhs, dB, ldb));
    }
    return ELX_SUCCESS;
}

elx_status_t elx_dgesv(void* stream, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                       double* dB, int64_t ldb, int64_t nb, int32_t* dDests, int32_t* dInfo) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG(nb >= 1 && nb <= 128 && dInfo != nullptr);
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
    ELX_CHECK(getrf_impl(s, n, n, dA, lda, nb, dests, dInfo, nrhs > 0 ? dB : nullptr, nrhs, ldb, true));
    if (nrhs > 0)
        ELX_CHECK(dtrsm(s, kLeft, kUpper, kNormal, kNonUnit, n, nrhs, 1.0, dA, lda, dB, ldb));
    return ELX_SUCCESS;
}

elx_status_t elx_dpotrf(void* stream, int uplo, int64_t n, double* dA, int64_t lda, int64_t nb,
                        int32_t* dInfo) {
    ELX_ARG(n >= 0 && lda >= std::max<int64_t>(n, 1) && nb >= 1 && nb <= 128 && dInfo);
    ELX_ARG(uplo == kLower || uplo == kUpper);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {
        ELX_HIP(hipMemsetAsync(dInfo, 0, sizeof(int32_t), s));
        return ELX_SUCCESS;
    }
    return potrf_impl(s, uplo, n, dA, lda, nb, dInfo);
}

elx_status_t elx_dpotrs(void* stream, int uplo, int orient, int64_t n, int64_t nrhs,
                        const double* dA, int64_t lda, double* dB, int64_t ldb) {
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) && ldb >= std::max<int64_t>(n, 1));
    ELX_ARG((uplo == kLower || uplo == kUpper) && orient >= 0 && orient <= 2);
    if (n == 0 || nrhs == 0) return ELX_SUCCESS;
    // TRANSPOSE conjugates B before and after: the identity for real data.
    return potrs_impl(static_cast<hipStream_t>(stream), uplo, n, nrhs, dA, lda, dB, ldb);
}

elx_status_t elx_dposv(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, double* dA,
                       int64_t lda, double* dB, int64_t ldb, int64_t nb, int32_t* dInfo) {
    ELX_CHECK(elx_dpotrf(stream, uplo, n, dA, lda, nb, dInfo));
    return elx_dpotrs(stream, uplo, orient, n, nrhs, dA, lda, dB, ldb);
}

elx_status_t elx_dgemm(void* stream, int transA, int transB, int64_t m, int64_t n, int64_t k,
                       double alpha, const double* dA, int64_t lda, const double* dB,
                       int64_t ldb, double beta, double* dC, int64_t ldc) {
    ELX_ARG(m >= 0 && n >= 0 && k >= 0);
    ELX_ARG(ldc >= std::max<int64_t>(m, 1));
    ELX_ARG(lda >= std::max<int64_t>(transA == kNormal ? m : k, 1));
    ELX_ARG(ldb >= std::max<int64_t>(transB == kNormal ? k : n, 1));
    return launch_dgemm(static_cast<hipStream_t>(stream), transA != kNormal, transB != kNormal, m,
                        n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, 0);
}

elx_status_t elx_dtrsm(void* stream, int side, int uplo, int orient, int diag, int64_t m,
                       int64_t n, double alpha, const double* dA, int64_t lda, double* dB,
                       int64_t ldb) {
    ELX_ARG(m >= 0 && n >= 0 && ldb >= std::max<int64_t>(m, 1));
    ELX_ARG(lda >= std::max<int64_t>(side == kLeft ? m : n, 1));
    ELX_ARG((side == kLeft || side == kRight) && (uplo == kLower || uplo == kUpper));
    ELX_ARG(orient >= 0 && orient <= 2 && (diag == kUnit || diag == kNonUnit));
    return dtrsm(static_cast<hipStream_t>(stream), side, uplo, orient, diag, m, n, alpha, dA, lda,
                 dB, ldb);
}

elx_status_t elx_dsyrk(void* stream, int uplo, int orient, int64_t n, int64_t k, double alpha,
                       const double* dA, int64_t lda, double beta, double* dC, int64_t ldc) {
    ELX_ARG(n >= 0 && k >= 0 && ldc >= std::max<int64_t>(n, 1));
    ELX_ARG(lda >= std::max<int64_t>(orient == kNormal ? n : k, 1));
    ELX_ARG(uplo == kLower || uplo == kUpper);
    const bool t = orient != kNormal;
    // NORMAL: C := alpha A A^T + beta C ; otherwise C := alpha A^T A + beta C
    return launch_dgemm(static_cast<hipStream_t>(stream), t, !t, n, n, k, alpha, dA, lda, dA, lda,
                        beta, dC, ldc, uplo == kLower ? 1 : 2);
}

elx_status_t elx_dpermute_rows(void* stream, int inverse, const int32_t* dDests, int64_t nswaps,
                               int64_t m, int64_t n, double* dA, int64_t lda) {
    ELX_ARG(m >= 0 && n >= 0 && nswaps >= 0 && nswaps <= m && lda >= std::max<int64_t>(m, 1));
    return permute_rows_impl(static_cast<hipStream_t>(stream), inverse, dDests, nswaps, m, n, dA,
                             lda);
}

}  // extern "C"
