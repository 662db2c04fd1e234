// Shared helpers of the HIP library: status macros, device properties, scratch pools.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/elemental_amd.h"

typedef int64_t i64;

#define ELX_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "elemental_amd: %s failed: %s (%s:%d)\n", #call,          \
                         hipGetErrorString(e_), __FILE__, __LINE__);                      \
            return ELX_ERROR_HIP;                                                         \
        }                                                                                 \
    } while (0)

#define ELX_CHECK(status)                         \
    do {                                          \
        elx_status_t s_ = (status);               \
        if (s_ != ELX_SUCCESS) return s_;         \
    } while (0)

#define ELX_ARG(cond)                                                                  \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::fprintf(stderr, "elemental_amd: invalid argument: %s (%s:%d)\n", #cond, \
                         __FILE__, __LINE__);                                          \
            return ELX_ERROR_INVALID_ARG;                                              \
        }                                                                              \
    } while (0)

namespace elx {

// Enum values of Elemental's C interface (include/El/core/types.h:185-225).
enum { kNormal = 0, kTranspose = 1, kAdjoint = 2 };
enum { kLower = 0, kUpper = 1 };
enum { kLeft = 0, kRight = 1 };
enum { kNonUnit = 0, kUnit = 1 };

struct DeviceInfo {
    int device = -1;
    int num_cus = 0;
    int max_lds_per_block = 0;
    int coop = 0;
};

// Properties of the current device (cached per device).
const DeviceInfo& device_info();

// Per-device scratch buffer that grows on demand and is kept for reuse. Growth
// synchronizes the device (the old buffer may still be in use by queued work).
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};
elx_status_t scratch_get(int slot, size_t bytes, void** out);
// scratch slot of the Ruiz equilibration's work vectors
constexpr int kSlotRuizWork = 64;
// scratch slot of the full-pivoting LU's per-workgroup pivot candidates
constexpr int kSlotFullPiv = 65;

// Per-device persistent state of the cooperative panel kernel's exchange area.
struct PanelSync {
    int64_t* flags = nullptr;  // one per workgroup, monotonically increasing
    void* slots = nullptr;     // candidate exchange area
    size_t slot_bytes = 0;
    int64_t epoch = 0;         // host-side launch counter (flags compare against it)
    int max_groups = 0;
};
elx_status_t panel_sync_get(int max_groups, size_t slot_bytes, PanelSync** out);
// The exchange area of the one-XCD panel: ordinary device memory (it stays in one XCD's L2).
elx_status_t xcd_sync_get(int max_groups, size_t slot_bytes, PanelSync** out);

inline unsigned cdiv(i64 a, i64 b) { return (unsigned)((a + b - 1) / b); }

// Optional per-kernel-class timing with HIP events on the launching stream (off by default;
// see elx_profile_enable). `work` is the algorithmic flops or bytes of the launch.
enum KernelClass {
    kKPanel = 0, kKDgemm, kKLpgemm, kKTrsm, kKRowMoves, kKTrsv, kKResidual, kKPack, kKBatched,
    kKBatchedSolve, kKOther,
    kKCount
};
bool prof_on();
void prof_begin(hipStream_t s, int cls);
void prof_end(hipStream_t s, int cls, double work);
struct ProfScope {
    hipStream_t s;
    int cls;
    double work;
    bool on;
    ProfScope(hipStream_t st, int c, double w) : s(st), cls(c), work(w), on(prof_on()) {
        if (on) prof_begin(s, cls);
    }
    ~ProfScope() {
        if (on) prof_end(s, cls, work);
    }
};

}  // namespace elx

// Kernel launchers shared between translation units (all enqueue on `stream`).
namespace elx {

// C := alpha op(A) op(B) + beta C  (FP64, MFMA). tri: 0 full, 1 lower only, 2 upper only.
// Checks, before a cooperative launch, that all `grid` workgroups of `fn` (block threads,
// `lds` bytes of dynamic LDS) can be resident at once; querying the kernel's attributes also
// loads its code object before the first launch.
elx_status_t coop_fits(const void* fn, int block, size_t lds, i64 grid);

// Launches a kernel whose workgroups wait on each other (every wait bounded): a cooperative
// launch, so that all workgroups are resident together. When the runtime cannot give this
// process a cooperative launch (hipErrorIllegalState, seen while another process on the same
// GPU holds the cooperative resources) the kernel is launched normally: with the grid no
// larger than what fits at once (checked by coop_fits) and an otherwise idle GPU, all
// workgroups are still resident, and if they are not, the bounded waits end the kernel with
// an error status instead of a hang.
elx_status_t launch_resident(const void* fn, i64 grid, int block, void** args, size_t lds,
                             hipStream_t s);

elx_status_t launch_dgemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                          const double* A, i64 lda, const double* B, i64 ldb, double beta,
                          double* C, i64 ldc, int tri);
// The same for FP64 and FP32 with the info word of a factorization sequence (`stop`, may be
// null): once it is non-zero the call leaves C unchanged. allow_library false keeps every
// product on the hand-written kernel (no library handle, no library workspace).
elx_status_t launch_gemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                         const double* A, i64 lda, const double* B, i64 ldb, double beta,
                         double* C, i64 ldc, int tri, const int32_t* stop,
                         bool allow_library = true);
elx_status_t launch_gemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, float alpha,
                         const float* A, i64 lda, const float* B, i64 ldb, float beta, float* C,
                         i64 ldc, int tri, const int32_t* stop, bool allow_library = true);
// Two device doubles per device for the coefficients of gated library calls.
elx_status_t gate_slot(double** out);

// Small triangular solves for many independent vectors (triangle order <= 128).
// side/uplo/orient/diag as in Elemental. B is m x n. `stop`: the info word of a
// factorization sequence (may be null); once it is non-zero the call changes nothing.
elx_status_t launch_trsm_small(hipStream_t s, int side, int uplo, int orient, int diag, i64 m,
                               i64 n, const double* A, i64 lda, double* B, i64 ldb,
                               const int32_t* stop = nullptr);
elx_status_t launch_trsm_small(hipStream_t s, int side, int uplo, int orient, int diag, i64 m,
                               i64 n, const float* A, i64 lda, float* B, i64 ldb,
                               const int32_t* stop = nullptr);
// maxbits (may be null): the largest |entry| of the result is folded into *maxbits (bits of a
// non-negative float, atomic max), for the exponent of the low-precision update that follows.
elx_status_t launch_trsm_small_f32(hipStream_t s, int side, int uplo, int orient, int diag,
                                   i64 m, i64 n, const float* A, i64 lda, float* B, i64 ldb,
                                   unsigned* maxbits = nullptr);

// General triangular solve (blocked by 128 with GEMM updates for larger triangles).
elx_status_t dtrsm(hipStream_t s, int side, int uplo, int orient, int diag, i64 m, i64 n,
                   double alpha, const double* A, i64 lda, double* B, i64 ldb);
elx_status_t strsm(hipStream_t s, int side, int uplo, int orient, int diag, i64 m, i64 n,
                   float alpha, const float* A, i64 lda, float* B, i64 ldb);

// LU panel of height mp, width w (<= 128) at A (column-major, lda). Records panel-relative
// pivots (plus offset `goff`) into dests, the row-move list into pairs (2*w int2), and the
// first zero pivot (global column + 1) into info.
elx_status_t launch_lu_panel_f64(hipStream_t s, i64 mp, i64 w, double* A, i64 lda,
                                 int32_t* dests, i64 goff, int2* pairs, int32_t* info);

elx_status_t launch_lu_panel(hipStream_t s, i64 mp, i64 w, double* A, i64 lda, int32_t* dests,
                             i64 goff, int2* pairs, int32_t* info);
elx_status_t launch_lu_panel(hipStream_t s, i64 mp, i64 w, float* A, i64 lda, int32_t* dests,
                             i64 goff, int2* pairs, int32_t* info);

// The same panel for an fp32 matrix stored row by row: element (r, j) at A[r*lda + j]. With
// xcd, the panel runs on the CUs of one XCD and exchanges its pivot candidates through that
// XCD's L2 (w <= 32 and at most 1024 rows per CU: lu_panel_xcd_fits).
elx_status_t launch_lu_panel_f32_rowmajor(hipStream_t s, i64 mp, i64 w, float* A, i64 lda,
                                          int32_t* dests, i64 goff, int2* pairs, int32_t* info,
                                          bool xcd = false);
bool lu_panel_xcd_fits(i64 mp, i64 w);
bool xcd_panels_enabled();
// Tournament-pivoting panel (communication-avoiding LU) of the mixed-precision factorization,
// same storage and outputs as launch_lu_panel_f32_rowmajor, factored ws (16 or 32) columns at
// a time; fits when the rows spread over at most one workgroup per CU, 128 rows each.
bool calu_panel_fits(i64 mp, i64 w, int ws);
elx_status_t launch_calu_panel_f32_rowmajor(hipStream_t s, i64 mp, i64 w, float* A, i64 lda,
                                            int32_t* dests, i64 goff, int2* pairs, int32_t* info,
                                            int ws);
// Panel of the mixed-precision LU: 0 partial pivoting, 16 or 32 the tournament block width
// (environment ELX_MIXED_PANEL = partial | calu16 | calu32, read at every solve; calu16 when
// unset), used for systems of order at least mixed_panel_min_order() (partial pivoting below).
int mixed_panel_mode();
i64 mixed_panel_min_order();
i64 mixed_panel_max_rhs();  // and for at most this many right-hand sides
// The row-move list applied to an fp32 matrix stored row by row (row i contiguous at
// A + i*lda): columns [c0,c1) and [c2,c3) of each moved row, starting at row r0; moves that
// reach outside the `rows` rows from r0 on are skipped.
elx_status_t launch_row_moves_f32_rowmajor(hipStream_t s, const int2* pairs, int npairs,
                                           float* A, i64 lda, i64 r0, i64 rows, i64 c0, i64 c1,
                                           i64 c2, i64 c3);

// Apply a row-move list (pairs[i] = {dst, src}, panel-relative, -1 = none) to columns
// [c0,c1) and [c2,c3) of A starting at row `r0`; moves outside rows [r0, r0+rows) are skipped.
elx_status_t launch_row_moves_f64(hipStream_t s, const int2* pairs, int npairs, double* A,
                                  i64 lda, i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3);
elx_status_t launch_row_moves(hipStream_t s, const int2* pairs, int npairs, double* A, i64 lda,
                              i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3,
                              const int32_t* stop);
elx_status_t launch_row_moves(hipStream_t s, const int2* pairs, int npairs, float* A, i64 lda,
                              i64 r0, i64 rows, i64 c0, i64 c1, i64 c2, i64 c3,
                              const int32_t* stop);

// Explicit permutation vector from a swap sequence (p[i] = source row of row i).
elx_status_t launch_swaps_to_perm(hipStream_t s, const int32_t* dests, i64 nswaps, i64 size,
                                  int32_t* perm, int inverse);

// B_out(i,:) = B_in(p[i],:)
elx_status_t launch_gather_rows_f64(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                    const double* Bin, i64 ldin, double* Bout, i64 ldout);
elx_status_t launch_gather_rows(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                const double* Bin, i64 ldin, double* Bout, i64 ldout);
elx_status_t launch_gather_rows(hipStream_t s, const int32_t* perm, i64 m, i64 n,
                                const float* Bin, i64 ldin, float* Bout, i64 ldout);

// C(fp32, column-major) += alpha * 2^-(ea+eb) * A * B^T with A (m x k) and B (n x k) fp16
// (prec 0) or bf16 (prec 1) stored K-major (ld a multiple of 8, >= k rounded up to 32,
// zero-padded); MFMA with fp32 accumulation. ea_dev / eb_dev: device ints, may be null.
elx_status_t launch_lpgemm_packed(hipStream_t s, int prec, i64 m, i64 n, i64 k, float alpha,
                                  const uint16_t* A, i64 lda, const uint16_t* B, i64 ldb,
                                  float* C, i64 ldc, const int* ea_dev, const int* eb_dev);
// The same product with B the FP32 operand itself (element (n, k) at B[n*ldb + k], readable up
// to the next multiple of 32 past k, 16-byte aligned rows): converted to fp16/bf16 while staged,
// no pack, no scaling of B.
elx_status_t launch_lpgemm_bf32(hipStream_t s, int prec, i64 m, i64 n, i64 k, float alpha,
                                const uint16_t* A, i64 lda, const float* B, i64 ldb, float* C,
                                i64 ldc, const int* ea_dev);
// H(r,k) = cvt(sign * 2^e * S(r,k)), S(r,k) = S[r + k*lds] (trans false) or S[k + r*lds];
// lo (fp16 only): the low part cvt((v - float(cvt(v))) * 2^11) of v = sign * 2^e * S(r,k).
// mx_in (may be null): e = scale_exp_of of the maximum folded into *mx_in by S's producer
// instead of *e_dev; the pack then stores e at e_out and clears *mx_clear.
elx_status_t launch_pack(hipStream_t s, int prec, bool trans, i64 R, i64 K, i64 Kpad,
                         const float* S, i64 lds, float sign, const int* e_dev, uint16_t* H,
                         i64 ldh, bool lo = false, const unsigned* mx_in = nullptr,
                         unsigned* mx_clear = nullptr, int* e_out = nullptr);
// C(m x n, column-major) += alpha * A B^T with A (m x k) and B (n x k) fp16 (prec 0) or
// bf16 (prec 1) packs (row r at A + r*lda), FP32 C, as a rocBLAS gemm_ex call; alpha_dev and
// beta_dev are device floats.
elx_status_t library_gemm_lowp(hipStream_t s, int prec, i64 m, i64 n, i64 k, const float* alpha_dev,
                               const uint16_t* A, i64 lda, const uint16_t* B, i64 ldb,
                               const float* beta_dev, float* C, i64 ldc);
// out[0] = -2^-(mul * e), out[1] = out[0] * 2^-11, out[2] = 1 on the device (e = 0 when
// e_dev is null): the alphas of the library GEMM for the packed, power-of-two scaled operands.
elx_status_t launch_gemm_alphas(hipStream_t s, const int* e_dev, int mul, float* out);
// FP64 C = alpha op(A) op(B) + beta C (rocblas_dgemm) and the symmetric rank-k update of one
// triangle of C (rocblas_dsyrk: trans false -> A A^T, true -> A^T A), host alpha and beta, or,
// when coef_dev is given, alpha = coef_dev[0] and beta = coef_dev[1] read on the device.
elx_status_t library_dgemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                           const double* A, i64 lda, const double* B, i64 ldb, double beta,
                           double* C, i64 ldc, const double* coef_dev = nullptr);
elx_status_t library_dsyrk(hipStream_t s, bool upper, bool trans, i64 n, i64 k, double alpha,
                           const double* A, i64 lda, double beta, double* C, i64 ldc,
                           const double* coef_dev = nullptr);
// FP32 B := op(A)^{-1} B from the left (rocblas_strsm), op = transpose when trans
elx_status_t library_strsm(hipStream_t s, bool upper, bool trans, bool unit, i64 m, i64 n,
                           const float* A, i64 lda, float* B, i64 ldb);
void library_gemm_release();
// Frees the batched solvers' schedules (pinned host and device copies) and queues.
void batched_release();
void batched_lu_release();
// The device-side batched mixed LU-IR (FP16 factors, AUTO refinement, FP64 LU fallback) for
// systems of order <= 2048; synchronous, stats per system.
elx_status_t gesv_mixed_batched_device(hipStream_t s, int64_t count, const int32_t* n,
                                       const int32_t* nrhs, const double* const* A_ptrs,
                                       const int32_t* lda, const double* const* B_ptrs,
                                       const int32_t* ldb, double* const* X_ptrs,
                                       const int32_t* ldx, double tol, elx_ir_stats_t* stats);
// e (device) with max|S| * 2^e < 2^15; ctl points at two zeroed unsigneds, zero again on return.
elx_status_t launch_scale_exponent(hipStream_t s, i64 m, i64 n, const float* S, i64 lds,
                                   unsigned* ctl, int* e_dev);

// Unblocked Cholesky of the diagonal block (n <= 128) in one workgroup.
elx_status_t launch_potf2_f64(hipStream_t s, int uplo, i64 n, double* A, i64 lda, i64 goff,
                              int32_t* info);
elx_status_t launch_potf2_f32(hipStream_t s, int uplo, i64 n, float* A, i64 lda, i64 goff,
                              int32_t* info);
elx_status_t launch_potf2(hipStream_t s, int uplo, i64 n, double* A, i64 lda, i64 goff,
                          int32_t* info);
elx_status_t launch_potf2(hipStream_t s, int uplo, i64 n, float* A, i64 lda, i64 goff,
                          int32_t* info);
// LU with full pivoting (lu::Full): p, q receive the row and column swap destinations.
elx_status_t getrf_full(hipStream_t s, i64 m, i64 n, double* A, i64 lda, int32_t* p, int32_t* q,
                        int32_t* info);
elx_status_t getrf_full(hipStream_t s, i64 m, i64 n, float* A, i64 lda, int32_t* p, int32_t* q,
                        int32_t* info);
// LU without pivoting of the n x n (n <= 128) block at A (lu::Unb); info as the panel's.
elx_status_t launch_getf2_nopiv(hipStream_t s, i64 n, double* A, i64 lda, i64 goff,
                                int32_t* info);
elx_status_t launch_getf2_nopiv(hipStream_t s, i64 n, float* A, i64 lda, i64 goff,
                                int32_t* info);

}  // namespace elx
