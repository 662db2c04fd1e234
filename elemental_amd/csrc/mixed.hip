// Mixed-precision dense solve: low-precision LU (fp16 / bf16 trailing updates with fp32
// accumulation, fp32 panels and factor storage) recovered to FP64 backward error by
// iterative refinement with FP64 residuals.
//
// The refinement follows the semantics of El::RefinedSolve
// (include/El/lapack_like/solve/Refined.hpp:44-119): the residual is formed in the higher
// precision, the correction comes from the low-precision solve, and the iteration stops at
// the relative tolerance or when the residual stops decreasing. There is no dense
// mixed-precision solver in the reference; this path has no reference counterpart beyond
// those semantics.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "elx_common.hpp"
#include "mixed.hpp"

namespace elx {

namespace {

// Scratch slots (elx_api.hip uses 0..3)
enum {
    kSlotW = 10, kSlotLh, kSlotUh, kSlotLinv, kSlotUinv, kSlotIdx, kSlotVec, kSlotPart,
    kSlotSmall, kSlotXcopy, kSlotKrylovV, kSlotKrylovZ, kSlotKrylovW, kSlotKrylovP,
    kSlotF64, kSlotF64Idx, kSlotSym, kSlotDiag, kSlotLl, kSlotUl, kSlotBlockY, kSlotBlockR,
    kSlotBlockN, kSlotBlockQ, kSlotEquilA, kSlotEquilV, kSlotEquilW, kSlotXsave, kSlotBlockXs
};

// Uncached buffers of the persistent triangular solves (never freed before the process
// ends, see panel_sync_get).
struct TrsvArea {
    void* xchunks = nullptr;  // 64 x 16 bytes per 128-row block
    int* err = nullptr;
    float* xa = nullptr;
    float* xb = nullptr;
    i64 n = 0;
    int64_t epoch = 0;  // last epoch used; grows for the life of the process
};
std::mutex g_tmu;
std::map<int, TrsvArea> g_trsv;

elx_status_t trsv_area(i64 n, TrsvArea** out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tmu);
    TrsvArea& a = g_trsv[dev];
    if (a.n < n) {
        // (a smaller earlier area is kept: uncached memory a kernel has read is never freed)
        if (a.xchunks) ELX_HIP(hipDeviceSynchronize());
        const i64 nblk = (n + 127) / 128;
        void* xchunks = nullptr;
        float *xa = nullptr, *xb = nullptr;
        ELX_HIP(hipExtMallocWithFlags(&xchunks, (size_t)nblk * 64 * 16, hipDeviceMallocUncached));
        ELX_HIP(hipMemset(xchunks, 0, (size_t)nblk * 64 * 16));
        int* err = nullptr;
        ELX_HIP(hipExtMallocWithFlags(reinterpret_cast<void**>(&err), 64, hipDeviceMallocUncached));
        ELX_HIP(hipMemset(err, 0, 64));
        a.err = err;
        ELX_HIP(hipExtMallocWithFlags(reinterpret_cast<void**>(&xa), sizeof(float) * n,
                                      hipDeviceMallocUncached));
        ELX_HIP(hipExtMallocWithFlags(reinterpret_cast<void**>(&xb), sizeof(float) * n,
                                      hipDeviceMallocUncached));
        ELX_HIP(hipDeviceSynchronize());
        a.xchunks = xchunks;
        a.xa = xa;
        a.xb = xb;
        a.n = n;
    }
    *out = &a;
    return ELX_SUCCESS;
}

template <typename T>
elx_status_t get(int slot, size_t count, T** p) {
    void* v = nullptr;
    ELX_CHECK(scratch_get(slot, sizeof(T) * std::max<size_t>(count, 1), &v));
    *p = static_cast<T*>(v);
    return ELX_SUCCESS;
}

inline i64 round_up(i64 a, i64 b) { return ((a + b - 1) / b) * b; }

struct Ctx {
    hipStream_t s;
    int prec;  // 0 fp16, 1 bf16, 2 FP32-accurate (fp16 high and low parts, three products)
    i64 n, NB;
    float* W;
    i64 ldw;
    uint16_t* Lh;
    uint16_t* Uh;
    uint16_t* Ll;  // low parts (prec 2)
    uint16_t* Ul;
    float* Linv;
    float* Uinv;
    int32_t* dests;
    int32_t* perm;
    int2* pairs;  // (NB/32) x 256
    bool xcd_panels;  // one-XCD 32-column panels where they fit
    int panel_mode;   // 0 partial pivoting, 16 / 32 tournament pivoting in blocks of that width
    i64 calu_min_order;  // tournament panels only for systems of at least this order
    int32_t* info;
    int* e_dev;
    unsigned* bits;
    unsigned* ring;  // two words: the U block maxima folded in by the TRSM / TRMM producing them
    float* alphas;  // device alphas of the library GEMM (3 floats)
};

// C -= L * U for blocks of the fp32 working matrix, which holds A row by row (A(i, j) at
// W[j + i*ld]): L (m x k) at Lt with L(i, kk) at Lt[kk + i*ld], U (k x n) at Ut with U(kk, j) at
// Ut[j + kk*ld], C (m x n) at Ct with C(i, j) at Ct[j + i*ld]. Stored this way C is C^T in
// column-major order, so the update runs as C^T -= U^T L^T with the operands rounded to
// fp16/bf16 (U scaled by a power of two for fp16). FP32-accurate form: both operands split
// into fp16 high and low parts, C^T -= Uh^T Lh^T + 2^-11 (Uh^T Ll^T + Ul^T Lh^T), the small
// products first.
// The rank-NB updates on rocBLAS (gemm_ex) instead of the hand-written kernel: environment
// ELX_LOWP_GEMM=library, read at every update.
bool lowp_library_selected() {
    const char* v = std::getenv("ELX_LOWP_GEMM");
    return v && std::strcmp(v, "library") == 0;
}

// umax (may be null): one of c.ring's words, into which the kernel that produced U folded its
// largest magnitude; the U pack then takes the exponent from it (no reduction launch) and clears
// the other ring word, which the next producer folds into.
elx_status_t lp_update(Ctx& c, i64 m, i64 n, i64 k, const float* Lt, const float* Ut, float* Ct,
                       i64 ld, i64 ldpack, unsigned* umax = nullptr) {
    if (m <= 0 || n <= 0 || k <= 0) return ELX_SUCCESS;
    const i64 kpad = round_up(k, 32);
    const bool split = (c.prec == 2);
    const int hp = split ? 0 : c.prec;  // element type of the MFMA operands
    const bool l_direct0 = !split && k < 512;
    if (umax && hp == 0 && l_direct0) {
        unsigned* other = (umax == c.ring) ? c.ring + 1 : c.ring;
        ELX_CHECK(launch_pack(c.s, 0, false, n, k, kpad, Ut, ld, 1.0f, nullptr, c.Uh, ldpack,
                              false, umax, other, c.e_dev));
        return launch_lpgemm_bf32(c.s, 0, n, m, k, -1.0f, c.Uh, ldpack, Lt, ld, Ct, ld, c.e_dev);
    }
    const int* e = nullptr;
    if (hp == 0) {
        // c.bits: zeroed with the solve's control words, left zero by each exponent launch
        ELX_CHECK(launch_scale_exponent(c.s, n, k, Ut, ld, c.bits, c.e_dev));
        e = c.e_dev;
    }
    // L needs no pack for the hand-written product: it is K-major in W already, and the GEMM
    // converts it while staging
    const bool l_direct = !split && k < 512;
    if (!l_direct)
        ELX_CHECK(launch_pack(c.s, hp, true, m, k, kpad, Lt, ld, 1.0f, nullptr, c.Lh, ldpack));
    ELX_CHECK(launch_pack(c.s, hp, false, n, k, kpad, Ut, ld, 1.0f, e, c.Uh, ldpack));
    if (split) {
        ELX_CHECK(launch_pack(c.s, 0, true, m, k, kpad, Lt, ld, 1.0f, nullptr, c.Ll, ldpack, true));
        ELX_CHECK(launch_pack(c.s, 0, false, n, k, kpad, Ut, ld, 1.0f, e, c.Ul, ldpack, true));
    }
    if (k >= 512 && lowp_library_selected()) {
        // the rank-NB update as a plain library GEMM (rocBLAS), same packs and scaling
        ELX_CHECK(launch_gemm_alphas(c.s, e, 1, c.alphas));
        if (split) {
            ELX_CHECK(library_gemm_lowp(c.s, 0, n, m, kpad, c.alphas + 1, c.Uh, ldpack, c.Ll,
                                        ldpack, c.alphas + 2, Ct, ld));
            ELX_CHECK(library_gemm_lowp(c.s, 0, n, m, kpad, c.alphas + 1, c.Ul, ldpack, c.Lh,
                                        ldpack, c.alphas + 2, Ct, ld));
        }
        return library_gemm_lowp(c.s, hp, n, m, kpad, c.alphas, c.Uh, ldpack, c.Lh, ldpack,
                                 c.alphas + 2, Ct, ld);
    }
    if (split) {
        ELX_CHECK(launch_lpgemm_packed(c.s, 0, n, m, k, -1.0f / 2048.0f, c.Uh, ldpack, c.Ll,
                                       ldpack, Ct, ld, e, nullptr));
        ELX_CHECK(launch_lpgemm_packed(c.s, 0, n, m, k, -1.0f / 2048.0f, c.Ul, ldpack, c.Lh,
                                       ldpack, Ct, ld, e, nullptr));
    }
    if (l_direct) return launch_lpgemm_bf32(c.s, hp, n, m, k, -1.0f, c.Uh, ldpack, Lt, ld, Ct, ld, e);
    return launch_lpgemm_packed(c.s, hp, n, m, k, -1.0f, c.Uh, ldpack, c.Lh, ldpack, Ct, ld, e,
                                nullptr);
}

// Two-level blocked LU with partial pivoting of W (fp32, A stored row by row, so a row
// interchange moves contiguous memory): cooperative panels inside NB-column outer blocks; the
// outer trailing update is one rank-NB low-precision GEMM. Panels are 32 columns wide and run
// on the CUs of one XCD, exchanging their pivot candidates through its L2, when the rows fit
// there (up to 1024 per CU), else 128 columns over all XCDs through memory.
elx_status_t getrf_mixed(Ctx& c) {
    const i64 n = c.n, ld = c.ldw;
    float* W = c.W;
    // A(i, j) of the working matrix
    auto at = [&](i64 i, i64 j) { return W + j + i * ld; };
    ELX_HIP(hipMemsetAsync(c.info, 0, sizeof(int32_t), c.s));
    // the producers of the rank-128 updates' U blocks fold their maxima into c.ring[rp], the
    // pack of that update clears c.ring[rp ^ 1] for the next one (both start at zero; with bf16
    // or split operands no exponent is taken and the words stay zero)
    const bool fold = c.prec == 0;
    int rp = 0;
    if (fold) ELX_HIP(hipMemsetAsync(c.ring, 0, 2 * sizeof(unsigned), c.s));
    for (i64 K0 = 0; K0 < n; K0 += c.NB) {
        const i64 nbk = std::min(c.NB, n - K0);
        const i64 K1 = K0 + nbk;
        const bool xcd = c.xcd_panels && lu_panel_xcd_fits(n - K0, 32);
        const i64 P = xcd ? 32 : 128;
        for (i64 j0 = K0; j0 < K1; j0 += P) {
            const i64 w = std::min<i64>(P, K1 - j0);
            int2* pr = c.pairs + ((j0 - K0) / P) * 256;
            if (c.panel_mode != 0 && !xcd && n >= c.calu_min_order &&
                calu_panel_fits(n - j0, w, c.panel_mode))
                ELX_CHECK(launch_calu_panel_f32_rowmajor(c.s, n - j0, w, at(j0, j0), ld,
                                                         c.dests + j0, j0, pr, c.info,
                                                         c.panel_mode));
            else
                ELX_CHECK(launch_lu_panel_f32_rowmajor(c.s, n - j0, w, at(j0, j0), ld,
                                                       c.dests + j0, j0, pr, c.info, xcd));
            ELX_CHECK(launch_row_moves_f32_rowmajor(c.s, pr, (int)(2 * w), W, ld, j0, n - j0, K0,
                                                    j0, j0 + w, K1));
            const i64 wr = K1 - (j0 + w);
            if (wr > 0) {
                // U12 = L11^{-1} A12, as U12^T L11^T = A12^T on the stored transposes
                unsigned* um = fold ? c.ring + rp : nullptr;
                ELX_CHECK(launch_trsm_small_f32(c.s, kRight, kUpper, kNormal, kUnit, wr, w,
                                                at(j0, j0), ld, at(j0, j0 + w), ld, um));
                ELX_CHECK(lp_update(c, n - (j0 + w), wr, w, at(j0 + w, j0), at(j0, j0 + w),
                                    at(j0 + w, j0 + w), ld, 128, um));
                rp ^= fold ? 1 : 0;
            }
        }
        // the outer block's swaps on the columns outside it, in panel order
        for (i64 j0 = K0; j0 < K1; j0 += P) {
            const i64 w = std::min<i64>(P, K1 - j0);
            const int2* pr = c.pairs + ((j0 - K0) / P) * 256;
            ELX_CHECK(launch_row_moves_f32_rowmajor(c.s, pr, (int)(2 * w), W, ld, j0, n - j0, 0,
                                                    K0, K1, n));
        }
        // Inverses of the outer block's unit-lower diagonal blocks: final now (later swaps only
        // move rows below K1), used for U12 here and by the triangular solves afterwards.
        ELX_CHECK(launch_diag_inverse(c.s, n, W, ld, 0, c.Linv, K0 / 128, (nbk + 127) / 128));
        const i64 n2 = n - K1;
        if (n2 <= 0) continue;
        // U12 = L11^{-1} A12, blocked by 128, each block by its diagonal inverse
        for (i64 jb = K0; jb < K1; jb += 128) {
            const i64 wb = std::min<i64>(128, K1 - jb);
            const i64 below = K1 - (jb + wb);
            unsigned* um = (fold && below > 0) ? c.ring + rp : nullptr;
            ELX_CHECK(launch_trmm_inv(c.s, wb, n2, c.Linv + (jb / 128) * 128 * 128, at(jb, K1),
                                      ld, um));
            if (below > 0) {
                ELX_CHECK(lp_update(c, below, n2, wb, at(jb + wb, jb), at(jb, K1),
                                    at(jb + wb, K1), ld, 128, um));
                rp ^= fold ? 1 : 0;
            }
        }
        // A22 -= L21 U12 (rank nbk)
        ELX_CHECK(lp_update(c, n - K1, n2, nbk, at(K1, K0), at(K0, K1), at(K1, K1), ld,
                            round_up(c.NB, 32)));
    }
    return ELX_SUCCESS;
}

// Blocked lower Cholesky of W (fp32): 128-column diagonal blocks unblocked in FP32
// (cholesky::LowerVariant3Unblocked), the block column below by an FP32 triangular solve, the
// trailing matrix by one low-precision rank-128 GEMM (the packed L21 serves as both operands,
// both scaled by the same power of two). The full square is updated; only the lower triangle
// is read afterwards.
elx_status_t potrf_mixed(Ctx& c) {
    const i64 n = c.n, ld = c.ldw;
    float* W = c.W;
    ELX_HIP(hipMemsetAsync(c.info, 0, sizeof(int32_t), c.s));
    for (i64 k0 = 0; k0 < n; k0 += 128) {
        const i64 kb = std::min<i64>(128, n - k0);
        ELX_CHECK(launch_potf2_f32(c.s, kLower, kb, W + k0 + k0 * ld, ld, k0, c.info));
        const i64 n2 = n - k0 - kb;
        if (n2 <= 0) break;
        float* L21 = W + (k0 + kb) + k0 * ld;
        ELX_CHECK(launch_trsm_small_f32(c.s, kRight, kLower, kAdjoint, kNonUnit, n2, kb,
                                        W + k0 + k0 * ld, ld, L21, ld));
        const int* e = nullptr;
        const bool split = (c.prec == 2);
        const int hp = split ? 0 : c.prec;
        if (hp == 0) {
            ELX_CHECK(launch_scale_exponent(c.s, n2, kb, L21, ld, c.bits, c.e_dev));
            e = c.e_dev;
        }
        ELX_CHECK(launch_pack(c.s, hp, false, n2, kb, 128, L21, ld, 1.0f, e, c.Lh, 128));
        if (split) {
            float* C22 = W + (k0 + kb) + (k0 + kb) * ld;
            ELX_CHECK(launch_pack(c.s, 0, false, n2, kb, 128, L21, ld, 1.0f, e, c.Ll, 128, true));
            ELX_CHECK(launch_lpgemm_packed(c.s, 0, n2, n2, kb, -1.0f / 2048.0f, c.Lh, 128, c.Ll,
                                           128, C22, ld, e, e));
            ELX_CHECK(launch_lpgemm_packed(c.s, 0, n2, n2, kb, -1.0f / 2048.0f, c.Ll, 128, c.Lh,
                                           128, C22, ld, e, e));
        }
        ELX_CHECK(launch_lpgemm_packed(c.s, hp, n2, n2, kb, -1.0f, c.Lh, 128, c.Lh, 128,
                                       W + (k0 + kb) + (k0 + kb) * ld, ld, e, e));
    }
    return ELX_SUCCESS;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, s);
    }
    double stop() {
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms * 1e-3;
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

template <typename T>
elx_status_t d2h(T* dst, const T* src, size_t count, hipStream_t s) {
    ELX_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    ELX_HIP(hipStreamSynchronize(s));
    return ELX_SUCCESS;
}

}  // namespace

}  // namespace elx

using namespace elx;

extern "C" {

// The mixed-precision solve shared by the LU (kind 0) and Cholesky (kind 1) forms: both
// leave W as a unit-lower / upper pair with a row permutation, which the preconditioner
// applies; dA is the full FP64 matrix for the residuals.
static elx_status_t mixed_solve(int kind, void* stream, int64_t n, int64_t nrhs, const double* dA,
                                int64_t lda, const double* dB, int64_t ldb, double* dX,
                                int64_t ldx, int factor_flags, int refine, double tol,
                                int32_t max_iters, elx_ir_stats_t* stats) {
    const bool equil = (factor_flags & ELX_FACTOR_EQUILIBRATE) != 0;
    const int factor_type = factor_flags & ~ELX_FACTOR_EQUILIBRATE;
    ELX_ARG(n >= 0 && nrhs >= 0 && lda >= std::max<int64_t>(n, 1) &&
            ldb >= std::max<int64_t>(n, 1) && ldx >= std::max<int64_t>(n, 1));
    ELX_ARG(factor_type == ELX_FACTOR_FP16 || factor_type == ELX_FACTOR_BF16 ||
            factor_type == ELX_FACTOR_FP32);
    ELX_ARG(refine >= ELX_REFINE_NONE && refine <= ELX_REFINE_AUTO);
    elx_ir_stats_t st;
    std::memset(&st, 0, sizeof(st));
    if (n == 0 || nrhs == 0) {
        st.converged = 1;
        if (stats) *stats = st;
        return ELX_SUCCESS;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double u64 = std::ldexp(1.0, -53);
    if (!(tol > 0.0)) tol = (double)n * u64;
    if (max_iters <= 0) max_iters = 30;

    Ctx c;
    c.s = s;
    c.prec = (factor_type == ELX_FACTOR_FP16) ? 0 : (factor_type == ELX_FACTOR_BF16) ? 1 : 2;
    c.n = n;
    c.NB = std::min<i64>(1024, round_up(n, 128));
    c.ldw = round_up(n, 128);  // whole 128-row blocks and 16-byte row slices for the solves
    const i64 nblk = (n + 127) / 128;
    const i64 nbpad = round_up(c.NB, 32);
    const int nch = (int)std::max<i64>(1, (n + 1023) / 1024);
    ELX_CHECK(get(kSlotW, (size_t)c.ldw * (size_t)n, &c.W));
    ELX_CHECK(get(kSlotLh, (size_t)n * (size_t)nbpad, &c.Lh));
    ELX_CHECK(get(kSlotUh, (size_t)n * (size_t)nbpad, &c.Uh));
    c.Ll = c.Ul = nullptr;
    if (c.prec == 2) {
        ELX_CHECK(get(kSlotLl, (size_t)n * (size_t)nbpad, &c.Ll));
        ELX_CHECK(get(kSlotUl, (size_t)n * (size_t)nbpad, &c.Ul));
    }
    ELX_CHECK(get(kSlotLinv, (size_t)nblk * 128 * 128, &c.Linv));
    ELX_CHECK(get(kSlotUinv, (size_t)nblk * 128 * 128, &c.Uinv));
    int32_t* idx = nullptr;
    ELX_CHECK(get(kSlotIdx, (size_t)(2 * n + 16 + (c.NB / 32 + 1) * 512), &idx));
    c.xcd_panels = xcd_panels_enabled();
    c.panel_mode = mixed_panel_mode();
    c.calu_min_order = mixed_panel_min_order();
    // the refinement's cost grows with the right-hand sides and the tournament's weaker
    // factors need more steps: beyond a few columns partial pivoting pays
    if (nrhs > mixed_panel_max_rhs()) c.panel_mode = 0;
    c.dests = idx;
    c.perm = idx + n;
    c.info = idx + 2 * n;
    c.e_dev = idx + 2 * n + 1;
    c.bits = reinterpret_cast<unsigned*>(idx + 2 * n + 2);
    c.alphas = reinterpret_cast<float*>(idx + 2 * n + 4);
    c.ring = reinterpret_cast<unsigned*>(idx + 2 * n + 8);
    c.pairs = reinterpret_cast<int2*>(idx + 2 * n + 16);
    double* vec = nullptr;  // r, norms(4), anorm
    ELX_CHECK(get(kSlotVec, (size_t)(n + 8), &vec));
    double* r = vec;
    double* norms = vec + n;   // [0..2]
    double* anorm_d = vec + n + 4;
    double* part = nullptr;
    ELX_CHECK(get(kSlotPart, (size_t)nch * (size_t)n, &part));
    float* y = nullptr;
    ELX_CHECK(get(kSlotSmall, (size_t)n, &y));
    double* xs = nullptr;  // the iterate before an IR correction (accepted only if it helps)
    ELX_CHECK(get(kSlotXsave, (size_t)n, &xs));
    TrsvArea* ta = nullptr;
    ELX_CHECK(trsv_area(n, &ta));
    int* err = ta->err;
    ELX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));

    // ---- factorization ----
    Timer tf(s);
    ELX_HIP(hipMemsetAsync(idx + 2 * n, 0, sizeof(int32_t) * 16, s));
    ELX_CHECK(launch_convert(s, n, dA, lda, c.W, c.ldw, part, nch));
    ELX_HIP(hipMemsetAsync(anorm_d, 0, sizeof(double), s));
    ELX_CHECK(launch_rowabs_max(s, n, nch, part, anorm_d));
    // Equilibrated: W from diag(dRow)^-1 A diag(dCol)^-1 (an FP64 copy; symmetric, dRow = dCol,
    // for the Cholesky); the preconditioner scales by sr = 1/dRow before and sc = 1/dCol after
    // the triangular solves.
    double* sr = nullptr;
    double* sc = nullptr;
    if (equil) {
        double* Aeq = nullptr;
        double* ev = nullptr;
        double* ew = nullptr;
        ELX_CHECK(get(kSlotEquilA, (size_t)n * (size_t)n, &Aeq));
        ELX_CHECK(get(kSlotEquilV, (size_t)(4 * n), &ev));
        ELX_CHECK(get(kSlotEquilW, ruiz_work_doubles(n, n), &ew));
        ELX_HIP(hipMemcpy2DAsync(Aeq, sizeof(double) * n, dA, sizeof(double) * lda,
                                 sizeof(double) * n, n, hipMemcpyDeviceToDevice, s));
        sr = ev + 2 * n;
        if (kind == 0) {
            ELX_CHECK(launch_ruiz_equil(s, n, n, Aeq, n, ev, ev + n, ew));
            sc = ev + 3 * n;
            ELX_CHECK(launch_reciprocal(s, n, ev, sr));
            ELX_CHECK(launch_reciprocal(s, n, ev + n, sc));
        } else {
            // symmetric scaling keeps the Cholesky's input symmetric: A = D Aeq D
            ELX_CHECK(launch_symmetric_ruiz_equil(s, n, Aeq, n, ev, 3, ew));
            ELX_CHECK(launch_reciprocal(s, n, ev, sr));
            sc = sr;
        }
        ELX_CHECK(launch_convert(s, n, Aeq, n, c.W, c.ldw, part, nch));
    }
    if (kind == 0) {
        ELX_CHECK(getrf_mixed(c));
        // the explicit permutation (Permutation::ExplicitVector): the swaps replayed on the host,
        // which is quicker than one device thread doing n dependent steps
        std::vector<int32_t> dh((size_t)n), ph((size_t)n);
        ELX_CHECK(d2h(dh.data(), c.dests, (size_t)n, s));
        for (i64 i = 0; i < n; ++i) ph[(size_t)i] = (int32_t)i;
        for (i64 i = 0; i < n; ++i) {
            const int32_t d = dh[(size_t)i];
            if (d >= 0 && d < n && d != i) std::swap(ph[(size_t)i], ph[(size_t)d]);
        }
        ELX_HIP(hipMemcpyAsync(c.perm, ph.data(), sizeof(int32_t) * (size_t)n,
                               hipMemcpyHostToDevice, s));
        ELX_HIP(hipStreamSynchronize(s));
    } else {
        ELX_CHECK(potrf_mixed(c));
        float* dg = nullptr;
        ELX_CHECK(get(kSlotDiag, (size_t)n, &dg));
        ELX_CHECK(launch_chol_to_lu(s, n, c.W, c.ldw, dg));
        ELX_CHECK(launch_iota(s, n, c.perm));
        ELX_CHECK(launch_diag_inverse(s, n, c.W, c.ldw, 0, c.Linv, 0, 0));
    }
    ELX_CHECK(launch_diag_inverse(s, n, c.W, c.ldw, 1, c.Uinv, 0, 0));
    st.factor_seconds = tf.stop();
    int32_t finfo = 0;
    double anorm = 0.0;
    ELX_CHECK(d2h(&finfo, c.info, 1, s));
    ELX_CHECK(d2h(&anorm, anorm_d, 1, s));
    st.factor_info = finfo;
    // the cooperative panel gave up waiting for a peer workgroup: a device error, not a
    // numerical outcome
    if (finfo < 0) {
        if (stats) *stats = st;
        return ELX_ERROR_HIP;
    }

    // r = b - A x in FP64 and its normwise backward error
    auto backward = [&](const double* x, const double* b, double* eta,
                        double* rnorm = nullptr) -> elx_status_t {
        ELX_HIP(hipMemsetAsync(norms, 0, sizeof(double) * 3, s));
        ELX_CHECK(launch_residual(s, n, dA, lda, x, b, r, part, nch, norms));
        double hn[3];
        ELX_CHECK(d2h(hn, norms, 3, s));
        const double den = anorm * hn[1] + hn[2];
        *eta = den > 0 ? hn[0] / den : (hn[0] > 0 ? INFINITY : 0.0);
        if (rnorm) *rnorm = hn[0];
        return ELX_SUCCESS;
    };
    // FP64 fallback (refine = AUTO): El::LU of a copy of A, lu::SolveAfter, then refinement
    // with the FP64 factors. Taken when the low-precision factorization breaks down or the
    // refinement stalls.
    double* F64 = nullptr;
    int32_t* d64 = nullptr;
    bool f64_ready = false;
    bool f64_failed = false;  // the FP64 factorization met an exactly zero pivot: not retried
    auto fp64_solve = [&](const double* b, double* x, double* eta, bool* ok) -> elx_status_t {
        *ok = false;
        if (f64_failed) return ELX_SUCCESS;
        if (!f64_ready) {
            ELX_CHECK(get(kSlotF64, (size_t)n * (size_t)n, &F64));
            ELX_CHECK(get(kSlotF64Idx, (size_t)n + 8, &d64));
            ELX_HIP(hipMemcpy2DAsync(F64, sizeof(double) * n, dA, sizeof(double) * lda,
                                     sizeof(double) * n, n, hipMemcpyDeviceToDevice, s));
            ELX_HIP(hipMemsetAsync(d64 + n, 0, sizeof(int32_t), s));
            if (kind == 0) ELX_CHECK(elx_dgetrf(s, n, n, F64, n, 128, d64, d64 + n));
            else ELX_CHECK(elx_dpotrf(s, kLower, n, F64, n, 128, d64 + n));
            int32_t info64 = 0;
            ELX_CHECK(d2h(&info64, d64 + n, 1, s));
            if (info64 < 0) return ELX_ERROR_HIP;  // the device panel did not complete
            if (info64 != 0) {                     // exactly singular in FP64 as well
                f64_failed = true;
                return ELX_SUCCESS;
            }
            f64_ready = true;
        }
        auto f64_apply = [&](double* v) -> elx_status_t {
            if (kind == 0) return elx_dgetrs(s, 0, n, 1, F64, n, d64, v, n);
            return elx_dpotrs(s, kLower, 0, n, 1, F64, n, v, n);
        };
        ELX_HIP(hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        ELX_CHECK(f64_apply(x));
        for (int k = 0; k <= 5; ++k) {
            ELX_CHECK(backward(x, b, eta));
            if (*eta <= tol) {
                *ok = true;
                break;
            }
            if (k == 5) break;
            ELX_CHECK(f64_apply(r));
            ELX_CHECK(launch_axpy(s, n, 1.0, r, x));
        }
        return ELX_SUCCESS;
    };
    if (finfo != 0) {
        // an exactly zero pivot in the low-precision factors
        if (refine != ELX_REFINE_AUTO) {
            if (stats) *stats = st;
            return ELX_ERROR_NUMERIC;
        }
        Timer tr(s);
        bool all_ok = true;
        double worst = 0.0;
        for (i64 col = 0; col < nrhs; ++col) {
            double eta = INFINITY;
            bool ok = false;
            ELX_CHECK(fp64_solve(dB + col * ldb, dX + col * ldx, &eta, &ok));
            worst = std::max(worst, eta);
            all_ok = all_ok && ok;
        }
        st.fell_back = 1;
        st.refine_seconds = tr.stop();
        st.converged = all_ok ? 1 : 0;
        st.backward_error = worst;
        if (stats) *stats = st;
        return all_ok ? ELX_SUCCESS : ELX_ERROR_NUMERIC;
    }

    // ---- refinement, one right-hand side at a time ----
    // An IR step and a GMRES step cost the same (one preconditioner application, one FP64
    // product with A), so AUTO keeps IR while a step reduces the backward error at least 3x and
    // only then switches to GMRES-IR; plain IR stops once a step gains less than 2x.
    constexpr double kIrKeep = 1.0 / 3.0;
    const double ir_rate = (refine == ELX_REFINE_AUTO) ? kIrKeep : 0.5;
    const int m_restart = 50;
    double* V = nullptr;
    double* Z = nullptr;
    double* gw = nullptr;   // w, d, h(2 x 64)
    double* gpart = nullptr;
    if (refine == ELX_REFINE_GMRES || refine == ELX_REFINE_AUTO) {
        ELX_CHECK(get(kSlotKrylovV, (size_t)n * (m_restart + 1), &V));
        ELX_CHECK(get(kSlotKrylovZ, (size_t)n * m_restart, &Z));
        ELX_CHECK(get(kSlotKrylovW, (size_t)2 * n + 256, &gw));
        ELX_CHECK(get(kSlotKrylovP, (size_t)((n + 1023) / 1024) * 64 + 64, &gpart));
    }
    Timer tr(s);
    bool all_ok = true;
    double worst = 0.0;
    int inner_total = 0;
    bool escalate = false;
    // Several right-hand sides (refine NONE, IR, AUTO): block refinement first. A step is one
    // row gather, two FP32 triangular solves over the active columns (rocBLAS strsm on the
    // row-stored factors: W read column-major with leading dimension ldw is L^T in its strict
    // upper triangle and U^T in its lower one, so L y = P b is a transposed unit-upper solve and
    // U x = y a transposed lower one) and one FP64 residual product R = B - A X for the group.
    // A column leaves the block when its backward error reaches tol, or when a step reduces it
    // less than IR continues with (3x under AUTO, 2x for IR) or it runs out of steps: those
    // columns start again in the per-column loop below (GMRES, escalation, FP64 fallback).
    std::vector<char> done((size_t)nrhs, 0);
    if (nrhs > 1 && refine != ELX_REFINE_GMRES) {
        const i64 G = std::min<i64>(nrhs, 64);
        float* Yb = nullptr;
        double* Rb = nullptr;
        double* cn = nullptr;
        int32_t* qd = nullptr;
        ELX_CHECK(get(kSlotBlockY, (size_t)n * (size_t)G, &Yb));
        ELX_CHECK(get(kSlotBlockR, (size_t)n * (size_t)G, &Rb));
        ELX_CHECK(get(kSlotBlockN, (size_t)(3 * G), &cn));
        ELX_CHECK(get(kSlotBlockQ, (size_t)G, &qd));
        double* Xs = nullptr;  // the block's X before its last update
        ELX_CHECK(get(kSlotBlockXs, (size_t)n * (size_t)G, &Xs));
        // X(:, q) [+]= M^{-1} Src(:, q) for nq columns
        auto block_precond = [&](const double* Src, i64 lds, const int32_t* q, i64 nq, double* X0,
                                 int accumulate) -> elx_status_t {
            ELX_CHECK(launch_gather_rows_f32(s, n, nq, c.perm, q, Src, lds, sr, Yb, n));
            ELX_CHECK(library_strsm(s, true, true, true, n, nq, c.W, c.ldw, Yb, n));
            ELX_CHECK(library_strsm(s, false, true, false, n, nq, c.W, c.ldw, Yb, n));
            return launch_update_f64_2d(s, n, nq, q, Yb, n, sc, X0, ldx, accumulate);
        };
        std::vector<double> hn((size_t)(3 * G)), eta_prev((size_t)G), rn_prev((size_t)G);
        std::vector<char> active((size_t)G);
        std::vector<int32_t> q;
        for (i64 c0 = 0; c0 < nrhs; c0 += G) {
            const i64 g = std::min<i64>(G, nrhs - c0);
            const double* B0 = dB + c0 * ldb;
            double* X0 = dX + c0 * ldx;
            ELX_CHECK(block_precond(B0, ldb, nullptr, g, X0, 0));
            std::fill(active.begin(), active.end(), 1);
            for (int it = 0;; ++it) {
                ELX_HIP(hipMemcpy2DAsync(Rb, sizeof(double) * n, B0, sizeof(double) * ldb,
                                         sizeof(double) * n, g, hipMemcpyDeviceToDevice, s));
                ELX_CHECK(launch_dgemm(s, false, false, n, g, n, -1.0, dA, lda, X0, ldx, 1.0, Rb, n,
                                       0));
                ELX_CHECK(launch_col_absmax(s, n, g, Rb, n, cn));
                ELX_CHECK(launch_col_absmax(s, n, g, X0, ldx, cn + G));
                ELX_CHECK(launch_col_absmax(s, n, g, B0, ldb, cn + 2 * G));
                ELX_CHECK(d2h(hn.data(), cn, (size_t)(3 * G), s));
                q.clear();
                for (i64 k = 0; k < g; ++k) {
                    if (!active[(size_t)k]) continue;
                    const double den = anorm * hn[(size_t)(G + k)] + hn[(size_t)(2 * G + k)];
                    const double rn = hn[(size_t)k];
                    const double e = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
                    const bool conv = e <= tol;
                    if (conv || refine == ELX_REFINE_NONE) {
                        // NONE: one solve, reported as it stands
                        done[(size_t)(c0 + k)] = 1;
                        active[(size_t)k] = 0;
                        worst = std::max(worst, e);
                        all_ok = all_ok && conv;
                        st.outer_iters = std::max(st.outer_iters, it);
                        continue;
                    }
                    if (it > 0 && !(rn < rn_prev[(size_t)k])) {
                        // Refined.hpp:96-105: a candidate is kept only when the residual
                        // decreased; the column goes back to its previous iterate
                        ELX_HIP(hipMemcpyAsync(X0 + k * ldx, Xs + k * n, sizeof(double) * n,
                                               hipMemcpyDeviceToDevice, s));
                        active[(size_t)k] = 0;  // to the per-column loop
                        continue;
                    }
                    if (it >= max_iters || (it > 0 && !(e < ir_rate * eta_prev[(size_t)k]))) {
                        active[(size_t)k] = 0;  // to the per-column loop
                        continue;
                    }
                    eta_prev[(size_t)k] = e;
                    rn_prev[(size_t)k] = rn;
                    q.push_back((int32_t)k);
                }
                if (q.empty()) break;
                ELX_HIP(hipMemcpyAsync(qd, q.data(), sizeof(int32_t) * q.size(),
                                       hipMemcpyHostToDevice, s));
                ELX_HIP(hipMemcpy2DAsync(Xs, sizeof(double) * n, X0, sizeof(double) * ldx,
                                         sizeof(double) * n, g, hipMemcpyDeviceToDevice, s));
                ELX_CHECK(block_precond(Rb, n, qd, (i64)q.size(), X0, 1));
            }
        }
    }
    for (i64 col = 0; col < nrhs; ++col) {
        if (done[(size_t)col]) continue;
        const double* b = dB + col * ldb;
        double* x = dX + col * ldx;
        auto precond = [&](const double* rin, double* xout, int accumulate) -> elx_status_t {
            ELX_CHECK(launch_gather_to_f32(s, n, c.perm, rin, sr, y));
            const uint32_t t1 = (uint32_t)++ta->epoch;
            ELX_CHECK(launch_trsv(s, n, c.W, c.ldw, c.Linv, 1, y, ta->xa, ta->xchunks, t1, err));
            const uint32_t t2 = (uint32_t)++ta->epoch;
            ELX_CHECK(launch_trsv(s, n, c.W, c.ldw, c.Uinv, 0, ta->xa, ta->xb, ta->xchunks, t2,
                                  err));
            return launch_update_f64(s, n, ta->xb, sc, xout, accumulate);
        };
        // ||v||_2 on the device, read back
        auto nrm2 = [&](const double* v, double* out) -> elx_status_t {
            double* h = gw + 2 * n;
            ELX_CHECK(launch_multidot(s, n, 1, v, n, v, gpart, h));
            double hh = 0.0;
            ELX_CHECK(d2h(&hh, h, 1, s));
            *out = std::sqrt(std::max(hh, 0.0));
            return ELX_SUCCESS;
        };
        // Right-preconditioned flexible GMRES on A d = r (El::FGMRES semantics,
        // include/El/lapack_like/solve/FGMRES.hpp:37-291): z_j = M^{-1} v_j, w = A z_j,
        // orthogonalization (classical Gram-Schmidt applied twice), Givens rotations on H,
        // stop when the rotated residual |t_{j+1}| <= relTol * ||r||. d = Z y on exit.
        // *slow (AUTO with FP16/BF16 factors): the residual reduction of the first 10 steps
        // projects more than 80 steps for relTol, more than the FP32-accurate refactorization
        // costs (about 120 steps at n = 32768): the cycle stops and the caller escalates.
        auto fgmres = [&](const double* rr, double* d, double relTol, int* iters,
                          bool* slow) -> elx_status_t {
            *slow = false;
            double* w = gw;
            double* h = gw + 2 * n;
            double beta = 0.0;
            ELX_CHECK(nrm2(rr, &beta));
            *iters = 0;
            ELX_HIP(hipMemsetAsync(d, 0, sizeof(double) * n, s));
            if (!(beta > 0.0) || !std::isfinite(beta)) return ELX_SUCCESS;
            ELX_CHECK(launch_scal_copy(s, n, 1.0 / beta, rr, V));
            std::vector<double> H((size_t)(m_restart + 1) * m_restart, 0.0), cs(m_restart), sn(m_restart),
                t(m_restart + 1, 0.0);
            t[0] = beta;
            int jend = 0;
            for (int j = 0; j < m_restart; ++j) {
                ELX_CHECK(precond(V + (i64)j * n, Z + (i64)j * n, 0));
                ELX_HIP(hipMemsetAsync(norms, 0, sizeof(double) * 3, s));
                ELX_CHECK(launch_residual(s, n, dA, lda, Z + (i64)j * n, nullptr, w, part, nch, norms));
                std::vector<double> hv(j + 1, 0.0), h2(j + 1);
                for (int pass = 0; pass < 2; ++pass) {
                    ELX_CHECK(launch_multidot(s, n, j + 1, V, n, w, gpart, h));
                    ELX_CHECK(d2h(h2.data(), h, (size_t)(j + 1), s));
                    ELX_CHECK(launch_multiaxpy(s, n, j + 1, -1.0, V, n, h, w));
                    for (int i = 0; i <= j; ++i) hv[i] += h2[i];
                }
                double delta = 0.0;
                ELX_CHECK(nrm2(w, &delta));
                if (!std::isfinite(delta)) return ELX_ERROR_NUMERIC;
                if (j + 1 < m_restart && delta > 0.0)
                    ELX_CHECK(launch_scal_copy(s, n, 1.0 / delta, w, V + (i64)(j + 1) * n));
                double* Hj = H.data() + (size_t)j * (m_restart + 1);
                for (int i = 0; i <= j; ++i) Hj[i] = hv[i];
                Hj[j + 1] = delta;
                for (int i = 0; i < j; ++i) {
                    const double e0 = Hj[i], e1 = Hj[i + 1];
                    Hj[i] = cs[i] * e0 + sn[i] * e1;
                    Hj[i + 1] = -sn[i] * e0 + cs[i] * e1;
                }
                const double rho = std::hypot(Hj[j], Hj[j + 1]);
                if (!(rho > 0.0)) return ELX_ERROR_NUMERIC;
                cs[j] = Hj[j] / rho;
                sn[j] = Hj[j + 1] / rho;
                Hj[j] = rho;
                Hj[j + 1] = 0.0;
                const double t0 = t[j], t1 = t[j + 1];
                t[j] = cs[j] * t0 + sn[j] * t1;
                t[j + 1] = -sn[j] * t0 + cs[j] * t1;
                jend = j + 1;
                if (std::fabs(t[j + 1]) <= relTol * beta || delta == 0.0) break;
                // AUTO: a cycle that has not halved the residual in 20 steps is not worth
                // finishing; the caller sees the stall and escalates
                if (refine == ELX_REFINE_AUTO && jend >= 20 && std::fabs(t[j + 1]) > 0.5 * beta)
                    break;
                if (refine == ELX_REFINE_AUTO && c.prec != 2 && jend == 10) {
                    const double red = std::fabs(t[j + 1]) / beta;
                    const double need = (red > 0.0 && red < 1.0)
                                            ? 10.0 * std::log(relTol) / std::log(red)
                                            : INFINITY;
                    if (need > 80.0) {
                        *slow = true;
                        break;
                    }
                }
            }
            // y = H^{-1} t (upper triangular), d = Z y
            std::vector<double> yv(jend);
            for (int i = jend - 1; i >= 0; --i) {
                double v = t[i];
                for (int k = i + 1; k < jend; ++k) v -= H[(size_t)k * (m_restart + 1) + i] * yv[k];
                yv[i] = v / H[(size_t)i * (m_restart + 1) + i];
            }
            ELX_HIP(hipMemcpyAsync(h, yv.data(), sizeof(double) * jend, hipMemcpyHostToDevice, s));
            ELX_CHECK(launch_multiaxpy(s, n, jend, 1.0, Z, n, h, d));
            ELX_HIP(hipStreamSynchronize(s));
            *iters = jend;
            return ELX_SUCCESS;
        };
        // x0 = M^{-1} b
        ELX_CHECK(precond(b, x, 0));
        double eta = INFINITY, eta_prev = INFINITY, rn = INFINITY, rn_prev = INFINITY;
        int it = 0;
        bool ok = false;
        bool use_gmres = (refine == ELX_REFINE_GMRES);
        bool ir_step = false;  // the last update was an IR correction (x_prev saved in xs)
        for (;;) {
            ELX_CHECK(backward(x, b, &eta, &rn));
            if (ir_step && !(rn < rn_prev)) {
                // Refined.hpp:96-105: the candidate x + dx is kept only when the residual
                // decreased; otherwise the previous iterate stands
                ELX_HIP(hipMemcpyAsync(x, xs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                if (refine != ELX_REFINE_AUTO) {
                    eta = eta_prev;
                    break;
                }
                use_gmres = true;  // AUTO: continue from the previous iterate with GMRES-IR
                ELX_CHECK(backward(x, b, &eta, &rn));
                eta_prev = INFINITY;
            }
            ir_step = false;
            if (eta <= tol) {
                ok = true;
                break;
            }
            if (refine == ELX_REFINE_NONE || it >= max_iters) break;
            if (!use_gmres) {
                const bool slow = it > 0 && !(eta < kIrKeep * eta_prev);
                if (slow) {
                    if (refine == ELX_REFINE_AUTO) use_gmres = true;
                    else if (!(eta < 0.5 * eta_prev)) break;  // IR stagnated
                }
            } else if (it > 0 && !(eta < 0.5 * eta_prev)) {
                break;  // GMRES-IR stagnated
            }
            eta_prev = eta;
            rn_prev = rn;
            if (!use_gmres) {
                ELX_HIP(hipMemcpyAsync(xs, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                ELX_CHECK(precond(r, x, 1));
                ir_step = true;
            } else {
                // aim to reach tol in this outer step
                const double rel = std::min(1e-2, std::max(1e-13, 0.1 * tol / eta));
                int inner = 0;
                bool slow = false;
                double* d = gw + n;
                ELX_CHECK(fgmres(r, d, rel, &inner, &slow));
                inner_total += inner;
                if (slow) {
                    ++it;
                    break;  // escalate (not ok)
                }
                ELX_CHECK(launch_axpy(s, n, 1.0, d, x));
            }
            ++it;
        }
        st.outer_iters = std::max(st.outer_iters, it);
        if (!ok && refine == ELX_REFINE_AUTO) {
            if (c.prec != 2) {
                // too inaccurate a preconditioner for this matrix: refactor FP32-accurately
                escalate = true;
                break;
            }
            ELX_CHECK(fp64_solve(b, x, &eta, &ok));
            st.fell_back = 1;
        }
        worst = std::max(worst, eta);
        all_ok = all_ok && ok;
    }
    st.inner_iters = inner_total;
    st.refine_seconds = tr.stop();
    if (escalate) {
        elx_ir_stats_t s2;
        const elx_status_t r2 = mixed_solve(kind, stream, n, nrhs, dA, lda, dB, ldb, dX, ldx,
                                            ELX_FACTOR_FP32 | (factor_flags & ELX_FACTOR_EQUILIBRATE),
                                            refine, tol, max_iters, &s2);
        s2.factor_seconds += st.factor_seconds;
        s2.refine_seconds += st.refine_seconds;
        s2.inner_iters += st.inner_iters;
        s2.outer_iters += st.outer_iters;
        s2.fell_back = s2.fell_back ? 1 : 2;
        if (stats) *stats = s2;
        return r2;
    }
    int e_h = 0;
    ELX_CHECK(d2h(&e_h, err, 1, s));
    st.converged = all_ok ? 1 : 0;
    st.backward_error = worst;
    if (stats) *stats = st;
    if (e_h != 0) return ELX_ERROR_HIP;
    return all_ok ? ELX_SUCCESS : ELX_ERROR_NUMERIC;
}

elx_status_t elx_dgesv_mixed(void* stream, int64_t n, int64_t nrhs, const double* dA,
                             int64_t lda, const double* dB, int64_t ldb, double* dX,
                             int64_t ldx, int factor_type, int refine, double tol,
                             int32_t max_iters, elx_ir_stats_t* stats) {
    return mixed_solve(0, stream, n, nrhs, dA, lda, dB, ldb, dX, ldx, factor_type, refine, tol,
                       max_iters, stats);
}

elx_status_t elx_dgesv_mixed_batched(void* stream, int64_t count, const int32_t* n,
                                     const int32_t* nrhs, const double* const* A_ptrs,
                                     const int32_t* lda, const double* const* B_ptrs,
                                     const int32_t* ldb, double* const* X_ptrs,
                                     const int32_t* ldx, int factor_type, int refine, double tol,
                                     int32_t max_iters, elx_ir_stats_t* stats) {
    ELX_ARG(count >= 0);
    if (count == 0) return ELX_SUCCESS;
    ELX_ARG(n && nrhs && A_ptrs && lda && B_ptrs && ldb && X_ptrs && ldx);
    // FP16 factors with AUTO refinement and orders up to 2048: every system factored and
    // refined on the device, one workgroup each, FP64 LU fallback per system
    bool device_path = (factor_type == ELX_FACTOR_FP16 && refine == ELX_REFINE_AUTO &&
                        (max_iters <= 0 || max_iters == 30));
    for (int64_t i = 0; i < count && device_path; ++i) device_path = n[i] >= 0 && n[i] <= 2048;
    if (device_path)
        return gesv_mixed_batched_device(static_cast<hipStream_t>(stream), count, n, nrhs, A_ptrs,
                                         lda, B_ptrs, ldb, X_ptrs, ldx, tol, stats);
    // largest first: the workspaces grow once, to the largest order
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return n[a] > n[b]; });
    elx_status_t worst = ELX_SUCCESS;
    for (int64_t i : order) {
        elx_ir_stats_t st;
        std::memset(&st, 0, sizeof(st));
        const elx_status_t r = mixed_solve(0, stream, n[i], nrhs[i], A_ptrs[i], lda[i], B_ptrs[i],
                                           ldb[i], X_ptrs[i], ldx[i], factor_type, refine, tol,
                                           max_iters, &st);
        if (stats) stats[i] = st;
        if (r != ELX_SUCCESS && r != ELX_ERROR_NUMERIC) return r;
        if (r == ELX_ERROR_NUMERIC) worst = r;
    }
    return worst;
}

elx_status_t elx_dposv_mixed(void* stream, int uplo, int64_t n, int64_t nrhs, const double* dA,
                             int64_t lda, const double* dB, int64_t ldb, double* dX,
                             int64_t ldx, int factor_type, int refine, double tol,
                             int32_t max_iters, elx_ir_stats_t* stats) {
    ELX_ARG(uplo == kLower || uplo == kUpper);
    ELX_ARG(n >= 0 && lda >= std::max<int64_t>(n, 1));
    if (n == 0) return mixed_solve(1, stream, n, nrhs, dA, lda, dB, ldb, dX, ldx, factor_type,
                                   refine, tol, max_iters, stats);
    // the full symmetric matrix from the stored triangle (the residuals read all of it)
    double* S = nullptr;
    ELX_CHECK(get(kSlotSym, (size_t)n * (size_t)n, &S));
    ELX_CHECK(launch_symmetrize(static_cast<hipStream_t>(stream), n, dA, lda, uplo == kUpper, S));
    return mixed_solve(1, stream, n, nrhs, S, n, dB, ldb, dX, ldx, factor_type, refine, tol,
                       max_iters, stats);
}

}  // extern "C"
