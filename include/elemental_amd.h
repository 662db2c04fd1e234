/* elemental_amd.h -- C ABI of the MI355X-native dense factor-and-solve path.
 *
 * Every entry point takes DEVICE pointers to column-major data (element (i,j) at
 * A[i + j*lda], lda >= max(m,1)) and a HIP stream (hipStream_t passed as void*; NULL is
 * the default stream). Work is enqueued on that stream; calls return after enqueueing,
 * except where noted. Scratch memory is allocated by the library on first use per device
 * and kept until elx_finalize().
 *
 * Return value: ELX_SUCCESS or an elx_status_t error detected on the host (bad argument,
 * HIP error). Numerical failure is reported through the device word *dInfo, with the
 * LAPACK convention used by the reference's exceptions:
 *   *dInfo == 0      success
 *   *dInfo == j+1    column j (0-based, global) had an exactly zero pivot (LU) or a
 *                    non-positive diagonal (Cholesky) -- the point at which Elemental throws
 *                    SingularMatrixException / NonHPDMatrixException. The rest of the
 *                    enqueued sequence then changes nothing, so A (and B) are left as the
 *                    reference leaves them when it throws.
 *   *dInfo == -1     the cooperative LU panel gave up waiting for a peer workgroup (a GPU
 *                    shared with other work can delay one); A is unchanged from the start of
 *                    that panel on. A runtime error, not a numerical result.
 * Block sizes nb > 128 run as 128-column panels (the same pivots in exact arithmetic).
 *
 * Every routine of the plain path exists for FP64 (elx_d*) and FP32 (elx_s*), as the
 * reference instantiates its sequential path for both fields.
 * Pivots use Elemental's swap-sequence format: dests[j] is the 0-based row swapped with
 * row j at step j, i.e. Permutation::SwapDestinations() (LAPACK ipiv - 1).
 */
#ifndef ELEMENTAL_AMD_H
#define ELEMENTAL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ELX_SUCCESS = 0,
    ELX_ERROR_INVALID_ARG = 1,
    ELX_ERROR_HIP = 2,
    ELX_ERROR_ALLOC = 3,
    ELX_ERROR_NOT_SUPPORTED = 4,
    ELX_ERROR_NUMERIC = 5 /* returned only by the synchronous solvers below */
} elx_status_t;

/* Orientation, UpperOrLower, LeftOrRight and UnitOrNonUnit use Elemental's C enum values
 * (include/El/core/types.h:185-225): NORMAL=0, TRANSPOSE=1, ADJOINT=2; LOWER=0, UPPER=1;
 * LEFT=0, RIGHT=1; NON_UNIT=0, UNIT=1. */

/* Library information / lifetime. */
const char* elx_version(void);
int elx_device_count(void);
elx_status_t elx_finalize(void);

/* Kernel-class timing (no counterpart in the reference; instrumentation for measurement):
 * when enabled, every launch of the classes below is bracketed by HIP events on its own
 * stream, and elx_profile_query returns the number of launches, their summed device time
 * (ms) and summed algorithmic work (flops for GEMM classes, bytes otherwise). Classes:
 * 0 LU panel, 1 FP64 GEMM, 2 low-precision GEMM, 3 small TRSM, 4 row moves, 5 persistent
 * TRSV, 6 FP64 residual, 7 operand pack, 8 batched Cholesky factor, 9 batched Cholesky
 * solve. */
void elx_profile_enable(int on);
void elx_profile_reset(void);
elx_status_t elx_profile_query(int kernel_class, int64_t* launches, double* total_ms,
                               double* work);

/* ------------------------------------------------------------------------------------
 * Plain FP64 path (the drop-in)
 * ---------------------------------------------------------------------------------- */

/* PA = LU, blocked right-looking with block size nb (Elemental default 128).
 * Replaces El::LU(Matrix<double>&, Permutation&) -- src/lapack_like/factor/LU.cpp:104-142,
 * with lu::Panel -- src/lapack_like/factor/LU/Panel.hpp:15-60.
 * dDests: min(m,n) int32 (device). dInfo: one int32 (device). */
elx_status_t elx_dgetrf(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                        int64_t nb, int32_t* dDests, int32_t* dInfo);
elx_status_t elx_sgetrf(void* stream, int64_t m, int64_t n, float* dA, int64_t lda, int64_t nb,
                        int32_t* dDests, int32_t* dInfo);

/* A = LU without pivoting: blocked with lu::Unb on the diagonal blocks,
 * Trsm(RIGHT,UPPER,NORMAL,NON_UNIT) of A21, Trsm(LEFT,LOWER,NORMAL,UNIT) of A12, Gemm for A22.
 * Replaces El::LU(Matrix<F>&) -- src/lapack_like/factor/LU.cpp:21-44, LU/Local.hpp:44-61. */
elx_status_t elx_dgetrf_nopiv(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                              int64_t nb, int32_t* dInfo);
elx_status_t elx_sgetrf_nopiv(void* stream, int64_t m, int64_t n, float* dA, int64_t lda,
                              int64_t nb, int32_t* dInfo);

/* PAQ = LU with full (complete) pivoting: per step the first entry of largest magnitude of
 * the trailing block in column-major order, rows and columns swapped over the whole matrix.
 * dP, dQ: min(m,n) int32 row and column swap destinations (device).
 * Replaces El::LU(Matrix<F>&, Permutation& P, Permutation& Q) -- src/lapack_like/factor/
 * LU.cpp:144-152, LU/Full.hpp:15-58. */
elx_status_t elx_dgetrf_full(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                             int32_t* dP, int32_t* dQ, int32_t* dInfo);
elx_status_t elx_sgetrf_full(void* stream, int64_t m, int64_t n, float* dA, int64_t lda,
                             int32_t* dP, int32_t* dQ, int32_t* dInfo);

/* B := Q U^{-1} L^{-1} P B (NORMAL) or P^T L^{-T} U^{-T} Q^T B with the factors of
 * elx_dgetrf_full. Replaces lu::SolveAfter(Orientation, const Matrix<F>&, const Permutation&,
 * const Permutation&, Matrix<F>&) -- src/lapack_like/factor/LU/SolveAfter.hpp:123-151. */
elx_status_t elx_dgetrs_full(void* stream, int orient, int64_t n, int64_t nrhs, const double* dA,
                             int64_t lda, const int32_t* dP, const int32_t* dQ, double* dB,
                             int64_t ldb);
elx_status_t elx_sgetrs_full(void* stream, int orient, int64_t n, int64_t nrhs, const float* dA,
                             int64_t lda, const int32_t* dP, const int32_t* dQ, float* dB,
                             int64_t ldb);

/* B := op(A)^{-1} P B (orient NORMAL) or P^T op(A)^{-1} B (TRANSPOSE/ADJOINT) with the
 * factors of elx_dgetrf. Replaces lu::SolveAfter(Orientation, const Matrix<double>&,
 * const Permutation&, Matrix<double>&) -- src/lapack_like/factor/LU/SolveAfter.hpp:66-92.
 * dDests NULL: no permutation, the factors of elx_dgetrf_nopiv -- lu::SolveAfter(Orientation,
 * const Matrix<F>&, Matrix<F>&), src/lapack_like/factor/LU/SolveAfter.hpp:15-38. */
elx_status_t elx_dgetrs(void* stream, int orient, int64_t n, int64_t nrhs, const double* dA,
                        int64_t lda, const int32_t* dDests, double* dB, int64_t ldb);
elx_status_t elx_sgetrs(void* stream, int orient, int64_t n, int64_t nrhs, const float* dA,
                        int64_t lda, const int32_t* dDests, float* dB, int64_t ldb);

/* elx_dgetrs without any allocation, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same arguments and results as elx_dgetrs; the row
 * permutation's work space (4n bytes rounded up to 256, plus 8 n nrhs bytes; none when
 * dDests is NULL) comes from a fixed device array of ELX_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends, and every update runs on the library's own kernels. Larger
 * problems return ELX_ERROR_INVALID_ARG. Calls that overlap on one GPU share that array and
 * race: serialise them (one stream, or events). */
#define ELX_NOALLOC_SCRATCH_BYTES (32LL << 20)
elx_status_t elx_dgetrs_noalloc(void* stream, int orient, int64_t n, int64_t nrhs,
                                const double* dA, int64_t lda, const int32_t* dDests, double* dB,
                                int64_t ldb);

/* Short-circuited LU on [A|B] followed by one upper Trsm: B := A^{-1} B, A overwritten
 * with its (row-echelon) factors. Replaces lin_solve::Overwrite / RowEchelon --
 * src/lapack_like/solve/Linear.cpp:30-76,176-183 (El::LinearSolve, :212-218, is this after
 * a copy of A). dDests may be NULL. */
elx_status_t elx_dgesv(void* stream, int64_t n, int64_t nrhs, double* dA, int64_t lda,
                       double* dB, int64_t ldb, int64_t nb, int32_t* dDests, int32_t* dInfo);
elx_status_t elx_sgesv(void* stream, int64_t n, int64_t nrhs, float* dA, int64_t lda, float* dB,
                       int64_t ldb, int64_t nb, int32_t* dDests, int32_t* dInfo);

/* A = L L^T (uplo LOWER) or U^T U (UPPER), right-looking variant 3 with block size nb.
 * Replaces El::Cholesky(UpperOrLower, Matrix<double>&) -- src/lapack_like/factor/
 * Cholesky.cpp:26-38, Cholesky/LowerVariant3.hpp:15-68, Cholesky/UpperVariant3.hpp:15-73. */
elx_status_t elx_dpotrf(void* stream, int uplo, int64_t n, double* dA, int64_t lda, int64_t nb,
                        int32_t* dInfo);
elx_status_t elx_spotrf(void* stream, int uplo, int64_t n, float* dA, int64_t lda, int64_t nb,
                        int32_t* dInfo);

/* B := A^{-1} B with the Cholesky factor. Replaces cholesky::SolveAfter(UpperOrLower,
 * Orientation, const Matrix<double>&, Matrix<double>&) --
 * src/lapack_like/factor/Cholesky/SolveAfter.hpp:15-43. */
elx_status_t elx_dpotrs(void* stream, int uplo, int orient, int64_t n, int64_t nrhs,
                        const double* dA, int64_t lda, double* dB, int64_t ldb);
elx_status_t elx_spotrs(void* stream, int uplo, int orient, int64_t n, int64_t nrhs,
                        const float* dA, int64_t lda, float* dB, int64_t ldb);

/* Cholesky factor + solve in place (A overwritten). Replaces hpd_solve::Overwrite --
 * src/lapack_like/solve/HPD.cpp:15-25 (El::HPDSolve, :47-57, is this after a copy of A). */
elx_status_t elx_dposv(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, double* dA,
                       int64_t lda, double* dB, int64_t ldb, int64_t nb, int32_t* dInfo);
elx_status_t elx_sposv(void* stream, int uplo, int orient, int64_t n, int64_t nrhs, float* dA,
                       int64_t lda, float* dB, int64_t ldb, int64_t nb, int32_t* dInfo);

/* Level-3 pieces on device matrices. Replace El::Gemm / El::Trsm / El::Herk on
 * Matrix<double> -- src/blas_like/level3/Gemm.cpp:19-72, Trsm.cpp:24-64, Herk.cpp:14-21 ->
 * Syrk.cpp:20-55 (only the uplo triangle of C is written). */
elx_status_t elx_dgemm(void* stream, int transA, int transB, int64_t m, int64_t n, int64_t k,
                       double alpha, const double* dA, int64_t lda, const double* dB,
                       int64_t ldb, double beta, double* dC, int64_t ldc);
elx_status_t elx_dtrsm(void* stream, int side, int uplo, int orient, int diag, int64_t m,
                       int64_t n, double alpha, const double* dA, int64_t lda, double* dB,
                       int64_t ldb);
elx_status_t elx_dsyrk(void* stream, int uplo, int orient, int64_t n, int64_t k, double alpha,
                       const double* dA, int64_t lda, double beta, double* dC, int64_t ldc);
elx_status_t elx_sgemm(void* stream, int transA, int transB, int64_t m, int64_t n, int64_t k,
                       float alpha, const float* dA, int64_t lda, const float* dB, int64_t ldb,
                       float beta, float* dC, int64_t ldc);
elx_status_t elx_strsm(void* stream, int side, int uplo, int orient, int diag, int64_t m,
                       int64_t n, float alpha, const float* dA, int64_t lda, float* dB,
                       int64_t ldb);
elx_status_t elx_ssyrk(void* stream, int uplo, int orient, int64_t n, int64_t k, float alpha,
                       const float* dA, int64_t lda, float beta, float* dC, int64_t ldc);

/* Row permutation of a device matrix by a swap sequence (forward or inverse order).
 * Replaces Permutation::PermuteRows / InversePermuteRows on Matrix<double> --
 * src/lapack_like/perm/Permutation.cpp:544-631. */
elx_status_t elx_dpermute_rows(void* stream, int inverse, const int32_t* dDests,
                               int64_t nswaps, int64_t m, int64_t n, double* dA, int64_t lda);
elx_status_t elx_spermute_rows(void* stream, int inverse, const int32_t* dDests,
                               int64_t nswaps, int64_t m, int64_t n, float* dA, int64_t lda);

/* Ruiz equilibration in place: four sweeps scaling the columns, then the rows, by their
 * damped max norms; on return A holds diag(dRow)^-1 A0 diag(dCol)^-1 (bitwise the reference's
 * result). dRow (m) and dCol (n) are device vectors. Replaces El::RuizEquil(Matrix<double>& A,
 * Matrix<double>& dRow, Matrix<double>& dCol, bool) --
 * src/lapack_like/equilibrate/RuizEquil.cpp:23-60, include/El/lapack_like/equilibrate.hpp:17-21. */
elx_status_t elx_druiz_equil(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                             double* dRow, double* dCol);

/* Geometric equilibration in place: up to six sweeps scaling each column, then each row, by
 * max(sqrt(min nonzero |a| * max |a|), sqrt(1/1000) max |a|) (stopping after the fourth once
 * a sweep reduces max/min-nonzero by less than 10%), then every column to max |a| = 1; on
 * return A holds diag(dRow)^-1 A0 diag(dCol)^-1, bitwise the reference's result. Synchronous
 * (the stopping test reads two numbers per sweep). Replaces El::GeomEquil(Matrix<double>& A,
 * Matrix<double>& dRow, Matrix<double>& dCol, bool) --
 * src/lapack_like/equilibrate/GeomEquil.cpp:26-116, include/El/lapack_like/equilibrate.hpp. */
elx_status_t elx_dgeom_equil(void* stream, int64_t m, int64_t n, double* dA, int64_t lda,
                             double* dRow, double* dCol);

/* Symmetric Ruiz equilibration in place: maxiter sweeps (the reference's default is 3), each
 * scaling row and column j by the square root of column j's damped max norm; on return A
 * holds diag(d)^-1 A0 diag(d)^-1 (bitwise the reference's result) and dD (n) the scales.
 * Replaces El::SymmetricRuizEquil(Matrix<double>& A, Matrix<double>& d, Int maxiter, bool) --
 * src/lapack_like/equilibrate/SymmetricRuizEquil.cpp:29-55, include/El/lapack_like/
 * equilibrate.hpp:80-84. */
elx_status_t elx_dsymmetric_ruiz_equil(void* stream, int64_t n, double* dA, int64_t lda,
                                       double* dD, int32_t maxiter);

/* ------------------------------------------------------------------------------------
 * Mixed precision: low-precision LU + FP64 iterative refinement (no reference
 * counterpart for dense matrices; the refinement loop follows the semantics of
 * El::RefinedSolve, include/El/lapack_like/solve/Refined.hpp:44-119, and the Krylov
 * variant follows El::FGMRES, include/El/lapack_like/solve/FGMRES.hpp:37-291).
 * ---------------------------------------------------------------------------------- */
typedef enum {
    ELX_FACTOR_FP64 = 0,
    ELX_FACTOR_FP32 = 1,  /* FP32-accurate trailing updates: each operand split into fp16 high
                             and low parts, three fp16 MFMA products (hi*hi + hi*lo + lo*hi) */
    ELX_FACTOR_FP16 = 2,
    ELX_FACTOR_BF16 = 3,
    /* Flag, OR'ed into a factor type of the mixed solvers: the low-precision LU factors the
     * Ruiz-equilibrated diag(dRow)^-1 A diag(dCol)^-1 of an FP64 copy of A (elx_druiz_equil),
     * the Cholesky the symmetrically equilibrated diag(d)^-1 A diag(d)^-1
     * (elx_dsymmetric_ruiz_equil, 3 sweeps); the preconditioner applies the scales, and the
     * residuals and the FP64 fallback use A itself. For badly scaled A; costs an n x n FP64
     * copy and up to eight passes over it. */
    ELX_FACTOR_EQUILIBRATE = 0x100
} elx_factor_type_t;

typedef enum {
    ELX_REFINE_NONE = 0,
    ELX_REFINE_IR = 1,      /* classical iterative refinement */
    ELX_REFINE_GMRES = 2,   /* GMRES-IR: each correction by GMRES preconditioned with the LU */
    ELX_REFINE_AUTO = 3     /* IR while a step reduces the backward error at least 3x, then
                               GMRES-IR (an IR step and a GMRES step cost the same); when GMRES-IR
                               stalls, an FP16/BF16 factorization is redone FP32-accurately
                               (ELX_FACTOR_FP32) and refined again, and the FP64 LU
                               (elx_dgetrf of a copy of A) is the last resort; an exactly zero
                               pivot in the low-precision factors goes to the FP64 LU */
} elx_refine_t;

typedef struct {
    int32_t converged;        /* 1 when every column met the tolerance */
    int32_t outer_iters;      /* refinement steps (residual evaluations after the first) */
    int32_t inner_iters;      /* total GMRES iterations */
    int32_t fell_back;        /* 1: the FP64 LU fallback was used; 2: refactored FP32-accurately
                                 (ELX_REFINE_AUTO) and refined to tol */
    int32_t factor_info;      /* info of the low-precision factorization */
    int32_t near_tie_pivots;  /* unused by the mixed path (0) */
    double backward_error;    /* max over columns of ||b-Ax||inf/(||A||inf||x||inf+||b||inf) */
    double factor_seconds;    /* device time of the factorization */
    double refine_seconds;    /* device time of the refinement */
} elx_ir_stats_t;

/* Solve A X = B to FP64 backward error <= tol (tol <= 0 selects n * 2^-53) with a
 * factorization in factor_type and refinement in FP64. A (FP64, not modified) and B stay
 * on the device; X receives the solution. Synchronous: returns after the solve, with
 * *stats filled (may be NULL). Returns ELX_ERROR_NUMERIC if neither refinement nor the
 * FP64 fallback reached tol. */
elx_status_t elx_dgesv_mixed(void* stream, int64_t n, int64_t nrhs, const double* dA,
                             int64_t lda, const double* dB, int64_t ldb, double* dX,
                             int64_t ldx, int factor_type, int refine, double tol,
                             int32_t max_iters, elx_ir_stats_t* stats);

/* elx_dgesv_mixed for `count` independent systems: A_i X_i = B_i (n_i x n_i, n_i x nrhs_i).
 * The arrays are host arrays of `count` entries; A_ptrs / B_ptrs / X_ptrs hold device pointers.
 * Each system gets its own stats entry (stats may be NULL) and its own FP64 fallback under
 * ELX_REFINE_AUTO. The systems are solved one after another on the stream, largest first.
 * Returns ELX_ERROR_NUMERIC if any system missed the tolerance. No counterpart in the
 * reference (a caller would loop over El::LinearSolve). */
elx_status_t elx_dgesv_mixed_batched(void* stream, int64_t count, const int32_t* n,
                                     const int32_t* nrhs, const double* const* A_ptrs,
                                     const int32_t* lda, const double* const* B_ptrs,
                                     const int32_t* ldb, double* const* X_ptrs,
                                     const int32_t* ldx, int factor_type, int refine, double tol,
                                     int32_t max_iters, elx_ir_stats_t* stats);

/* Same for Hermitian positive definite A, whose `uplo` triangle is read (the other is never
 * touched; El::HPDSolve semantics, src/lapack_like/solve/HPD.cpp:47-57): Cholesky in FP32
 * diagonal blocks with FP16 / BF16 MFMA trailing updates (factor_type), refinement in FP64 as
 * above, FP64 Cholesky (elx_dpotrf / elx_dpotrs) as the AUTO fallback. factor_info > 0 reports
 * the first non-positive pivot of the low-precision factorization. */
elx_status_t elx_dposv_mixed(void* stream, int uplo, int64_t n, int64_t nrhs, const double* dA,
                             int64_t lda, const double* dB, int64_t ldb, double* dX,
                             int64_t ldx, int factor_type, int refine, double tol,
                             int32_t max_iters, elx_ir_stats_t* stats);

/* ------------------------------------------------------------------------------------
 * Batched variable-size Cholesky (lower). No counterpart in the reference, which factors one
 * Matrix at a time: each matrix i gets exactly what El::Cholesky(LOWER, A_i)
 * (src/lapack_like/factor/Cholesky.cpp:26-38, LowerVariant3.hpp:15-68) and
 * cholesky::SolveAfter(LOWER, NORMAL, L_i, B_i) (Cholesky/SolveAfter.hpp:15-43) compute, up to
 * rounding (32-column blocks instead of Blocksize()). n, nrhs, lda, ldb are host arrays of
 * `count` entries (0 <= n_i <= 512); dA_array / dB_array are device arrays of device
 * pointers to column-major matrices; dinfo (device, count entries) receives 0 or j+1 for the
 * first non-positive pivot of matrix i (NonHPDMatrixException in the reference). The
 * matrices are scheduled largest first over persistent workgroups. Asynchronous on stream.
 * ---------------------------------------------------------------------------------- */
elx_status_t elx_dpotrf_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 double* const* dA_array, const int32_t* lda, int32_t* dinfo);
elx_status_t elx_spotrf_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 float* const* dA_array, const int32_t* lda, int32_t* dinfo);
elx_status_t elx_dpotrs_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 const int32_t* nrhs, const double* const* dA_array,
                                 const int32_t* lda, double* const* dB_array, const int32_t* ldb);
elx_status_t elx_spotrs_vbatched(void* stream, int uplo, int64_t count, const int32_t* n,
                                 const int32_t* nrhs, const float* const* dA_array,
                                 const int32_t* lda, float* const* dB_array, const int32_t* ldb);

/* The batched factor + solve (elx_?potrf_vbatched, then elx_?potrs_vbatched on the same
 * matrices, skipping any matrix whose factorization failed: HPDSolve throws before it
 * solves) captured once into a HIP graph for this batch: sizes, leading dimensions and the
 * device pointer arrays are fixed at capture, the matrices' contents may change between
 * launches. elx_graph_launch replays it on a stream (on the device it was made on);
 * elx_graph_destroy waits for the device and frees it. No counterpart in the reference. */
typedef struct elx_graph_st* elx_graph_t;
elx_status_t elx_dposv_vbatched_graph(void* stream, int uplo, int64_t count, const int32_t* n,
                                      const int32_t* nrhs, double* const* dA_array,
                                      const int32_t* lda, double* const* dB_array,
                                      const int32_t* ldb, int32_t* dinfo, elx_graph_t* graph);
elx_status_t elx_sposv_vbatched_graph(void* stream, int uplo, int64_t count, const int32_t* n,
                                      const int32_t* nrhs, float* const* dA_array,
                                      const int32_t* lda, float* const* dB_array,
                                      const int32_t* ldb, int32_t* dinfo, elx_graph_t* graph);
elx_status_t elx_graph_launch(elx_graph_t graph, void* stream);
elx_status_t elx_graph_destroy(elx_graph_t graph);

/* Tuning hooks of the batched Cholesky (no counterpart in the reference; used by the
 * autotuner, elemental_amd/autotune_batched.py): the variant -- block-column width nb (16 or
 * 32; 64 for kind 2), persistent workgroups per CU occ, and kernel kind (0: block column
 * staged in LDS; 1: operands read straight into registers; 2: 64-column panels with the
 * diagonal block factored and inverted in LDS; kinds 1 and 2 in bins 1 and 2 only) -- of size
 * bin `bin` (0: n <= 128, 1: n <= 256, 2: n <= 512) for dtype 0 (double) or 1 (float).
 * nb = 0 restores the built-in table. */
elx_status_t elx_batched_set_variant(int dtype, int bin, int nb, int occ, int kind);
elx_status_t elx_batched_get_variant(int dtype, int bin, int* nb, int* occ, int* kind);

#ifdef __cplusplus
}
#endif

#endif /* ELEMENTAL_AMD_H */
