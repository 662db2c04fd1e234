// Launchers of the mixed-precision kernels (mixed_kernels.hip) used by the drivers in
// mixed.hip.
#pragma once
#include "elx_common.hpp"

namespace elx {

// W = float(A) stored row by row (A(i,j) at W[j + i*ldw], columns n..ldw-1 zeroed);
// rowabs[c][i] = sum over column chunk c of |A(i,j)| (nch chunks)
elx_status_t launch_convert(hipStream_t s, i64 n, const double* A, i64 lda, float* W, i64 ldw,
                            double* rowabs, int nch);
// *out = max_i sum_c rowabs[c][i]   (*out zeroed first)
elx_status_t launch_rowabs_max(hipStream_t s, i64 n, int nch, const double* rowabs, double* out);
// *out = max_i |v_i|   (*out zeroed first)
elx_status_t launch_vec_absmax(hipStream_t s, i64 n, const double* v, double* out);
// Inverses of the 128x128 diagonal blocks of the unit lower (upper = 0) or upper (upper = 1)
// factor stored row by row in W, to Inv (nblk x 128 x 128, column-major blocks).
// Blocks b0 .. b0+count-1 (count <= 0: through the last block).
elx_status_t launch_diag_inverse(hipStream_t s, i64 n, const float* W, i64 ldw, int upper,
                                 float* Inv, i64 b0, i64 count);
// S (n x n, ld n) = the full symmetric matrix stored in the lower (upper = 0) or upper
// triangle of A.
elx_status_t launch_symmetrize(hipStream_t s, i64 n, const double* A, i64 lda, int upper, double* S);
// Cholesky factor L in the lower triangle of W (column-major) -> W holds the LU pair
// L D^{-1} (strict lower) and D L^T (upper) stored row by row; d receives diag(L).
elx_status_t launch_chol_to_lu(hipStream_t s, i64 n, float* W, i64 ldw, float* d);
elx_status_t launch_iota(hipStream_t s, i64 n, int32_t* p);
// X := Inv * X for an nb-row block X (nb <= 128) of ncols columns stored row by row
// (X(k, j) at B[j + k*ldb]), in place.
// maxbits (may be null): the largest |entry| of the result folded into *maxbits (atomic max of
// the bits of a non-negative float).
elx_status_t launch_trmm_inv(hipStream_t s, i64 nb, i64 ncols, const float* Inv, float* B, i64 ldb,
                             unsigned* maxbits = nullptr);
// Persistent triangular solve with the factor stored row by row in W (fp32, ldw % 4 == 0,
// padding columns zero).
// xchunks: 64 x 16 bytes per 128-row block, uncached; epoch distinguishes this launch's
// chunks from every earlier one (it must differ from all epochs used before on the area).
elx_status_t launch_trsv(hipStream_t s, i64 n, const float* W, i64 ldw, const float* Inv,
                         int lower, const float* y, float* x, void* xchunks, uint32_t epoch,
                         int* err);
// y(i) = float(r(p[i]) * sr(p[i]))   (sr null: no scaling)
elx_status_t launch_gather_to_f32(hipStream_t s, i64 n, const int32_t* p, const double* r,
                                  const double* sr, float* y);
// Block forms for several right-hand sides, q a column list (null = identity):
// Y(i,k) = float(R(p[i],q[k]) * sr(p[i])); X(i,q[k]) = [X +] double(Y(i,k)) * sc(i);
// out[k] = max_i |M(i,k)|
elx_status_t launch_gather_rows_f32(hipStream_t s, i64 n, i64 nr, const int32_t* p,
                                   const int32_t* q, const double* R, i64 ldr, const double* sr,
                                   float* Y, i64 ldy);
elx_status_t launch_update_f64_2d(hipStream_t s, i64 n, i64 nr, const int32_t* q, const float* Y,
                                  i64 ldy, const double* sc, double* X, i64 ldx, int accumulate);
elx_status_t launch_col_absmax(hipStream_t s, i64 n, i64 nr, const double* M, i64 ldm, double* out);
// x(i) = [x(i) +] double(d(i)) * sc(i)   (sc null: no scaling)
elx_status_t launch_update_f64(hipStream_t s, i64 n, const float* d, const double* sc, double* x,
                               int accumulate);
// Ruiz equilibration (El::RuizEquil) of the m x n column-major A in place; dRow, dCol get
// the accumulated scales; work holds ruiz_work_doubles(m, n) doubles.
elx_status_t launch_ruiz_equil(hipStream_t s, i64 m, i64 n, double* A, i64 lda, double* dRow,
                               double* dCol, double* work);
size_t ruiz_work_doubles(i64 m, i64 n);
// El::GeomEquil of the m x n A in place; work holds geom_work_doubles(m, n) doubles
elx_status_t launch_geom_equil(hipStream_t s, i64 m, i64 n, double* A, i64 lda, double* dRow,
                               double* dCol, double* work);
size_t geom_work_doubles(i64 m, i64 n);
// El::SymmetricRuizEquil of the n x n A in place (maxiter sweeps); work holds 2n doubles
elx_status_t launch_symmetric_ruiz_equil(hipStream_t s, i64 n, double* A, i64 lda, double* d,
                                         int maxiter, double* work);
// out(i) = 1 / in(i)
elx_status_t launch_reciprocal(hipStream_t s, i64 n, const double* in, double* out);
// r = b - A x (or r = A x when b is null) with chunk partials in part (nch x n); max |r|,
// |x|, |b| into norms[0..2] (norms zeroed first). Requires ceil(n/nch) <= 1024.
elx_status_t launch_residual(hipStream_t s, i64 n, const double* A, i64 lda, const double* x,
                             const double* b, double* r, double* part, int nch, double* norms);

// h[c] = V(:,c)^T w for c < k (k <= 64); part needs ceil(n/1024)*k doubles
elx_status_t launch_multidot(hipStream_t s, i64 n, int k, const double* V, i64 ldv,
                             const double* w, double* part, double* h);
// w += alpha * V(:,0:k) h
elx_status_t launch_multiaxpy(hipStream_t s, i64 n, int k, double alpha, const double* V, i64 ldv,
                              const double* h, double* w);
// y += alpha x
elx_status_t launch_axpy(hipStream_t s, i64 n, double alpha, const double* x, double* y);
// y = alpha x
elx_status_t launch_scal_copy(hipStream_t s, i64 n, double alpha, const double* x, double* y);

}  // namespace elx
