// el_oracle.cpp -- CPU restatement of Elemental's sequential dense factor-and-solve path.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle: it is loaded by tests/, by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, as the checker or as
// the timed CPU baseline. The product (elemental_amd/) never links or calls it.
//
// It issues the same BLAS call sequence, on the same blocks, as the reference:
//   El::LU(Matrix&,Permutation&)        src/lapack_like/factor/LU.cpp:104-142
//   lu::Panel                           src/lapack_like/factor/LU/Panel.hpp:15-60
//   lu::SolveAfter(orient,A,P,B)        src/lapack_like/factor/LU/SolveAfter.hpp:66-92
//   El::LinearSolve / RowEchelon        src/lapack_like/solve/Linear.cpp:30-76,176-183,212-218
//   El::Cholesky (variant 3, L and U)   src/lapack_like/factor/Cholesky.cpp:26-38,
//                                       .../Cholesky/LowerVariant3.hpp:15-68,
//                                       .../Cholesky/UpperVariant3.hpp:15-73
//   cholesky::SolveAfter                src/lapack_like/factor/Cholesky/SolveAfter.hpp:15-43
//   El::HPDSolve                        src/lapack_like/solve/HPD.cpp:15-25,47-57
//   Permutation (swap sequence)         src/lapack_like/perm/Permutation.cpp:166-250,544-631,731-756
//   El::LU(Matrix&) without pivoting    src/lapack_like/factor/LU.cpp:21-44, LU/Local.hpp:44-61
//   lu::SolveAfter(orient,A,B)          src/lapack_like/factor/LU/SolveAfter.hpp:15-38
//   El::LU(A,P,Q) (full pivoting)       src/lapack_like/factor/LU/Full.hpp:15-58, LU.cpp:144-152
//   lu::SolveAfter(orient,A,P,Q,B)      src/lapack_like/factor/LU/SolveAfter.hpp:123-151
//   Gaussian / GEPPGrowth               src/matrices/random/independent/Gaussian.cpp:16-47,
//                                       include/El/core/random/impl.hpp:185-207,
//                                       src/core/random.cpp:24-32,
//                                       src/matrices/deterministic/misc/GEPPGrowth.cpp:15-30
//
// Every algorithm is a template over the field and is exported for double (orc_*) and float
// (orc_s*), as the reference instantiates its sequential path for both.
//
// The BLAS underneath is selectable (double; float always runs the naive form):
//   * "naive": the netlib reference-BLAS loop orders, with every a += b*c written as one
//     fused multiply-add (the form vectorised host BLAS kernels use). Deterministic,
//     single-threaded; this is the backend the parity tests use.
//   * "mkl" / "openblas": a host BLAS resolved with dlopen (Fortran ABI, 32-bit ints),
//     used for the timed CPU baseline and for library cross-checks.
#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dd.h"

typedef int64_t i64;

namespace {

// ---------------------------------------------------------------------------------------
// BLAS backend table
// ---------------------------------------------------------------------------------------
struct Blas {
    i64 (*idamax)(i64 n, const double* x, i64 incx);  // 0-based, first index on ties
    void (*dswap)(i64 n, double* x, i64 incx, double* y, i64 incy);
    void (*dscal)(i64 n, double a, double* x, i64 incx);
    void (*dger)(i64 m, i64 n, double alpha, const double* x, i64 incx, const double* y,
                 i64 incy, double* A, i64 lda);
    void (*dsyr)(char uplo, i64 n, double alpha, const double* x, i64 incx, double* A,
                 i64 lda);
    void (*dtrsm)(char side, char uplo, char trans, char diag, i64 m, i64 n, double alpha,
                  const double* A, i64 lda, double* B, i64 ldb);
    void (*dgemm)(char ta, char tb, i64 m, i64 n, i64 k, double alpha, const double* A,
                  i64 lda, const double* B, i64 ldb, double beta, double* C, i64 ldc);
    void (*dsyrk)(char uplo, char trans, i64 n, i64 k, double alpha, const double* A,
                  i64 lda, double beta, double* C, i64 ldc);
};

#define IDX(i, j, ld) ((i) + (i64)(j) * (ld))

// ---------------- naive (netlib loop order, FMA), for T = double and float ----------------
template <typename T>
i64 nv_idamax(i64 n, const T* x, i64 incx) {
    // Strict '>' keeps the first index among equal magnitudes (BLAS i?amax semantics),
    // src/core/imports/blas/MaxInd.hpp:26-46.
    T best = T(-1);
    i64 bi = -1;
    for (i64 i = 0; i < n; ++i) {
        T v = std::fabs(x[i * incx]);
        if (v > best) {
            best = v;
            bi = i;
        }
    }
    return bi;
}

template <typename T>
void nv_dswap(i64 n, T* x, i64 incx, T* y, i64 incy) {
    for (i64 i = 0; i < n; ++i) std::swap(x[i * incx], y[i * incy]);
}

template <typename T>
void nv_dscal(i64 n, T a, T* x, i64 incx) {
    for (i64 i = 0; i < n; ++i) x[i * incx] = a * x[i * incx];
}

template <typename T>
void nv_dger(i64 m, i64 n, T alpha, const T* x, i64 incx, const T* y, i64 incy, T* A,
             i64 lda) {
    for (i64 j = 0; j < n; ++j) {
        T yj = y[j * incy];
        if (yj == T(0)) continue;
        T temp = alpha * yj;
        T* a = A + j * lda;
        for (i64 i = 0; i < m; ++i) a[i] = std::fma(x[i * incx], temp, a[i]);
    }
}

template <typename T>
void nv_dsyr(char uplo, i64 n, T alpha, const T* x, i64 incx, T* A, i64 lda) {
    for (i64 j = 0; j < n; ++j) {
        T xj = x[j * incx];
        if (xj == T(0)) continue;
        T temp = alpha * xj;
        T* a = A + j * lda;
        if (uplo == 'L') {
            for (i64 i = j; i < n; ++i) a[i] = std::fma(x[i * incx], temp, a[i]);
        } else {
            for (i64 i = 0; i <= j; ++i) a[i] = std::fma(x[i * incx], temp, a[i]);
        }
    }
}

template <typename T>
void nv_dgemm(char ta, char tb, i64 m, i64 n, i64 k, T alpha, const T* A, i64 lda, const T* B,
              i64 ldb, T beta, T* C, i64 ldc) {
    if (m == 0 || n == 0) return;
    const bool nota = (ta == 'N'), notb = (tb == 'N');
    for (i64 j = 0; j < n; ++j) {
        T* c = C + j * ldc;
        if (notb && nota) {
            if (beta == T(0)) for (i64 i = 0; i < m; ++i) c[i] = T(0);
            else if (beta != T(1)) for (i64 i = 0; i < m; ++i) c[i] *= beta;
            for (i64 l = 0; l < k; ++l) {
                T temp = alpha * B[IDX(l, j, ldb)];
                const T* a = A + l * lda;
                for (i64 i = 0; i < m; ++i) c[i] = std::fma(temp, a[i], c[i]);
            }
        } else if (!notb && nota) {
            if (beta == T(0)) for (i64 i = 0; i < m; ++i) c[i] = T(0);
            else if (beta != T(1)) for (i64 i = 0; i < m; ++i) c[i] *= beta;
            for (i64 l = 0; l < k; ++l) {
                T temp = alpha * B[IDX(j, l, ldb)];
                const T* a = A + l * lda;
                for (i64 i = 0; i < m; ++i) c[i] = std::fma(temp, a[i], c[i]);
            }
        } else {
            for (i64 i = 0; i < m; ++i) {
                T temp = T(0);
                for (i64 l = 0; l < k; ++l) {
                    T av = A[IDX(l, i, lda)];
                    T bv = notb ? B[IDX(l, j, ldb)] : B[IDX(j, l, ldb)];
                    temp = std::fma(av, bv, temp);
                }
                c[i] = (beta == T(0)) ? alpha * temp : alpha * temp + beta * c[i];
            }
        }
    }
}

template <typename T>
void nv_dsyrk(char uplo, char trans, i64 n, i64 k, T alpha, const T* A, i64 lda, T beta, T* C,
              i64 ldc) {
    for (i64 j = 0; j < n; ++j) {
        T* c = C + j * ldc;
        i64 i0 = (uplo == 'L') ? j : 0, i1 = (uplo == 'L') ? n : j + 1;
        if (trans == 'N') {
            if (beta == T(0)) for (i64 i = i0; i < i1; ++i) c[i] = T(0);
            else if (beta != T(1)) for (i64 i = i0; i < i1; ++i) c[i] *= beta;
            for (i64 l = 0; l < k; ++l) {
                T ajl = A[IDX(j, l, lda)];
                if (ajl == T(0)) continue;
                T temp = alpha * ajl;
                const T* a = A + l * lda;
                for (i64 i = i0; i < i1; ++i) c[i] = std::fma(temp, a[i], c[i]);
            }
        } else {
            for (i64 i = i0; i < i1; ++i) {
                T temp = T(0);
                for (i64 l = 0; l < k; ++l)
                    temp = std::fma(A[IDX(l, i, lda)], A[IDX(l, j, lda)], temp);
                c[i] = (beta == T(0)) ? alpha * temp : alpha * temp + beta * c[i];
            }
        }
    }
}

template <typename T>
void nv_dtrsm(char side, char uplo, char trans, char diag, i64 m, i64 n, T alpha, const T* A,
              i64 lda, T* B, i64 ldb) {
    if (m == 0 || n == 0) return;
    const bool nounit = (diag == 'N');
    if (alpha != T(1))
        for (i64 j = 0; j < n; ++j)
            for (i64 i = 0; i < m; ++i) B[IDX(i, j, ldb)] *= alpha;
    if (side == 'L') {
        if (trans == 'N') {
            if (uplo == 'U') {
                for (i64 j = 0; j < n; ++j) {
                    T* b = B + j * ldb;
                    for (i64 k = m - 1; k >= 0; --k) {
                        if (b[k] == T(0)) continue;
                        if (nounit) b[k] = b[k] / A[IDX(k, k, lda)];
                        T bk = -b[k];
                        const T* a = A + k * lda;
                        for (i64 i = 0; i < k; ++i) b[i] = std::fma(bk, a[i], b[i]);
                    }
                }
            } else {
                for (i64 j = 0; j < n; ++j) {
                    T* b = B + j * ldb;
                    for (i64 k = 0; k < m; ++k) {
                        if (b[k] == T(0)) continue;
                        if (nounit) b[k] = b[k] / A[IDX(k, k, lda)];
                        T bk = -b[k];
                        const T* a = A + k * lda;
                        for (i64 i = k + 1; i < m; ++i) b[i] = std::fma(bk, a[i], b[i]);
                    }
                }
            }
        } else {  // B := inv(A^T) B
            if (uplo == 'U') {
                for (i64 j = 0; j < n; ++j) {
                    T* b = B + j * ldb;
                    for (i64 i = 0; i < m; ++i) {
                        T temp = b[i];
                        const T* a = A + i * lda;
                        for (i64 k = 0; k < i; ++k) temp = std::fma(-a[k], b[k], temp);
                        if (nounit) temp = temp / a[i];
                        b[i] = temp;
                    }
                }
            } else {
                for (i64 j = 0; j < n; ++j) {
                    T* b = B + j * ldb;
                    for (i64 i = m - 1; i >= 0; --i) {
                        T temp = b[i];
                        const T* a = A + i * lda;
                        for (i64 k = i + 1; k < m; ++k) temp = std::fma(-a[k], b[k], temp);
                        if (nounit) temp = temp / a[i];
                        b[i] = temp;
                    }
                }
            }
        }
    } else {  // side == 'R'
        if (trans == 'N') {
            if (uplo == 'U') {
                for (i64 j = 0; j < n; ++j) {
                    T* bj = B + j * ldb;
                    for (i64 k = 0; k < j; ++k) {
                        T akj = A[IDX(k, j, lda)];
                        if (akj == T(0)) continue;
                        const T* bk = B + k * ldb;
                        for (i64 i = 0; i < m; ++i) bj[i] = std::fma(-akj, bk[i], bj[i]);
                    }
                    if (nounit) {
                        T temp = T(1) / A[IDX(j, j, lda)];
                        for (i64 i = 0; i < m; ++i) bj[i] = temp * bj[i];
                    }
                }
            } else {
                for (i64 j = n - 1; j >= 0; --j) {
                    T* bj = B + j * ldb;
                    for (i64 k = j + 1; k < n; ++k) {
                        T akj = A[IDX(k, j, lda)];
                        if (akj == T(0)) continue;
                        const T* bk = B + k * ldb;
                        for (i64 i = 0; i < m; ++i) bj[i] = std::fma(-akj, bk[i], bj[i]);
                    }
                    if (nounit) {
                        T temp = T(1) / A[IDX(j, j, lda)];
                        for (i64 i = 0; i < m; ++i) bj[i] = temp * bj[i];
                    }
                }
            }
        } else {  // B := B inv(A^T)
            if (uplo == 'U') {
                for (i64 k = n - 1; k >= 0; --k) {
                    T* bk = B + k * ldb;
                    if (nounit) {
                        T temp = T(1) / A[IDX(k, k, lda)];
                        for (i64 i = 0; i < m; ++i) bk[i] = temp * bk[i];
                    }
                    for (i64 j = 0; j < k; ++j) {
                        T ajk = A[IDX(j, k, lda)];
                        if (ajk == T(0)) continue;
                        T* bj = B + j * ldb;
                        for (i64 i = 0; i < m; ++i) bj[i] = std::fma(-ajk, bk[i], bj[i]);
                    }
                }
            } else {
                for (i64 k = 0; k < n; ++k) {
                    T* bk = B + k * ldb;
                    if (nounit) {
                        T temp = T(1) / A[IDX(k, k, lda)];
                        for (i64 i = 0; i < m; ++i) bk[i] = temp * bk[i];
                    }
                    for (i64 j = k + 1; j < n; ++j) {
                        T ajk = A[IDX(j, k, lda)];
                        if (ajk == T(0)) continue;
                        T* bj = B + j * ldb;
                        for (i64 i = 0; i < m; ++i) bj[i] = std::fma(-ajk, bk[i], bj[i]);
                    }
                }
            }
        }
    }
}

Blas g_naive = {nv_idamax<double>, nv_dswap<double>, nv_dscal<double>, nv_dger<double>,
                nv_dsyr<double>,   nv_dtrsm<double>, nv_dgemm<double>, nv_dsyrk<double>};

// ---------------- host BLAS through dlopen (Fortran ABI, LP64) ----------------
typedef int fint;
struct FortranBlas {
    void* handle = nullptr;
    std::string path, prefix;
    fint (*idamax_)(const fint*, const double*, const fint*) = nullptr;
    void (*dswap_)(const fint*, double*, const fint*, double*, const fint*) = nullptr;
    void (*dscal_)(const fint*, const double*, double*, const fint*) = nullptr;
    void (*dger_)(const fint*, const fint*, const double*, const double*, const fint*,
                  const double*, const fint*, double*, const fint*) = nullptr;
    void (*dsyr_)(const char*, const fint*, const double*, const double*, const fint*,
                  double*, const fint*) = nullptr;
    void (*dtrsm_)(const char*, const char*, const char*, const char*, const fint*,
                   const fint*, const double*, const double*, const fint*, double*,
                   const fint*) = nullptr;
    void (*dgemm_)(const char*, const char*, const fint*, const fint*, const fint*,
                   const double*, const double*, const fint*, const double*, const fint*,
                   const double*, double*, const fint*) = nullptr;
    void (*dsyrk_)(const char*, const char*, const fint*, const fint*, const double*,
                   const double*, const fint*, const double*, double*, const fint*) = nullptr;
    void (*dgetrf_)(const fint*, const fint*, double*, const fint*, fint*, fint*) = nullptr;
    void (*dpotrf_)(const char*, const fint*, double*, const fint*, fint*) = nullptr;
    void (*set_threads)(int) = nullptr;
    int (*get_threads)() = nullptr;
};
FortranBlas g_fb;

i64 fb_idamax(i64 n, const double* x, i64 incx) {
    fint nn = (fint)n, ii = (fint)incx;
    return (i64)g_fb.idamax_(&nn, x, &ii) - 1;
}
void fb_dswap(i64 n, double* x, i64 incx, double* y, i64 incy) {
    if (n <= 0) return;
    fint nn = (fint)n, a = (fint)incx, b = (fint)incy;
    g_fb.dswap_(&nn, x, &a, y, &b);
}
void fb_dscal(i64 n, double al, double* x, i64 incx) {
    if (n <= 0) return;
    fint nn = (fint)n, a = (fint)incx;
    g_fb.dscal_(&nn, &al, x, &a);
}
void fb_dger(i64 m, i64 n, double alpha, const double* x, i64 incx, const double* y, i64 incy,
             double* A, i64 lda) {
    if (m <= 0 || n <= 0) return;
    fint mm = (fint)m, nn = (fint)n, ix = (fint)incx, iy = (fint)incy, ld = (fint)lda;
    g_fb.dger_(&mm, &nn, &alpha, x, &ix, y, &iy, A, &ld);
}
void fb_dsyr(char uplo, i64 n, double alpha, const double* x, i64 incx, double* A, i64 lda) {
    if (n <= 0) return;
    fint nn = (fint)n, ix = (fint)incx, ld = (fint)lda;
    g_fb.dsyr_(&uplo, &nn, &alpha, x, &ix, A, &ld);
}
void fb_dtrsm(char side, char uplo, char trans, char diag, i64 m, i64 n, double alpha,
              const double* A, i64 lda, double* B, i64 ldb) {
    if (m <= 0 || n <= 0) return;
    fint mm = (fint)m, nn = (fint)n, la = (fint)std::max<i64>(lda, 1), lb = (fint)std::max<i64>(ldb, 1);
    g_fb.dtrsm_(&side, &uplo, &trans, &diag, &mm, &nn, &alpha, A, &la, B, &lb);
}
void fb_dgemm(char ta, char tb, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda,
              const double* B, i64 ldb, double beta, double* C, i64 ldc) {
    if (m <= 0 || n <= 0) return;
    fint mm = (fint)m, nn = (fint)n, kk = (fint)k, la = (fint)std::max<i64>(lda, 1),
         lb = (fint)std::max<i64>(ldb, 1), lc = (fint)std::max<i64>(ldc, 1);
    g_fb.dgemm_(&ta, &tb, &mm, &nn, &kk, &alpha, A, &la, B, &lb, &beta, C, &lc);
}
void fb_dsyrk(char uplo, char trans, i64 n, i64 k, double alpha, const double* A, i64 lda,
              double beta, double* C, i64 ldc) {
    if (n <= 0) return;
    fint nn = (fint)n, kk = (fint)k, la = (fint)std::max<i64>(lda, 1), lc = (fint)std::max<i64>(ldc, 1);
    g_fb.dsyrk_(&uplo, &trans, &nn, &kk, &alpha, A, &la, &beta, C, &lc);
}

Blas g_host = {fb_idamax, fb_dswap, fb_dscal, fb_dger, fb_dsyr, fb_dtrsm, fb_dgemm, fb_dsyrk};

Blas* g_blas = &g_naive;
std::string g_backend = "naive";

template <typename T>
bool load_sym(void* h, const std::string& prefix, const char* name, T& out) {
    std::string s = prefix + name;
    out = reinterpret_cast<T>(dlsym(h, s.c_str()));
    return out != nullptr;
}

bool open_fortran_blas(const char* path, const char* prefix) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return false;
    FortranBlas fb;
    fb.handle = h;
    fb.path = path;
    fb.prefix = prefix;
    std::string p = prefix;
    bool ok = load_sym(h, p, "idamax_", fb.idamax_) && load_sym(h, p, "dswap_", fb.dswap_) &&
              load_sym(h, p, "dscal_", fb.dscal_) && load_sym(h, p, "dger_", fb.dger_) &&
              load_sym(h, p, "dsyr_", fb.dsyr_) && load_sym(h, p, "dtrsm_", fb.dtrsm_) &&
              load_sym(h, p, "dgemm_", fb.dgemm_) && load_sym(h, p, "dsyrk_", fb.dsyrk_);
    if (!ok) {
        dlclose(h);
        return false;
    }
    load_sym(h, p, "dgetrf_", fb.dgetrf_);
    load_sym(h, p, "dpotrf_", fb.dpotrf_);
    if (!load_sym(h, "", "MKL_Set_Num_Threads", fb.set_threads))
        if (!load_sym(h, p, "openblas_set_num_threads", fb.set_threads))
            load_sym(h, "", "openblas_set_num_threads", fb.set_threads);
    if (!load_sym(h, "", "MKL_Get_Max_Threads", fb.get_threads))
        if (!load_sym(h, p, "openblas_get_num_threads", fb.get_threads))
            load_sym(h, "", "openblas_get_num_threads", fb.get_threads);
    g_fb = fb;
    return true;
}

// ---------------------------------------------------------------------------------------
// Typed BLAS dispatch: double through the selected backend, float through the naive one (the
// FP32 parity tests use the naive backend only).
// ---------------------------------------------------------------------------------------
template <typename T>
struct Ops;
template <>
struct Ops<double> {
    static i64 iamax(i64 n, const double* x, i64 incx) { return g_blas->idamax(n, x, incx); }
    static void swap(i64 n, double* x, i64 incx, double* y, i64 incy) {
        g_blas->dswap(n, x, incx, y, incy);
    }
    static void scal(i64 n, double a, double* x, i64 incx) { g_blas->dscal(n, a, x, incx); }
    static void ger(i64 m, i64 n, double al, const double* x, i64 incx, const double* y,
                    i64 incy, double* A, i64 lda) {
        g_blas->dger(m, n, al, x, incx, y, incy, A, lda);
    }
    static void syr(char u, i64 n, double al, const double* x, i64 incx, double* A, i64 lda) {
        g_blas->dsyr(u, n, al, x, incx, A, lda);
    }
    static void trsm(char sd, char u, char t, char d, i64 m, i64 n, double al, const double* A,
                     i64 lda, double* B, i64 ldb) {
        g_blas->dtrsm(sd, u, t, d, m, n, al, A, lda, B, ldb);
    }
    static void gemm(char ta, char tb, i64 m, i64 n, i64 k, double al, const double* A, i64 lda,
                     const double* B, i64 ldb, double be, double* C, i64 ldc) {
        g_blas->dgemm(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
    }
    static void syrk(char u, char t, i64 n, i64 k, double al, const double* A, i64 lda,
                     double be, double* C, i64 ldc) {
        g_blas->dsyrk(u, t, n, k, al, A, lda, be, C, ldc);
    }
};
template <>
struct Ops<float> {
    static i64 iamax(i64 n, const float* x, i64 incx) { return nv_idamax<float>(n, x, incx); }
    static void swap(i64 n, float* x, i64 incx, float* y, i64 incy) {
        nv_dswap<float>(n, x, incx, y, incy);
    }
    static void scal(i64 n, float a, float* x, i64 incx) { nv_dscal<float>(n, a, x, incx); }
    static void ger(i64 m, i64 n, float al, const float* x, i64 incx, const float* y, i64 incy,
                    float* A, i64 lda) {
        nv_dger<float>(m, n, al, x, incx, y, incy, A, lda);
    }
    static void syr(char u, i64 n, float al, const float* x, i64 incx, float* A, i64 lda) {
        nv_dsyr<float>(u, n, al, x, incx, A, lda);
    }
    static void trsm(char sd, char u, char t, char d, i64 m, i64 n, float al, const float* A,
                     i64 lda, float* B, i64 ldb) {
        nv_dtrsm<float>(sd, u, t, d, m, n, al, A, lda, B, ldb);
    }
    static void gemm(char ta, char tb, i64 m, i64 n, i64 k, float al, const float* A, i64 lda,
                     const float* B, i64 ldb, float be, float* C, i64 ldc) {
        nv_dgemm<float>(ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
    }
    static void syrk(char u, char t, i64 n, i64 k, float al, const float* A, i64 lda, float be,
                     float* C, i64 ldc) {
        nv_dsyrk<float>(u, t, n, k, al, A, lda, be, C, ldc);
    }
};

std::mt19937& generator() {
    static std::mt19937 gen((21L << 16) | 0L);
    return gen;
}

// ---------------------------------------------------------------------------------------
// Elemental's sequential algorithms (templates over the field, instantiated for double and
// float as the reference's PROTO macros do, e.g. src/lapack_like/factor/LU.cpp:233-313)
// ---------------------------------------------------------------------------------------

// lu::Panel, src/lapack_like/factor/LU/Panel.hpp:15-60. piv[c] receives the swap
// destination of step c relative to the panel top (PB.Swap(c, iPiv)). Returns 0, or c+1 for
// an exactly zero pivot at panel column c (the point where SingularMatrixException is thrown).
// When trace_col is in [0,n), the candidate column of that step (rows k..m-1, before the
// pivot search) is copied to trace_out and the panel stops, returning -1.
template <typename T>
int lu_panel(T* A, i64 m, i64 n, i64 lda, int32_t* piv, i64 trace_col = -1,
             T* trace_out = nullptr) {
    for (i64 k = 0; k < n; ++k) {
        const i64 horz = n - (k + 1), vert = m - (k + 1);
        if (k == trace_col) {
            for (i64 i = k; i < m; ++i) trace_out[i - k] = A[IDX(i, k, lda)];
            return -1;
        }
        const i64 maxInd = Ops<T>::iamax(vert + 1, &A[IDX(k, k, lda)], 1);
        const i64 iPiv = maxInd + k;
        piv[k] = (int32_t)iPiv;
        if (iPiv != k) Ops<T>::swap(n, &A[k], lda, &A[iPiv], lda);
        const T alpha = A[IDX(k, k, lda)];
        if (alpha == T(0)) return (int)(k + 1);
        const T alphaInv = T(1) / alpha;
        Ops<T>::scal(vert, alphaInv, &A[IDX(k + 1, k, lda)], 1);
        Ops<T>::ger(vert, horz, T(-1), &A[IDX(k + 1, k, lda)], 1, &A[IDX(k, k + 1, lda)], lda,
                    &A[IDX(k + 1, k + 1, lda)], lda);
    }
    return 0;
}

// lu::Unb (LU without pivoting), src/lapack_like/factor/LU/Local.hpp:44-61: Scal by
// 1/alpha, then Geru(-1). Returns 0, or j+1 for an exactly zero pivot at column j.
template <typename T>
int lu_unb(T* A, i64 m, i64 n, i64 lda) {
    for (i64 j = 0; j < std::min(m, n); ++j) {
        const T alpha = A[IDX(j, j, lda)];
        if (alpha == T(0)) return (int)(j + 1);
        Ops<T>::scal(m - (j + 1), T(1) / alpha, &A[IDX(j + 1, j, lda)], 1);
        Ops<T>::ger(m - (j + 1), n - (j + 1), T(-1), &A[IDX(j + 1, j, lda)], 1,
                    &A[IDX(j, j + 1, lda)], lda, &A[IDX(j + 1, j + 1, lda)], lda);
    }
    return 0;
}

// Permutation::PermuteRows with implicit origins j (Permutation.cpp:544-580) applied
// through RowSwap (src/blas_like/level1/Swap.cpp:99-116).
template <typename T>
void permute_rows(const int32_t* dests, i64 nswaps, T* A, i64 height, i64 width, i64 lda) {
    if (height == 0 || width == 0) return;
    for (i64 j = 0; j < nswaps; ++j)
        if (dests[j] != j) Ops<T>::swap(width, &A[j], lda, &A[dests[j]], lda);
}

// Permutation::InversePermuteRows (Permutation.cpp:586-620): the swaps in reverse order.
template <typename T>
void inverse_permute_rows(const int32_t* dests, i64 nswaps, T* A, i64 height, i64 width,
                          i64 lda) {
    if (height == 0 || width == 0) return;
    for (i64 j = nswaps - 1; j >= 0; --j)
        if (dests[j] != j) Ops<T>::swap(width, &A[j], lda, &A[dests[j]], lda);
}

template <typename T>
int chol_unb_lower(T* A, i64 n, i64 lda) {
    // LowerVariant3Unblocked, Cholesky/LowerVariant3.hpp:15-41
    for (i64 j = 0; j < n; ++j) {
        T alpha = A[IDX(j, j, lda)];
        if (alpha <= T(0)) return (int)(j + 1);
        alpha = std::sqrt(alpha);
        A[IDX(j, j, lda)] = alpha;
        const i64 h = n - (j + 1);
        Ops<T>::scal(h, T(1) / alpha, &A[IDX(j + 1, j, lda)], 1);
        Ops<T>::syr('L', h, T(-1), &A[IDX(j + 1, j, lda)], 1, &A[IDX(j + 1, j + 1, lda)], lda);
    }
    return 0;
}

template <typename T>
int chol_unb_upper(T* A, i64 n, i64 lda) {
    // UpperVariant3Unblocked, Cholesky/UpperVariant3.hpp:15-46 (Conj is the identity for
    // real data)
    for (i64 j = 0; j < n; ++j) {
        T alpha = A[IDX(j, j, lda)];
        if (alpha <= T(0)) return (int)(j + 1);
        alpha = std::sqrt(alpha);
        A[IDX(j, j, lda)] = alpha;
        const i64 w = n - (j + 1);
        Ops<T>::scal(w, T(1) / alpha, &A[IDX(j, j + 1, lda)], lda);
        Ops<T>::syr('U', w, T(-1), &A[IDX(j, j + 1, lda)], lda, &A[IDX(j + 1, j + 1, lda)], lda);
    }
    return 0;
}

// El::LU(Matrix<F>&, Permutation&), src/lapack_like/factor/LU.cpp:104-142.
// dests[0..minDim) receives Permutation::SwapDestinations() (0-based). At an exactly zero
// pivot the function returns (global column + 1) with A as the reference leaves it when
// lu::Panel throws: the earlier blocks finished, the current panel's rows swapped up to the
// failing column, the columns outside the panel not yet permuted for this block.
template <typename T>
int lu_impl(T* A, i64 m, i64 n, i64 lda, i64 bsize, int32_t* dests, i64 trace_col,
            T* trace_out) {
    const i64 minDim = std::min(m, n);
    std::vector<int32_t> piv(bsize);
    for (i64 k = 0; k < minDim; k += bsize) {
        const i64 nb = std::min(bsize, minDim - k);
        T* AB0 = A + k;
        T* AB1 = A + IDX(k, k, lda);
        T* AB2 = A + IDX(k, k + nb, lda);
        const i64 tc = (trace_col >= k && trace_col < k + nb) ? trace_col - k : -1;
        const int pinfo = lu_panel(AB1, m - k, nb, lda, piv.data(), tc, trace_out);
        if (pinfo < 0) return -1;
        for (i64 c = 0; c < nb; ++c) dests[k + c] = (int32_t)(k + piv[c]);
        if (pinfo) return (int)(k + pinfo);
        permute_rows(piv.data(), nb, AB0, m - k, k, lda);
        permute_rows(piv.data(), nb, AB2, m - k, n - k - nb, lda);
        T* A11 = AB1;
        T* A12 = AB2;
        T* A21 = A + IDX(k + nb, k, lda);
        T* A22 = A + IDX(k + nb, k + nb, lda);
        Ops<T>::trsm('L', 'L', 'N', 'U', nb, n - k - nb, T(1), A11, lda, A12, lda);
        Ops<T>::gemm('N', 'N', m - k - nb, n - k - nb, nb, T(-1), A21, lda, A12, lda, T(1), A22,
                     lda);
    }
    return 0;
}

// El::LU(Matrix<F>&) without pivoting, src/lapack_like/factor/LU.cpp:21-44: lu::Unb on A11,
// Trsm(RIGHT,UPPER,NORMAL,NON_UNIT) for A21, Trsm(LEFT,LOWER,NORMAL,UNIT) for A12, Gemm.
template <typename T>
int lu_nopiv_impl(T* A, i64 m, i64 n, i64 lda, i64 bsize) {
    const i64 minDim = std::min(m, n);
    for (i64 k = 0; k < minDim; k += bsize) {
        const i64 nb = std::min(bsize, minDim - k);
        T* A11 = A + IDX(k, k, lda);
        T* A12 = A + IDX(k, k + nb, lda);
        T* A21 = A + IDX(k + nb, k, lda);
        T* A22 = A + IDX(k + nb, k + nb, lda);
        const int info = lu_unb(A11, nb, nb, lda);
        if (info) return (int)(k + info);
        Ops<T>::trsm('R', 'U', 'N', 'N', m - k - nb, nb, T(1), A11, lda, A21, lda);
        Ops<T>::trsm('L', 'L', 'N', 'U', nb, n - k - nb, T(1), A11, lda, A12, lda);
        Ops<T>::gemm('N', 'N', m - k - nb, n - k - nb, nb, T(-1), A21, lda, A12, lda, T(1), A22,
                     lda);
    }
    return 0;
}

// lu::Full (LU with full pivoting), src/lapack_like/factor/LU/Full.hpp:15-58: per step k,
// MaxAbsLoc of the trailing block A(k:,k:) (src/blas_like/level1/MaxAbsLoc.cpp:131-163: a
// column-major scan with strict '>', so ties go to the first entry in column-major order and
// an all-zero block gives its top-left entry), RowSwap(A,k,iPiv) and ColSwap(A,k,jPiv) over the
// whole matrix, then Scal by 1/alpha and Geru(-1). p and q receive P's and Q's swap
// destinations. Returns 0, or k+1 for an exactly zero pivot (SingularMatrixException).
template <typename T>
int lu_full_impl(T* A, i64 m, i64 n, i64 lda, int32_t* p, int32_t* q) {
    const i64 minDim = std::min(m, n);
    for (i64 k = 0; k < minDim; ++k) {
        T best = T(0);
        i64 bi = 0, bj = 0;
        for (i64 j = k; j < n; ++j)
            for (i64 i = k; i < m; ++i) {
                const T v = std::fabs(A[IDX(i, j, lda)]);
                if (v > best) {
                    best = v;
                    bi = i - k;
                    bj = j - k;
                }
            }
        const i64 iPiv = k + bi, jPiv = k + bj;
        p[k] = (int32_t)iPiv;
        q[k] = (int32_t)jPiv;
        if (iPiv != k) Ops<T>::swap(n, &A[k], lda, &A[iPiv], lda);
        if (jPiv != k) Ops<T>::swap(m, &A[IDX(0, k, lda)], 1, &A[IDX(0, jPiv, lda)], 1);
        const T alpha = A[IDX(k, k, lda)];
        if (alpha == T(0)) return (int)(k + 1);
        Ops<T>::scal(m - (k + 1), T(1) / alpha, &A[IDX(k + 1, k, lda)], 1);
        Ops<T>::ger(m - (k + 1), n - (k + 1), T(-1), &A[IDX(k + 1, k, lda)], 1,
                    &A[IDX(k, k + 1, lda)], lda, &A[IDX(k + 1, k + 1, lda)], lda);
    }
    return 0;
}

// lu::SolveAfter(orientation, A, P, Q, B), src/lapack_like/factor/LU/SolveAfter.hpp:123-151.
template <typename T>
int lu_full_solve_after(int orient, const T* A, i64 n, i64 lda, const int32_t* p,
                        const int32_t* q, i64 nswaps, T* B, i64 nrhs, i64 ldb) {
    if (orient == 0) {
        permute_rows(p, nswaps, B, n, nrhs, ldb);
        Ops<T>::trsm('L', 'L', 'N', 'U', n, nrhs, T(1), A, lda, B, ldb);
        Ops<T>::trsm('L', 'U', 'N', 'N', n, nrhs, T(1), A, lda, B, ldb);
        inverse_permute_rows(q, nswaps, B, n, nrhs, ldb);
    } else {
        permute_rows(q, nswaps, B, n, nrhs, ldb);
        Ops<T>::trsm('L', 'U', 'T', 'N', n, nrhs, T(1), A, lda, B, ldb);
        Ops<T>::trsm('L', 'L', 'T', 'U', n, nrhs, T(1), A, lda, B, ldb);
        inverse_permute_rows(p, nswaps, B, n, nrhs, ldb);
    }
    return 0;
}

// lu::SolveAfter(orientation, A, [P,] B), src/lapack_like/factor/LU/SolveAfter.hpp:15-38
// (no pivots: dests null) and :66-92. orient: 0 NORMAL, 1 TRANSPOSE, 2 ADJOINT (identical
// for real data).
template <typename T>
int lu_solve_after(int orient, const T* A, i64 n, i64 lda, const int32_t* dests, i64 ndests,
                   T* B, i64 nrhs, i64 ldb) {
    if (orient == 0) {
        if (dests) permute_rows(dests, ndests, B, n, nrhs, ldb);
        Ops<T>::trsm('L', 'L', 'N', 'U', n, nrhs, T(1), A, lda, B, ldb);
        Ops<T>::trsm('L', 'U', 'N', 'N', n, nrhs, T(1), A, lda, B, ldb);
    } else {
        Ops<T>::trsm('L', 'U', 'T', 'N', n, nrhs, T(1), A, lda, B, ldb);
        Ops<T>::trsm('L', 'L', 'T', 'U', n, nrhs, T(1), A, lda, B, ldb);
        if (dests) inverse_permute_rows(dests, ndests, B, n, nrhs, ldb);
    }
    return 0;
}

// El::LinearSolve(A, B): ACopy(A), RowEchelon(ACopy, B), Trsm(L,U,N,N)
// (src/lapack_like/solve/Linear.cpp:30-76,176-183,212-218). dests (optional, may be null)
// receives the swap destinations RowEchelon records.
template <typename T>
int linear_solve(const T* A, i64 n, i64 lda, T* B, i64 nrhs, i64 ldb, i64 bsize,
                 int32_t* dests) {
    const i64 ld = std::max<i64>(n, 1);
    std::vector<T> C((size_t)(ld * n));
    for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < n; ++i) C[IDX(i, j, ld)] = A[IDX(i, j, lda)];
    T* Ac = C.data();
    std::vector<int32_t> piv(bsize);
    for (i64 k = 0; k < n; k += bsize) {
        const i64 nb = std::min(bsize, n - k);
        T* AB1 = Ac + IDX(k, k, ld);
        T* AB2 = Ac + IDX(k, k + nb, ld);
        T* BB = B + k;
        const int pinfo = lu_panel(AB1, n - k, nb, ld, piv.data());
        if (dests)
            for (i64 c = 0; c < nb; ++c) dests[k + c] = (int32_t)(k + piv[c]);
        if (pinfo) return (int)(k + pinfo);
        permute_rows(piv.data(), nb, AB2, n - k, n - k - nb, ld);
        permute_rows(piv.data(), nb, BB, n - k, nrhs, ldb);
        T* A11 = AB1;
        T* A12 = AB2;
        T* A21 = Ac + IDX(k + nb, k, ld);
        T* A22 = Ac + IDX(k + nb, k + nb, ld);
        T* B1 = B + k;
        T* B2 = B + k + nb;
        Ops<T>::trsm('L', 'L', 'N', 'U', nb, n - k - nb, T(1), A11, ld, A12, ld);
        Ops<T>::trsm('L', 'L', 'N', 'U', nb, nrhs, T(1), A11, ld, B1, ldb);
        Ops<T>::gemm('N', 'N', n - k - nb, n - k - nb, nb, T(-1), A21, ld, A12, ld, T(1), A22, ld);
        Ops<T>::gemm('N', 'N', n - k - nb, nrhs, nb, T(-1), A21, ld, B1, ldb, T(1), B2, ldb);
    }
    Ops<T>::trsm('L', 'U', 'N', 'N', n, nrhs, T(1), Ac, ld, B, ldb);
    return 0;
}

// El::Cholesky(uplo, A): uplo 0 LOWER, 1 UPPER (Cholesky.cpp:26-38, variant 3).
template <typename T>
int cholesky(int uplo, T* A, i64 n, i64 lda, i64 bsize) {
    for (i64 k = 0; k < n; k += bsize) {
        const i64 nb = std::min(bsize, n - k);
        T* A11 = A + IDX(k, k, lda);
        T* A22 = A + IDX(k + nb, k + nb, lda);
        if (uplo == 0) {
            const int info = chol_unb_lower(A11, nb, lda);
            if (info) return (int)(k + info);
            T* A21 = A + IDX(k + nb, k, lda);
            Ops<T>::trsm('R', 'L', 'T', 'N', n - k - nb, nb, T(1), A11, lda, A21, lda);
            Ops<T>::syrk('L', 'N', n - k - nb, nb, T(-1), A21, lda, T(1), A22, lda);
        } else {
            const int info = chol_unb_upper(A11, nb, lda);
            if (info) return (int)(k + info);
            T* A12 = A + IDX(k, k + nb, lda);
            Ops<T>::trsm('L', 'U', 'T', 'N', nb, n - k - nb, T(1), A11, lda, A12, lda);
            Ops<T>::syrk('U', 'T', n - k - nb, nb, T(-1), A12, lda, T(1), A22, lda);
        }
    }
    return 0;
}

// cholesky::SolveAfter(uplo, orientation, A, B), Cholesky/SolveAfter.hpp:15-43.
// (TRANSPOSE conjugates B before and after, the identity for real data.)
template <typename T>
int cholesky_solve_after(int uplo, const T* A, i64 n, i64 lda, T* B, i64 nrhs, i64 ldb) {
    if (uplo == 0) {
        Ops<T>::trsm('L', 'L', 'N', 'N', n, nrhs, T(1), A, lda, B, ldb);
        Ops<T>::trsm('L', 'L', 'T', 'N', n, nrhs, T(1), A, lda, B, ldb);
    } else {
        Ops<T>::trsm('L', 'U', 'T', 'N', n, nrhs, T(1), A, lda, B, ldb);
        Ops<T>::trsm('L', 'U', 'N', 'N', n, nrhs, T(1), A, lda, B, ldb);
    }
    return 0;
}

// El::HPDSolve(uplo, orientation, A, B), src/lapack_like/solve/HPD.cpp:15-25,47-57.
template <typename T>
int hpd_solve(int uplo, const T* A, i64 n, i64 lda, T* B, i64 nrhs, i64 ldb, i64 bsize) {
    const i64 ld = std::max<i64>(n, 1);
    std::vector<T> C((size_t)(ld * n));
    for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < n; ++i) C[IDX(i, j, ld)] = A[IDX(i, j, lda)];
    const int info = cholesky(uplo, C.data(), n, ld, bsize);
    if (info) return info;
    return cholesky_solve_after(uplo, C.data(), n, ld, B, nrhs, ldb);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------
// Backend control
// ---------------------------------------------------------------------------------------

// name: "naive", "mkl", "openblas" or a path to a Fortran-ABI BLAS. Returns 0 on success.
int orc_set_backend(const char* name) {
    std::string s = name ? name : "naive";
    if (s == "naive") {
        g_blas = &g_naive;
        g_backend = "naive";
        return 0;
    }
    std::vector<std::pair<std::string, std::string>> cands;
    if (s == "mkl" || s == "auto") {
        cands.push_back({"libmkl_rt.so", ""});
        cands.push_back({"/opt/conda/lib/libmkl_rt.so", ""});
        cands.push_back({"libmkl_rt.so.1", ""});
    }
    if (s == "openblas" || s == "auto") {
        cands.push_back({"libopenblas.so.0", ""});
        cands.push_back({"libopenblas.so", ""});
    }
    if (s.find('/') != std::string::npos) {
        std::string prefix = (s.find("scipy_openblas") != std::string::npos) ? "scipy_" : "";
        cands.push_back({s, prefix});
    }
    for (auto& c : cands) {
        if (open_fortran_blas(c.first.c_str(), c.second.c_str())) {
            g_blas = &g_host;
            g_backend = c.first;
            return 0;
        }
    }
    return -1;
}

const char* orc_backend() { return g_backend.c_str(); }

void orc_set_threads(int nt) {
    if (g_blas == &g_host && g_fb.set_threads) g_fb.set_threads(nt);
    omp_set_num_threads(nt);
}

int orc_get_threads() {
    if (g_blas == &g_host && g_fb.get_threads) return g_fb.get_threads();
    return 1;
}

// ---------------------------------------------------------------------------------------
// Generators (Elemental's deterministic stream: std::mt19937 seeded with (21<<16)|rank)
// ---------------------------------------------------------------------------------------
void orc_rng_seed(uint64_t seed) { generator().seed((std::mt19937::result_type)seed); }

// MakeGaussian: one fresh normal_distribution per entry, column-major fill
// (Gaussian.cpp:16-23, random/impl.hpp:185-203, EntrywiseFill.hpp:14-23).
void orc_gaussian(double* A, i64 m, i64 n, i64 lda, double mean, double stddev) {
    for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < m; ++i) {
            std::normal_distribution<double> d(mean, stddev);
            A[IDX(i, j, lda)] = d(generator());
        }
}

// GEPPGrowth, src/matrices/deterministic/misc/GEPPGrowth.cpp:15-30.
void orc_geppgrowth(double* A, i64 n, i64 lda) {
    for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < n; ++i) A[IDX(i, j, lda)] = (i == j) ? 1.0 : 0.0;
    if (n <= 1) return;
    for (i64 i = 0; i < n; ++i) A[IDX(i, n - 1, lda)] = 1.0;
    for (i64 j = 1; j < n; ++j)
        for (i64 i = j; i < n; ++i) A[IDX(i, i - j, lda)] = -1.0;
}

// ---------------------------------------------------------------------------------------
// Factorizations and solves. Return 0, or (global column index + 1) at the point where
// the reference throws SingularMatrixException / NonHPDMatrixException.
// ---------------------------------------------------------------------------------------

int orc_lu(double* A, i64 m, i64 n, i64 lda, i64 bsize, int32_t* dests) {
    return lu_impl(A, m, n, lda, bsize, dests, -1, (double*)nullptr);
}
int orc_slu(float* A, i64 m, i64 n, i64 lda, i64 bsize, int32_t* dests) {
    return lu_impl(A, m, n, lda, bsize, dests, -1, (float*)nullptr);
}

// Runs El::LU up to global step `col` and copies that step's candidate column (rows
// col..m-1 of column col, before the pivot search) to out. Used to classify pivot
// disagreements as near-ties. A is overwritten with the partial factorization.
int orc_lu_trace_column(double* A, i64 m, i64 n, i64 lda, i64 bsize, i64 col, double* out) {
    std::vector<int32_t> d((size_t)std::min(m, n));
    return lu_impl(A, m, n, lda, bsize, d.data(), col, out);
}
int orc_slu_trace_column(float* A, i64 m, i64 n, i64 lda, i64 bsize, i64 col, float* out) {
    std::vector<int32_t> d((size_t)std::min(m, n));
    return lu_impl(A, m, n, lda, bsize, d.data(), col, out);
}

// El::LU(A) without pivoting (LU.cpp:21-44).
int orc_lu_nopiv(double* A, i64 m, i64 n, i64 lda, i64 bsize) {
    return lu_nopiv_impl(A, m, n, lda, bsize);
}
int orc_slu_nopiv(float* A, i64 m, i64 n, i64 lda, i64 bsize) {
    return lu_nopiv_impl(A, m, n, lda, bsize);
}

// El::LU(A, P, Q) with full pivoting and its SolveAfter.
int orc_lu_full(double* A, i64 m, i64 n, i64 lda, int32_t* p, int32_t* q) {
    return lu_full_impl(A, m, n, lda, p, q);
}
int orc_slu_full(float* A, i64 m, i64 n, i64 lda, int32_t* p, int32_t* q) {
    return lu_full_impl(A, m, n, lda, p, q);
}
int orc_lu_full_solve_after(int orient, const double* A, i64 n, i64 lda, const int32_t* p,
                            const int32_t* q, i64 nswaps, double* B, i64 nrhs, i64 ldb) {
    return lu_full_solve_after(orient, A, n, lda, p, q, nswaps, B, nrhs, ldb);
}
int orc_slu_full_solve_after(int orient, const float* A, i64 n, i64 lda, const int32_t* p,
                             const int32_t* q, i64 nswaps, float* B, i64 nrhs, i64 ldb) {
    return lu_full_solve_after(orient, A, n, lda, p, q, nswaps, B, nrhs, ldb);
}

// lu::SolveAfter(orientation, A, P, B) (dests non-null) or (orientation, A, B) (dests null).
int orc_lu_solve_after(int orient, const double* A, i64 n, i64 lda, const int32_t* dests,
                       i64 ndests, double* B, i64 nrhs, i64 ldb) {
    return lu_solve_after(orient, A, n, lda, dests, ndests, B, nrhs, ldb);
}
int orc_slu_solve_after(int orient, const float* A, i64 n, i64 lda, const int32_t* dests,
                        i64 ndests, float* B, i64 nrhs, i64 ldb) {
    return lu_solve_after(orient, A, n, lda, dests, ndests, B, nrhs, ldb);
}

int orc_linear_solve(const double* A, i64 n, i64 lda, double* B, i64 nrhs, i64 ldb, i64 bsize,
                     int32_t* dests) {
    return linear_solve(A, n, lda, B, nrhs, ldb, bsize, dests);
}
int orc_slinear_solve(const float* A, i64 n, i64 lda, float* B, i64 nrhs, i64 ldb, i64 bsize,
                      int32_t* dests) {
    return linear_solve(A, n, lda, B, nrhs, ldb, bsize, dests);
}

int orc_cholesky(int uplo, double* A, i64 n, i64 lda, i64 bsize) {
    return cholesky(uplo, A, n, lda, bsize);
}
int orc_scholesky(int uplo, float* A, i64 n, i64 lda, i64 bsize) {
    return cholesky(uplo, A, n, lda, bsize);
}

int orc_cholesky_solve_after(int uplo, int orient, const double* A, i64 n, i64 lda, double* B,
                             i64 nrhs, i64 ldb) {
    (void)orient;
    return cholesky_solve_after(uplo, A, n, lda, B, nrhs, ldb);
}
int orc_scholesky_solve_after(int uplo, int orient, const float* A, i64 n, i64 lda, float* B,
                              i64 nrhs, i64 ldb) {
    (void)orient;
    return cholesky_solve_after(uplo, A, n, lda, B, nrhs, ldb);
}

int orc_hpd_solve(int uplo, int orient, const double* A, i64 n, i64 lda, double* B, i64 nrhs,
                  i64 ldb, i64 bsize) {
    (void)orient;
    return hpd_solve(uplo, A, n, lda, B, nrhs, ldb, bsize);
}
int orc_shpd_solve(int uplo, int orient, const float* A, i64 n, i64 lda, float* B, i64 nrhs,
                   i64 ldb, i64 bsize) {
    (void)orient;
    return hpd_solve(uplo, A, n, lda, B, nrhs, ldb, bsize);
}

// Permutation::ExplicitVector (Permutation.cpp:731-745): p with (PA)(i,:) = A(p(i),:).
void orc_perm_explicit(const int32_t* dests, i64 nswaps, i64 size, int32_t* p) {
    for (i64 i = 0; i < size; ++i) p[i] = (int32_t)i;
    for (i64 j = 0; j < nswaps; ++j)
        if (dests[j] != j) std::swap(p[j], p[dests[j]]);
}

// Parity of the swap sequence (Permutation.cpp:213-214): number of non-trivial swaps mod 2.
int orc_perm_parity(const int32_t* dests, i64 nswaps) {
    int par = 0;
    for (i64 j = 0; j < nswaps; ++j)
        if (dests[j] != j) par ^= 1;
    return par;
}

// ---------------------------------------------------------------------------------------
// Double-double residuals and error measures
// ---------------------------------------------------------------------------------------

// R = B - op(A) X with every product and sum carried in double-double; (Rhi + Rlo) is the
// residual to about 2^-104 relative to the terms. A is m x k (op = N) ; X k x nrhs.
void orc_dd_residual(i64 m, i64 k, const double* A, i64 lda, const double* X, i64 ldx,
                     i64 nrhs, const double* B, i64 ldb, double* Rhi, double* Rlo, i64 ldr) {
    for (i64 c = 0; c < nrhs; ++c) {
        const double* x = X + c * ldx;
        const double* b = B + c * ldb;
        double* rh = Rhi + c * ldr;
        double* rl = Rlo + c * ldr;
        const i64 blk = 256;
        const i64 nblk = (m + blk - 1) / blk;
#pragma omp parallel for schedule(static)
        for (i64 bi = 0; bi < nblk; ++bi) {
            const i64 i0 = bi * blk, i1 = std::min(m, i0 + blk);
            std::vector<orc::dd> acc((size_t)(i1 - i0));
            for (i64 i = i0; i < i1; ++i) acc[i - i0] = orc::dd_from(b[i]);
            for (i64 j = 0; j < k; ++j) {
                const double xj = x[j];
                const double* a = A + j * lda;
                for (i64 i = i0; i < i1; ++i)
                    acc[i - i0] = orc::dd_sub(acc[i - i0], orc::two_prod(a[i], xj));
            }
            for (i64 i = i0; i < i1; ++i) {
                rh[i] = acc[i - i0].hi;
                rl[i] = acc[i - i0].lo;
            }
        }
    }
}

// Normwise backward error of each column (Rigal-Gaches, infinity norms):
//   eta_c = ||b_c - A x_c||_inf / (||A||_inf ||x_c||_inf + ||b_c||_inf)
// with the residual in double-double. Returns max_c eta_c; per-column values in eta (may be
// null).
double orc_backward_error(i64 n, const double* A, i64 lda, const double* X, i64 ldx, i64 nrhs,
                          const double* B, i64 ldb, double* eta) {
    std::vector<double> rh((size_t)(n * nrhs)), rl((size_t)(n * nrhs));
    orc_dd_residual(n, n, A, lda, X, ldx, nrhs, B, ldb, rh.data(), rl.data(), n);
    std::vector<double> rows((size_t)n, 0.0);
    for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < n; ++i) rows[i] += std::fabs(A[IDX(i, j, lda)]);
    double anorm = 0.0;
    for (i64 i = 0; i < n; ++i) anorm = std::max(anorm, rows[i]);
    double worst = 0.0;
    for (i64 c = 0; c < nrhs; ++c) {
        double rn = 0.0, xn = 0.0, bn = 0.0;
        for (i64 i = 0; i < n; ++i) {
            rn = std::max(rn, std::fabs(rh[c * n + i] + rl[c * n + i]));
            xn = std::max(xn, std::fabs(X[IDX(i, c, ldx)]));
            bn = std::max(bn, std::fabs(B[IDX(i, c, ldb)]));
        }
        const double den = anorm * xn + bn;
        const double e = den > 0 ? rn / den : (rn > 0 ? INFINITY : 0.0);
        if (eta) eta[c] = e;
        worst = std::max(worst, e);
    }
    return worst;
}

// ||P A - L U||_inf / ||A||_inf for a packed LU factor (unit L strictly below the diagonal,
// U on and above), with each entry of L U accumulated in double-double. O(n^3): small n.
double orc_lu_factor_error(i64 n, const double* A0, i64 lda0, const double* F, i64 ldf,
                           const int32_t* dests, i64 nswaps) {
    std::vector<int32_t> p((size_t)n);
    orc_perm_explicit(dests, nswaps, n, p.data());
    double num = 0.0, anorm = 0.0;
    for (i64 i = 0; i < n; ++i) {
        double rowerr = 0.0, rowa = 0.0;
        for (i64 j = 0; j < n; ++j) {
            const double a = A0[IDX(p[i], j, lda0)];
            orc::dd acc = orc::dd_from(a);
            const i64 kmax = std::min(i, j);
            for (i64 k = 0; k <= kmax; ++k) {
                const double l = (k == i) ? 1.0 : F[IDX(i, k, ldf)];
                const double u = F[IDX(k, j, ldf)];
                acc = orc::dd_sub(acc, orc::two_prod(l, u));
            }
            rowerr += std::fabs(acc.hi + acc.lo);
            rowa += std::fabs(a);
        }
        num = std::max(num, rowerr);
        anorm = std::max(anorm, rowa);
    }
    return anorm > 0 ? num / anorm : num;
}

// ||A - L L^T||_inf / ||A||_inf (uplo 0) or ||A - U^T U|| (uplo 1) in double-double.
double orc_chol_factor_error(int uplo, i64 n, const double* A0, i64 lda0, const double* F,
                             i64 ldf) {
    double num = 0.0, anorm = 0.0;
    for (i64 i = 0; i < n; ++i) {
        double rowerr = 0.0, rowa = 0.0;
        for (i64 j = 0; j < n; ++j) {
            const double a = (uplo == 0) ? ((i >= j) ? A0[IDX(i, j, lda0)] : A0[IDX(j, i, lda0)])
                                         : ((i <= j) ? A0[IDX(i, j, lda0)] : A0[IDX(j, i, lda0)]);
            orc::dd acc = orc::dd_from(a);
            const i64 kmax = std::min(i, j);
            for (i64 k = 0; k <= kmax; ++k) {
                const double fi = (uplo == 0) ? F[IDX(i, k, ldf)] : F[IDX(k, i, ldf)];
                const double fj = (uplo == 0) ? F[IDX(j, k, ldf)] : F[IDX(k, j, ldf)];
                acc = orc::dd_sub(acc, orc::two_prod(fi, fj));
            }
            rowerr += std::fabs(acc.hi + acc.lo);
            rowa += std::fabs(a);
        }
        num = std::max(num, rowerr);
        anorm = std::max(anorm, rowa);
    }
    return anorm > 0 ? num / anorm : num;
}

// Exposes the double-double primitives so tests can pin them against exact rationals.
void orc_dd_two_prod(double a, double b, double* hi, double* lo) {
    orc::dd r = orc::two_prod(a, b);
    *hi = r.hi;
    *lo = r.lo;
}
void orc_dd_add(double ahi, double alo, double bhi, double blo, double* hi, double* lo) {
    orc::dd r = orc::dd_add({ahi, alo}, {bhi, blo});
    *hi = r.hi;
    *lo = r.lo;
}

// Direct BLAS entry points, for cross-checks of the backends.
void orc_dgemm(char ta, char tb, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda,
               const double* B, i64 ldb, double beta, double* C, i64 ldc) {
    g_blas->dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
void orc_dtrsm(char side, char uplo, char trans, char diag, i64 m, i64 n, double alpha,
               const double* A, i64 lda, double* B, i64 ldb) {
    g_blas->dtrsm(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}
void orc_dsyrk(char uplo, char trans, i64 n, i64 k, double alpha, const double* A, i64 lda,
               double beta, double* C, i64 ldc) {
    g_blas->dsyrk(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}
void orc_sgemm(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda,
               const float* B, i64 ldb, float beta, float* C, i64 ldc) {
    Ops<float>::gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
void orc_strsm(char side, char uplo, char trans, char diag, i64 m, i64 n, float alpha,
               const float* A, i64 lda, float* B, i64 ldb) {
    Ops<float>::trsm(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
}
void orc_ssyrk(char uplo, char trans, i64 n, i64 k, float alpha, const float* A, i64 lda,
               float beta, float* C, i64 ldc) {
    Ops<float>::syrk(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

// Host LAPACK dgetrf / dpotrf of the loaded backend (library cross-check and secondary CPU
// baseline). Returns -100 when the backend has no LAPACK. ipiv is 0-based on return.
int orc_host_dgetrf(i64 n, double* A, i64 lda, int32_t* ipiv) {
    if (g_blas != &g_host || !g_fb.dgetrf_) return -100;
    fint nn = (fint)n, la = (fint)lda, info = 0;
    std::vector<fint> p((size_t)n);
    g_fb.dgetrf_(&nn, &nn, A, &la, p.data(), &info);
    for (i64 i = 0; i < n; ++i) ipiv[i] = (int32_t)(p[i] - 1);
    return (int)info;
}
int orc_host_dpotrf(int uplo, i64 n, double* A, i64 lda) {
    if (g_blas != &g_host || !g_fb.dpotrf_) return -100;
    char u = uplo == 0 ? 'L' : 'U';
    fint nn = (fint)n, la = (fint)lda, info = 0;
    g_fb.dpotrf_(&u, &nn, A, &la, &info);
    return (int)info;
}

}  // extern "C"
