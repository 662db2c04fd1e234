// Elemental's C++ and C interfaces for the dense factor-and-solve path, on top of the device
// ABI (elemental_amd.h). Host matrices are staged in HBM per call; results, pivots and
// failures come back with the reference's types and exceptions:
//   El::LU                    src/lapack_like/factor/LU.cpp:104-142
//   El::lu::SolveAfter        src/lapack_like/factor/LU/SolveAfter.hpp:66-92
//   El::LinearSolve           src/lapack_like/solve/Linear.cpp:212-218
//   El::Cholesky              src/lapack_like/factor/Cholesky.cpp:26-38
//   El::cholesky::SolveAfter  src/lapack_like/factor/Cholesky/SolveAfter.hpp:15-43
//   El::HPDSolve              src/lapack_like/solve/HPD.cpp:47-57
//   C wrappers + EL_TRY       src/lapack_like/factor-C.cpp:70-135,268-323,
//                             src/lapack_like/solve-C.cpp:18-46, include/El/core/CReflect.hpp:12-26
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/El.h"
#include "../../include/El.hpp"
#include "elx_common.hpp"

namespace El {

const Range ALL(0, END);

void LogicError(const std::string& msg) { throw std::logic_error(msg); }
void RuntimeError(const std::string& msg) { throw std::runtime_error(msg); }

// ------------------------------------------------------------------ Matrix<T>
template <typename T>
Matrix<T>::Matrix(const Matrix<T>& A) : Matrix() {
    Resize(A.Height(), A.Width());
    for (Int j = 0; j < width_; ++j)
        std::memcpy(buf_ + (int64_t)j * ldim_, A.LockedBuffer(0, j), sizeof(T) * height_);
}

template <typename T>
Matrix<T>& Matrix<T>::operator=(const Matrix<T>& A) {
    if (this == &A) return *this;
    if (viewing_ && (height_ != A.Height() || width_ != A.Width()))
        LogicError("Cannot resize a view in assignment");
    if (!viewing_) Resize(A.Height(), A.Width());
    for (Int j = 0; j < width_; ++j)
        std::memcpy(Buffer() + (int64_t)j * ldim_, A.LockedBuffer(0, j), sizeof(T) * height_);
    return *this;
}

template <typename T>
void Matrix<T>::Empty() {
    mem_.clear();
    buf_ = nullptr;
    height_ = width_ = 0;
    ldim_ = 1;
    viewing_ = locked_ = false;
}

template <typename T>
void Matrix<T>::Resize(Int height, Int width) {
    Resize(height, width, std::max(height, 1));
}

template <typename T>
void Matrix<T>::Resize(Int height, Int width, Int ldim) {
    if (height < 0 || width < 0) LogicError("Height and width must be non-negative");
    if (ldim < std::max(height, 1)) LogicError("Leading dimension must be >= max(height,1)");
    if (viewing_) {
        if (height != height_ || width != width_) LogicError("Cannot resize a view");
        return;
    }
    height_ = height;
    width_ = width;
    ldim_ = ldim;
    mem_.assign((size_t)ldim * (size_t)width, T(0));
    buf_ = mem_.empty() ? nullptr : mem_.data();
}

template <typename T>
void Matrix<T>::Attach(Int height, Int width, T* buffer, Int ldim) {
    if (ldim < std::max(height, 1)) LogicError("Leading dimension must be >= max(height,1)");
    mem_.clear();
    height_ = height;
    width_ = width;
    ldim_ = ldim;
    buf_ = buffer;
    viewing_ = true;
    locked_ = false;
}

template <typename T>
void Matrix<T>::LockedAttach(Int height, Int width, const T* buffer, Int ldim) {
    Attach(height, width, const_cast<T*>(buffer), ldim);
    locked_ = true;
}

template <typename T>
Matrix<T> Matrix<T>::operator()(Range I, Range J) {
    const Int i1 = (I.end == END) ? height_ : I.end, j1 = (J.end == END) ? width_ : J.end;
    if (I.beg < 0 || i1 > height_ || I.beg > i1 || J.beg < 0 || j1 > width_ || J.beg > j1)
        LogicError("View out of bounds");
    Matrix<T> V;
    V.Attach(i1 - I.beg, j1 - J.beg, buf_ + I.beg + (int64_t)J.beg * ldim_, ldim_);
    V.locked_ = locked_;
    return V;
}

template <typename T>
const Matrix<T> Matrix<T>::operator()(Range I, Range J) const {
    return const_cast<Matrix<T>*>(this)->operator()(I, J);
}

template class Matrix<double>;
template class Matrix<float>;
template class Matrix<Int>;

// ------------------------------------------------------------------ Permutation
void Permutation::Empty() {
    size_ = 0;
    parity_ = false;
    dests_.clear();
}
void Permutation::MakeIdentity(Int size) {
    size_ = size;
    parity_ = false;
    dests_.clear();
}
void Permutation::ReserveSwaps(Int maxSwaps) { dests_.reserve(maxSwaps); }
void Permutation::Swap(Int origin, Int dest) {
    if (origin < 0 || origin >= size_ || dest < 0 || dest >= size_)
        LogicError("Attempted swap out of range");
    if (origin != (Int)dests_.size())
        LogicError("Only swap sequences with origins 0,1,2,... are supported");
    if (origin != dest) parity_ = !parity_;
    dests_.push_back(dest);
}
void Permutation::SetSwapDestinations(Int size, const std::vector<Int>& dests) {
    size_ = size;
    dests_ = dests;
    parity_ = false;
    for (size_t j = 0; j < dests_.size(); ++j)
        if (dests_[j] != (Int)j) parity_ = !parity_;
}
void Permutation::ExplicitVector(Matrix<Int>& p) const {
    p.Resize(size_, 1);
    for (Int i = 0; i < size_; ++i) p(i, 0) = i;
    for (size_t j = 0; j < dests_.size(); ++j)
        if (dests_[j] != (Int)j) std::swap(p((Int)j, 0), p(dests_[j], 0));
}
void Permutation::ExplicitMatrix(Matrix<Int>& P) const {
    Matrix<Int> p;
    ExplicitVector(p);
    P.Resize(size_, size_);
    for (Int i = 0; i < size_; ++i) P(i, p(i, 0)) = 1;
}

// ------------------------------------------------------------------ blocksize stack
static std::vector<Int> g_blocksizes{128};
Int Blocksize() { return g_blocksizes.back(); }
void SetBlocksize(Int b) { g_blocksizes.back() = b; }
void PushBlocksizeStack(Int b) { g_blocksizes.push_back(b); }
void PopBlocksizeStack() {
    if (g_blocksizes.size() > 1) g_blocksizes.pop_back();
}

// ------------------------------------------------------------------ device staging
namespace {

void check(elx_status_t s, const char* what) {
    if (s == ELX_SUCCESS) return;
    if (s == ELX_ERROR_ALLOC) throw std::bad_alloc();
    if (s == ELX_ERROR_INVALID_ARG) LogicError(std::string(what) + ": invalid argument");
    RuntimeError(std::string(what) + ": device error " + std::to_string((int)s));
}

void hcheck(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    RuntimeError(std::string(what) + ": " + hipGetErrorString(e));
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (bytes) hcheck(hipMalloc(&p, bytes), "hipMalloc");
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Device copy of a host matrix with leading dimension max(height,1).
template <typename T>
struct DevMatrix {
    Int m, n;
    int64_t ld;
    DevBuf buf;
    explicit DevMatrix(const Matrix<T>& A)
        : m(A.Height()), n(A.Width()), ld(std::max(A.Height(), 1)),
          buf(sizeof(T) * (size_t)ld * (size_t)A.Width()) {
        if (m > 0 && n > 0)
            hcheck(hipMemcpy2D(buf.p, sizeof(T) * ld, A.LockedBuffer(), sizeof(T) * A.LDim(),
                               sizeof(T) * m, n, hipMemcpyHostToDevice),
                   "hipMemcpy2D");
    }
    T* ptr() { return static_cast<T*>(buf.p); }
    void to_host(Matrix<T>& A) {
        if (m > 0 && n > 0)
            hcheck(hipMemcpy2D(A.Buffer(), sizeof(T) * A.LDim(), buf.p, sizeof(T) * ld,
                               sizeof(T) * m, n, hipMemcpyDeviceToHost),
                   "hipMemcpy2D");
    }
};

int read_info(const DevBuf& info) {
    int32_t v = 0;
    hcheck(hipMemcpy(&v, info.p, sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy");
    return v;
}

// The device ABI for each field (elemental_amd.h).
struct Dev {
    static elx_status_t getrf(int64_t m, int64_t n, double* A, int64_t ld, int64_t nb, int32_t* d,
                              int32_t* info) {
        return elx_dgetrf(nullptr, m, n, A, ld, nb, d, info);
    }
    static elx_status_t getrf(int64_t m, int64_t n, float* A, int64_t ld, int64_t nb, int32_t* d,
                              int32_t* info) {
        return elx_sgetrf(nullptr, m, n, A, ld, nb, d, info);
    }
    static elx_status_t getrf_nopiv(int64_t m, int64_t n, double* A, int64_t ld, int64_t nb,
                                    int32_t* info) {
        return elx_dgetrf_nopiv(nullptr, m, n, A, ld, nb, info);
    }
    static elx_status_t getrf_nopiv(int64_t m, int64_t n, float* A, int64_t ld, int64_t nb,
                                    int32_t* info) {
        return elx_sgetrf_nopiv(nullptr, m, n, A, ld, nb, info);
    }
    static elx_status_t getrf_full(int64_t m, int64_t n, double* A, int64_t ld, int32_t* p,
                                   int32_t* q, int32_t* info) {
        return elx_dgetrf_full(nullptr, m, n, A, ld, p, q, info);
    }
    static elx_status_t getrf_full(int64_t m, int64_t n, float* A, int64_t ld, int32_t* p,
                                   int32_t* q, int32_t* info) {
        return elx_sgetrf_full(nullptr, m, n, A, ld, p, q, info);
    }
    static elx_status_t getrs_full(int o, int64_t n, int64_t k, const double* A, int64_t lda,
                                   const int32_t* p, const int32_t* q, double* B, int64_t ldb) {
        return elx_dgetrs_full(nullptr, o, n, k, A, lda, p, q, B, ldb);
    }
    static elx_status_t getrs_full(int o, int64_t n, int64_t k, const float* A, int64_t lda,
                                   const int32_t* p, const int32_t* q, float* B, int64_t ldb) {
        return elx_sgetrs_full(nullptr, o, n, k, A, lda, p, q, B, ldb);
    }
    static elx_status_t getrs(int o, int64_t n, int64_t k, const double* A, int64_t lda,
                              const int32_t* d, double* B, int64_t ldb) {
        return elx_dgetrs(nullptr, o, n, k, A, lda, d, B, ldb);
    }
    static elx_status_t getrs(int o, int64_t n, int64_t k, const float* A, int64_t lda,
                              const int32_t* d, float* B, int64_t ldb) {
        return elx_sgetrs(nullptr, o, n, k, A, lda, d, B, ldb);
    }
    static elx_status_t gesv(int64_t n, int64_t k, double* A, int64_t lda, double* B, int64_t ldb,
                             int64_t nb, int32_t* info) {
        return elx_dgesv(nullptr, n, k, A, lda, B, ldb, nb, nullptr, info);
    }
    static elx_status_t gesv(int64_t n, int64_t k, float* A, int64_t lda, float* B, int64_t ldb,
                             int64_t nb, int32_t* info) {
        return elx_sgesv(nullptr, n, k, A, lda, B, ldb, nb, nullptr, info);
    }
    static elx_status_t potrf(int u, int64_t n, double* A, int64_t lda, int64_t nb, int32_t* info) {
        return elx_dpotrf(nullptr, u, n, A, lda, nb, info);
    }
    static elx_status_t potrf(int u, int64_t n, float* A, int64_t lda, int64_t nb, int32_t* info) {
        return elx_spotrf(nullptr, u, n, A, lda, nb, info);
    }
    static elx_status_t potrs(int u, int o, int64_t n, int64_t k, const double* A, int64_t lda,
                              double* B, int64_t ldb) {
        return elx_dpotrs(nullptr, u, o, n, k, A, lda, B, ldb);
    }
    static elx_status_t potrs(int u, int o, int64_t n, int64_t k, const float* A, int64_t lda,
                              float* B, int64_t ldb) {
        return elx_spotrs(nullptr, u, o, n, k, A, lda, B, ldb);
    }
    static elx_status_t posv(int u, int o, int64_t n, int64_t k, double* A, int64_t lda, double* B,
                             int64_t ldb, int64_t nb, int32_t* info) {
        return elx_dposv(nullptr, u, o, n, k, A, lda, B, ldb, nb, info);
    }
    static elx_status_t posv(int u, int o, int64_t n, int64_t k, float* A, int64_t lda, float* B,
                             int64_t ldb, int64_t nb, int32_t* info) {
        return elx_sposv(nullptr, u, o, n, k, A, lda, B, ldb, nb, info);
    }
    static elx_status_t gemm(int ta, int tb, int64_t m, int64_t n, int64_t k, double al,
                             const double* A, int64_t lda, const double* B, int64_t ldb, double be,
                             double* C, int64_t ldc) {
        return elx_dgemm(nullptr, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
    }
    static elx_status_t gemm(int ta, int tb, int64_t m, int64_t n, int64_t k, float al,
                             const float* A, int64_t lda, const float* B, int64_t ldb, float be,
                             float* C, int64_t ldc) {
        return elx_sgemm(nullptr, ta, tb, m, n, k, al, A, lda, B, ldb, be, C, ldc);
    }
    static elx_status_t trsm(int sd, int u, int o, int dg, int64_t m, int64_t n, double al,
                             const double* A, int64_t lda, double* B, int64_t ldb) {
        return elx_dtrsm(nullptr, sd, u, o, dg, m, n, al, A, lda, B, ldb);
    }
    static elx_status_t trsm(int sd, int u, int o, int dg, int64_t m, int64_t n, float al,
                             const float* A, int64_t lda, float* B, int64_t ldb) {
        return elx_strsm(nullptr, sd, u, o, dg, m, n, al, A, lda, B, ldb);
    }
    static elx_status_t syrk(int u, int o, int64_t n, int64_t k, double al, const double* A,
                             int64_t lda, double be, double* C, int64_t ldc) {
        return elx_dsyrk(nullptr, u, o, n, k, al, A, lda, be, C, ldc);
    }
    static elx_status_t syrk(int u, int o, int64_t n, int64_t k, float al, const float* A,
                             int64_t lda, float be, float* C, int64_t ldc) {
        return elx_ssyrk(nullptr, u, o, n, k, al, A, lda, be, C, ldc);
    }
    static elx_status_t permute(int inv, const int32_t* d, int64_t ns, int64_t m, int64_t n,
                                double* A, int64_t lda) {
        return elx_dpermute_rows(nullptr, inv, d, ns, m, n, A, lda);
    }
    static elx_status_t permute(int inv, const int32_t* d, int64_t ns, int64_t m, int64_t n,
                                float* A, int64_t lda) {
        return elx_spermute_rows(nullptr, inv, d, ns, m, n, A, lda);
    }
};

// Status of a device call as the reference's exception. Device-side numerical failures
// arrive through the info word instead (see read_info).
void check_status(elx_status_t st, const char* what) { check(st, what); }

Int blocksize_checked() {
    if (Blocksize() < 1) LogicError("Blocksize() must be positive");
    return Blocksize();
}

// The info word of a device factorization as the reference reports it: 0 success, j+1 the
// first failing column (the caller throws its exception), and -1 (the device panel gave up
// waiting for a peer workgroup) a runtime error.
int read_info_checked(const DevBuf& info, const char* what) {
    const int v = read_info(info);
    if (v < 0) RuntimeError(std::string(what) + ": device panel did not complete");
    return v;
}

}  // namespace

template <typename T>
void Permutation::PermuteRows(Matrix<T>& A, Int offset) const {
    if (A.Height() == 0 || A.Width() == 0 || dests_.empty()) return;
    // origins j+offset, destinations dests[j]+offset (Permutation.cpp:544-580): the swap
    // sequence applied to the rows from `offset` down
    if (offset < 0 || offset > A.Height()) LogicError("Permutation offset out of range");
    if ((Int)dests_.size() > A.Height() - offset) LogicError("Permutation larger than matrix");
    for (Int d : dests_)
        if (d < 0 || d >= A.Height() - offset) LogicError("Swap destination out of range");
    DevMatrix<T> dA(A);
    DevBuf dd(sizeof(int32_t) * dests_.size());
    hcheck(hipMemcpy(dd.p, dests_.data(), sizeof(int32_t) * dests_.size(), hipMemcpyHostToDevice), "hipMemcpy");
    check(Dev::permute(0, static_cast<int32_t*>(dd.p), (int64_t)dests_.size(), dA.m - offset, dA.n,
                       dA.ptr() + offset, dA.ld),
          "PermuteRows");
    hcheck(hipDeviceSynchronize(), "sync");
    dA.to_host(A);
}

template <typename T>
void Permutation::InversePermuteRows(Matrix<T>& A, Int offset) const {
    if (A.Height() == 0 || A.Width() == 0 || dests_.empty()) return;
    if (offset < 0 || offset > A.Height()) LogicError("Permutation offset out of range");
    if ((Int)dests_.size() > A.Height() - offset) LogicError("Permutation larger than matrix");
    for (Int d : dests_)
        if (d < 0 || d >= A.Height() - offset) LogicError("Swap destination out of range");
    DevMatrix<T> dA(A);
    DevBuf dd(sizeof(int32_t) * dests_.size());
    hcheck(hipMemcpy(dd.p, dests_.data(), sizeof(int32_t) * dests_.size(), hipMemcpyHostToDevice), "hipMemcpy");
    check(Dev::permute(1, static_cast<int32_t*>(dd.p), (int64_t)dests_.size(), dA.m - offset, dA.n,
                       dA.ptr() + offset, dA.ld),
          "InversePermuteRows");
    hcheck(hipDeviceSynchronize(), "sync");
    dA.to_host(A);
}
template void Permutation::PermuteRows(Matrix<double>&, Int) const;
template void Permutation::InversePermuteRows(Matrix<double>&, Int) const;
template void Permutation::PermuteRows(Matrix<float>&, Int) const;
template void Permutation::InversePermuteRows(Matrix<float>&, Int) const;

// ------------------------------------------------------------------ level 3
template <typename T>
void Gemm(Orientation oA, Orientation oB, T alpha, const Matrix<T>& A, const Matrix<T>& B, T beta,
          Matrix<T>& C) {
    const Int m = C.Height(), n = C.Width();
    const Int k = (oA == NORMAL) ? A.Width() : A.Height();
    if (((oA == NORMAL) ? A.Height() : A.Width()) != m || ((oB == NORMAL) ? B.Width() : B.Height()) != n ||
        ((oB == NORMAL) ? B.Height() : B.Width()) != k)
        LogicError("Nonconformal Gemm");
    DevMatrix<T> dA(A), dB(B), dC(C);
    check(Dev::gemm(oA, oB, m, n, k, alpha, dA.ptr(), dA.ld, dB.ptr(), dB.ld, beta, dC.ptr(), dC.ld), "Gemm");
    hcheck(hipDeviceSynchronize(), "sync");
    dC.to_host(C);
}

template <typename F>
void Trsm(LeftOrRight side, UpperOrLower uplo, Orientation orientation, UnitOrNonUnit diag,
          F alpha, const Matrix<F>& A, Matrix<F>& B, bool checkIfSingular) {
    if (A.Height() != A.Width()) LogicError("Triangular matrix must be square");
    if ((side == LEFT ? B.Height() : B.Width()) != A.Height()) LogicError("Nonconformal Trsm");
    if (checkIfSingular && diag != UNIT)
        for (Int j = 0; j < A.Height(); ++j)
            if (A.Get(j, j) == F(0)) throw SingularMatrixException();
    DevMatrix<F> dA(A), dB(B);
    check(Dev::trsm(side, uplo, orientation, diag, B.Height(), B.Width(), alpha, dA.ptr(), dA.ld,
                    dB.ptr(), dB.ld),
          "Trsm");
    hcheck(hipDeviceSynchronize(), "sync");
    dB.to_host(B);
}

template <typename T>
void Syrk(UpperOrLower uplo, Orientation orientation, T alpha, const Matrix<T>& A, T beta,
          Matrix<T>& C, bool) {
    const Int n = C.Height();
    const Int k = (orientation == NORMAL) ? A.Width() : A.Height();
    if (C.Width() != n || ((orientation == NORMAL) ? A.Height() : A.Width()) != n)
        LogicError("Nonconformal Syrk");
    DevMatrix<T> dA(A), dC(C);
    check(Dev::syrk(uplo, orientation, n, k, alpha, dA.ptr(), dA.ld, beta, dC.ptr(), dC.ld), "Syrk");
    hcheck(hipDeviceSynchronize(), "sync");
    dC.to_host(C);
}

template <typename T>
void Herk(UpperOrLower uplo, Orientation orientation, T alpha, const Matrix<T>& A, T beta,
          Matrix<T>& C) {
    Syrk(uplo, orientation, alpha, A, beta, C, true);
}

// ------------------------------------------------------------------ factorizations
template <typename F>
void LU(Matrix<F>& A) {
    const Int nb = blocksize_checked();
    const Int m = A.Height(), n = A.Width();
    if (std::min(m, n) == 0) return;
    DevMatrix<F> dA(A);
    DevBuf info(sizeof(int32_t));
    check(Dev::getrf_nopiv(m, n, dA.ptr(), dA.ld, nb, static_cast<int32_t*>(info.p)), "LU");
    hcheck(hipDeviceSynchronize(), "sync");
    const int inf = read_info_checked(info, "LU");
    dA.to_host(A);  // partially overwritten when the factorization stopped
    if (inf > 0) throw SingularMatrixException();
}

template <typename F>
void LU(Matrix<F>& A, Permutation& P) {
    const Int nb = blocksize_checked();
    const Int m = A.Height(), n = A.Width(), minDim = std::min(m, n);
    P.MakeIdentity(m);
    if (minDim == 0) return;
    DevMatrix<F> dA(A);
    DevBuf dests(sizeof(int32_t) * minDim), info(sizeof(int32_t));
    check(Dev::getrf(m, n, dA.ptr(), dA.ld, nb, static_cast<int32_t*>(dests.p),
                     static_cast<int32_t*>(info.p)),
          "LU");
    hcheck(hipDeviceSynchronize(), "sync");
    const int inf = read_info_checked(info, "LU");
    std::vector<Int> d(minDim);
    hcheck(hipMemcpy(d.data(), dests.p, sizeof(int32_t) * minDim, hipMemcpyDeviceToHost), "hipMemcpy");
    dA.to_host(A);
    if (inf > 0) {
        // the swaps recorded before the throw: those of every step up to the failing column
        d.resize(inf);
        P.SetSwapDestinations(m, d);
        throw SingularMatrixException();
    }
    P.SetSwapDestinations(m, d);
}

// El::LU(A, P, Q): full pivoting (LU.cpp:144-152, LU/Full.hpp:15-58).
template <typename F>
void LU(Matrix<F>& A, Permutation& P, Permutation& Q) {
    const Int m = A.Height(), n = A.Width(), minDim = std::min(m, n);
    P.MakeIdentity(m);
    Q.MakeIdentity(n);
    if (minDim == 0) return;
    DevMatrix<F> dA(A);
    DevBuf dp(sizeof(int32_t) * minDim), dq(sizeof(int32_t) * minDim), info(sizeof(int32_t));
    check(Dev::getrf_full(m, n, dA.ptr(), dA.ld, static_cast<int32_t*>(dp.p),
                          static_cast<int32_t*>(dq.p), static_cast<int32_t*>(info.p)),
          "LU");
    hcheck(hipDeviceSynchronize(), "sync");
    const int inf = read_info_checked(info, "LU");
    std::vector<Int> p(minDim), q(minDim);
    hcheck(hipMemcpy(p.data(), dp.p, sizeof(int32_t) * minDim, hipMemcpyDeviceToHost), "hipMemcpy");
    hcheck(hipMemcpy(q.data(), dq.p, sizeof(int32_t) * minDim, hipMemcpyDeviceToHost), "hipMemcpy");
    dA.to_host(A);
    const Int nsw = inf > 0 ? inf : minDim;  // the swaps made before a throw
    p.resize(nsw);
    q.resize(nsw);
    P.SetSwapDestinations(m, p);
    Q.SetSwapDestinations(n, q);
    if (inf > 0) throw SingularMatrixException();
}

namespace lu {
template <typename F>
void SolveAfter(Orientation orientation, const Matrix<F>& A, const Permutation& P,
                const Permutation& Q, Matrix<F>& B) {
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (A.Height() != B.Height()) LogicError("A and B must be the same height");
    const Int n = A.Height();
    if (n == 0 || B.Width() == 0) return;
    auto padded = [&](const Permutation& X) {
        const std::vector<Int>& d = X.SwapDestinations();
        if ((Int)d.size() > n) LogicError("Permutation larger than A");
        std::vector<Int> full(n);
        for (Int i = 0; i < n; ++i) full[i] = (i < (Int)d.size()) ? d[i] : i;
        return full;
    };
    const std::vector<Int> p = padded(P), q = padded(Q);
    DevMatrix<F> dA(A), dB(B);
    DevBuf dp(sizeof(int32_t) * n), dq(sizeof(int32_t) * n);
    hcheck(hipMemcpy(dp.p, p.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hcheck(hipMemcpy(dq.p, q.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice), "hipMemcpy");
    check(Dev::getrs_full(orientation, n, B.Width(), dA.ptr(), dA.ld, static_cast<int32_t*>(dp.p),
                          static_cast<int32_t*>(dq.p), dB.ptr(), dB.ld),
          "lu::SolveAfter");
    hcheck(hipDeviceSynchronize(), "sync");
    dB.to_host(B);
}

template <typename F>
void SolveAfter(Orientation orientation, const Matrix<F>& A, Matrix<F>& B) {
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (A.Height() != B.Height()) LogicError("A and B must be the same height");
    const Int n = A.Height();
    if (n == 0 || B.Width() == 0) return;
    DevMatrix<F> dA(A), dB(B);
    check(Dev::getrs(orientation, n, B.Width(), dA.ptr(), dA.ld, nullptr, dB.ptr(), dB.ld),
          "lu::SolveAfter");
    hcheck(hipDeviceSynchronize(), "sync");
    dB.to_host(B);
}

template <typename F>
void SolveAfter(Orientation orientation, const Matrix<F>& A, const Permutation& P, Matrix<F>& B) {
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (A.Height() != B.Height()) LogicError("A and B must be the same height");
    const Int n = A.Height();
    if (n == 0 || B.Width() == 0) return;
    const std::vector<Int>& d = P.SwapDestinations();
    if ((Int)d.size() > n) LogicError("Permutation larger than A");
    // A partial swap sequence is padded with identity swaps.
    std::vector<Int> full(n);
    for (Int i = 0; i < n; ++i) full[i] = (i < (Int)d.size()) ? d[i] : i;
    DevMatrix<F> dA(A), dB(B);
    DevBuf df(sizeof(int32_t) * n);
    hcheck(hipMemcpy(df.p, full.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice), "hipMemcpy");
    check(Dev::getrs(orientation, n, B.Width(), dA.ptr(), dA.ld, static_cast<int32_t*>(df.p),
                     dB.ptr(), dB.ld),
          "lu::SolveAfter");
    hcheck(hipDeviceSynchronize(), "sync");
    dB.to_host(B);
}
}  // namespace lu

template <typename F>
void Cholesky(UpperOrLower uplo, Matrix<F>& A) {
    const Int nb = blocksize_checked();
    if (A.Height() != A.Width()) LogicError("A must be square");
    const Int n = A.Height();
    if (n == 0) return;
    DevMatrix<F> dA(A);
    DevBuf info(sizeof(int32_t));
    check(Dev::potrf(uplo, n, dA.ptr(), dA.ld, nb, static_cast<int32_t*>(info.p)), "Cholesky");
    hcheck(hipDeviceSynchronize(), "sync");
    const int inf = read_info_checked(info, "Cholesky");
    dA.to_host(A);
    if (inf > 0) throw NonHPDMatrixException("A was not numerically HPD");
}

namespace cholesky {
template <typename F>
void SolveAfter(UpperOrLower uplo, Orientation orientation, const Matrix<F>& A, Matrix<F>& B) {
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (A.Height() != B.Height()) LogicError("A and B must be the same height");
    if (A.Height() == 0 || B.Width() == 0) return;
    DevMatrix<F> dA(A), dB(B);
    check(Dev::potrs(uplo, orientation, A.Height(), B.Width(), dA.ptr(), dA.ld, dB.ptr(), dB.ld),
          "cholesky::SolveAfter");
    hcheck(hipDeviceSynchronize(), "sync");
    dB.to_host(B);
}
}  // namespace cholesky

template <typename F>
void LinearSolve(const Matrix<F>& A, Matrix<F>& B) {
    const Int nb = blocksize_checked();
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (A.Height() != B.Height()) LogicError("A and B must be the same height");
    const Int n = A.Height();
    if (n == 0) return;
    DevMatrix<F> dA(A), dB(B);  // the device copy plays the role of ACopy
    DevBuf info(sizeof(int32_t));
    check(Dev::gesv(n, B.Width(), dA.ptr(), dA.ld, dB.ptr(), dB.ld, nb,
                    static_cast<int32_t*>(info.p)),
          "LinearSolve");
    hcheck(hipDeviceSynchronize(), "sync");
    const int inf = read_info_checked(info, "LinearSolve");
    // B is partially updated when the elimination stopped, as in the reference
    dB.to_host(B);
    if (inf > 0) throw SingularMatrixException();
}

template <typename F>
void HPDSolve(UpperOrLower uplo, Orientation orientation, const Matrix<F>& A, Matrix<F>& B) {
    const Int nb = blocksize_checked();
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (A.Height() != B.Height()) LogicError("A and B must be the same height");
    const Int n = A.Height();
    if (n == 0) return;
    DevMatrix<F> dA(A), dB(B);
    DevBuf info(sizeof(int32_t));
    check(Dev::posv(uplo, orientation, n, B.Width(), dA.ptr(), dA.ld, dB.ptr(), dB.ld, nb,
                    static_cast<int32_t*>(info.p)),
          "HPDSolve");
    hcheck(hipDeviceSynchronize(), "sync");
    if (read_info_checked(info, "HPDSolve") > 0) throw NonHPDMatrixException("A was not numerically HPD");
    dB.to_host(B);
}

// Explicit instantiations, as the reference's PROTO macros provide (LU.cpp:233-313,
// Cholesky.cpp:206-266, Linear.cpp:320-343, HPD.cpp:109-133, Gemm.cpp:261-298,
// Trsm.cpp:400-435, Syrk.cpp, Herk.cpp).
#define EL_AMD_PROTO(F)                                                                        \
    template void Gemm(Orientation, Orientation, F, const Matrix<F>&, const Matrix<F>&, F,       \
                       Matrix<F>&);                                                            \
    template void Trsm(LeftOrRight, UpperOrLower, Orientation, UnitOrNonUnit, F,               \
                       const Matrix<F>&, Matrix<F>&, bool);                                    \
    template void Syrk(UpperOrLower, Orientation, F, const Matrix<F>&, F, Matrix<F>&, bool);  \
    template void Herk(UpperOrLower, Orientation, F, const Matrix<F>&, F, Matrix<F>&);        \
    template void LU(Matrix<F>&);                                                              \
    template void LU(Matrix<F>&, Permutation&);                                                \
    template void LU(Matrix<F>&, Permutation&, Permutation&);                                  \
    template void lu::SolveAfter(Orientation, const Matrix<F>&, const Permutation&,            \
                                 const Permutation&, Matrix<F>&);                              \
    template void lu::SolveAfter(Orientation, const Matrix<F>&, Matrix<F>&);                   \
    template void lu::SolveAfter(Orientation, const Matrix<F>&, const Permutation&, Matrix<F>&); \
    template void Cholesky(UpperOrLower, Matrix<F>&);                                          \
    template void cholesky::SolveAfter(UpperOrLower, Orientation, const Matrix<F>&, Matrix<F>&); \
    template void LinearSolve(const Matrix<F>&, Matrix<F>&);                                   \
    template void HPDSolve(UpperOrLower, Orientation, const Matrix<F>&, Matrix<F>&);
EL_AMD_PROTO(double)
EL_AMD_PROTO(float)
#undef EL_AMD_PROTO

// El::RuizEquil (src/lapack_like/equilibrate/RuizEquil.cpp:23-60): A becomes
// diag(dRow)^{-1} A diag(dCol)^{-1}; dRow, dCol are resized to m x 1 and n x 1.
void RuizEquil(Matrix<double>& A, Matrix<double>& dRow, Matrix<double>& dCol, bool progress) {
    (void)progress;
    const Int m = A.Height(), n = A.Width();
    dRow.Resize(m, 1);
    dCol.Resize(n, 1);
    DevMatrix<double> dA(A), dR(dRow), dC(dCol);
    check(elx_druiz_equil(nullptr, m, n, dA.ptr(), dA.ld, dR.ptr(), dC.ptr()), "RuizEquil");
    hcheck(hipDeviceSynchronize(), "sync");
    dA.to_host(A);
    dR.to_host(dRow);
    dC.to_host(dCol);
}

// El::GeomEquil (src/lapack_like/equilibrate/GeomEquil.cpp:26-116): A becomes
// diag(dRow)^{-1} A diag(dCol)^{-1}; dRow, dCol are resized to m x 1 and n x 1.
void GeomEquil(Matrix<double>& A, Matrix<double>& dRow, Matrix<double>& dCol, bool progress) {
    (void)progress;
    const Int m = A.Height(), n = A.Width();
    dRow.Resize(m, 1);
    dCol.Resize(n, 1);
    DevMatrix<double> dA(A), dR(dRow), dC(dCol);
    check(elx_dgeom_equil(nullptr, m, n, dA.ptr(), dA.ld, dR.ptr(), dC.ptr()), "GeomEquil");
    hcheck(hipDeviceSynchronize(), "sync");
    dA.to_host(A);
    dR.to_host(dRow);
    dC.to_host(dCol);
}

// El::SymmetricRuizEquil (src/lapack_like/equilibrate/SymmetricRuizEquil.cpp:29-55): A becomes
// diag(d)^{-1} A diag(d)^{-1}; d is resized to n x 1.
void SymmetricRuizEquil(Matrix<double>& A, Matrix<double>& d, Int maxiter, bool progress) {
    (void)progress;
    if (A.Height() != A.Width()) LogicError("A must be square");
    if (maxiter < 0) LogicError("maxiter must be non-negative");
    const Int n = A.Height();
    d.Resize(n, 1);
    DevMatrix<double> dA(A), dD(d);
    check(elx_dsymmetric_ruiz_equil(nullptr, n, dA.ptr(), dA.ld, dD.ptr(), (int32_t)maxiter),
          "SymmetricRuizEquil");
    hcheck(hipDeviceSynchronize(), "sync");
    dA.to_host(A);
    dD.to_host(d);
}

}  // namespace El

// ====================================================================== C interface
using El::Int;
using El::Matrix;
using El::Permutation;

#define EL_CATCH                                                     \
    catch (std::bad_alloc & e) {                                     \
        std::fprintf(stderr, "[elemental_amd] %s\n", e.what());      \
        return EL_ALLOC_ERROR;                                       \
    }                                                                \
    catch (El::ArgException & e) {                                   \
        std::fprintf(stderr, "[elemental_amd] %s\n", e.what());      \
        return EL_ARG_ERROR;                                         \
    }                                                                \
    catch (std::logic_error & e) {                                   \
        std::fprintf(stderr, "[elemental_amd] %s\n", e.what());      \
        return EL_LOGIC_ERROR;                                       \
    }                                                                \
    catch (std::runtime_error & e) {                                 \
        std::fprintf(stderr, "[elemental_amd] %s\n", e.what());      \
        return EL_RUNTIME_ERROR;                                     \
    }                                                                \
    catch (std::exception & e) {                                     \
        std::fprintf(stderr, "[elemental_amd] %s\n", e.what());      \
        return EL_ERROR;                                             \
    }
#define EL_TRY(payload) \
    try {               \
        payload;        \
    }                   \
    EL_CATCH            \
    return EL_SUCCESS;

static Matrix<double>* M(ElMatrix_d A) { return reinterpret_cast<Matrix<double>*>(A); }
static const Matrix<double>* M(ElConstMatrix_d A) { return reinterpret_cast<const Matrix<double>*>(A); }
static Matrix<float>* M(ElMatrix_s A) { return reinterpret_cast<Matrix<float>*>(A); }
static const Matrix<float>* M(ElConstMatrix_s A) { return reinterpret_cast<const Matrix<float>*>(A); }
static Matrix<Int>* Mi(ElMatrix_i A) { return reinterpret_cast<Matrix<Int>*>(A); }
static const Matrix<Int>* Mi(ElConstMatrix_i A) { return reinterpret_cast<const Matrix<Int>*>(A); }
static Permutation* Pm(ElPermutation P) { return reinterpret_cast<Permutation*>(P); }
static const Permutation* Pm(ElConstPermutation P) { return reinterpret_cast<const Permutation*>(P); }
static El::Orientation O(ElOrientation o) { return static_cast<El::Orientation>(o); }
static El::UpperOrLower U(ElUpperOrLower u) { return static_cast<El::UpperOrLower>(u); }

extern "C" {

const char* ElErrorString(ElError error) {
    switch (error) {
        case EL_SUCCESS: return "EL_SUCCESS";
        case EL_ALLOC_ERROR: return "EL_ALLOC_ERROR";
        case EL_OUT_OF_BOUNDS_ERROR: return "EL_OUT_OF_BOUNDS_ERROR";
        case EL_ARG_ERROR: return "EL_ARG_ERROR";
        case EL_LOGIC_ERROR: return "EL_LOGIC_ERROR";
        case EL_RUNTIME_ERROR: return "EL_RUNTIME_ERROR";
        default: return "EL_ERROR";
    }
}

ElError ElMatrixCreate_d(ElMatrix_d* A) { EL_TRY(*A = reinterpret_cast<ElMatrix_d>(new Matrix<double>())) }
ElError ElMatrixDestroy_d(ElConstMatrix_d A) { EL_TRY(delete M(A)) }
ElError ElMatrixResize_d(ElMatrix_d A, ElInt h, ElInt w) { EL_TRY(M(A)->Resize(h, w)) }
ElError ElMatrixResizeWithLDim_d(ElMatrix_d A, ElInt h, ElInt w, ElInt ld) { EL_TRY(M(A)->Resize(h, w, ld)) }
ElError ElMatrixAttach_d(ElMatrix_d A, ElInt h, ElInt w, double* buf, ElInt ld) { EL_TRY(M(A)->Attach(h, w, buf, ld)) }
ElError ElMatrixHeight_d(ElConstMatrix_d A, ElInt* h) { EL_TRY(*h = M(A)->Height()) }
ElError ElMatrixWidth_d(ElConstMatrix_d A, ElInt* w) { EL_TRY(*w = M(A)->Width()) }
ElError ElMatrixLDim_d(ElConstMatrix_d A, ElInt* ld) { EL_TRY(*ld = M(A)->LDim()) }
ElError ElMatrixBuffer_d(ElMatrix_d A, double** b) { EL_TRY(*b = M(A)->Buffer()) }
ElError ElMatrixLockedBuffer_d(ElConstMatrix_d A, const double** b) { EL_TRY(*b = M(A)->LockedBuffer()) }

ElError ElMatrixCreate_s(ElMatrix_s* A) { EL_TRY(*A = reinterpret_cast<ElMatrix_s>(new Matrix<float>())) }
ElError ElMatrixDestroy_s(ElConstMatrix_s A) { EL_TRY(delete M(A)) }
ElError ElMatrixResize_s(ElMatrix_s A, ElInt h, ElInt w) { EL_TRY(M(A)->Resize(h, w)) }
ElError ElMatrixResizeWithLDim_s(ElMatrix_s A, ElInt h, ElInt w, ElInt ld) { EL_TRY(M(A)->Resize(h, w, ld)) }
ElError ElMatrixAttach_s(ElMatrix_s A, ElInt h, ElInt w, float* buf, ElInt ld) { EL_TRY(M(A)->Attach(h, w, buf, ld)) }
ElError ElMatrixHeight_s(ElConstMatrix_s A, ElInt* h) { EL_TRY(*h = M(A)->Height()) }
ElError ElMatrixWidth_s(ElConstMatrix_s A, ElInt* w) { EL_TRY(*w = M(A)->Width()) }
ElError ElMatrixLDim_s(ElConstMatrix_s A, ElInt* ld) { EL_TRY(*ld = M(A)->LDim()) }
ElError ElMatrixBuffer_s(ElMatrix_s A, float** b) { EL_TRY(*b = M(A)->Buffer()) }
ElError ElMatrixLockedBuffer_s(ElConstMatrix_s A, const float** b) { EL_TRY(*b = M(A)->LockedBuffer()) }

ElError ElMatrixCreate_i(ElMatrix_i* A) { EL_TRY(*A = reinterpret_cast<ElMatrix_i>(new Matrix<Int>())) }
ElError ElMatrixDestroy_i(ElConstMatrix_i A) { EL_TRY(delete Mi(A)) }
ElError ElMatrixHeight_i(ElConstMatrix_i A, ElInt* h) { EL_TRY(*h = Mi(A)->Height()) }
ElError ElMatrixLockedBuffer_i(ElConstMatrix_i A, const ElInt** b) { EL_TRY(*b = Mi(A)->LockedBuffer()) }

ElError ElPermutationCreate(ElPermutation* P) { EL_TRY(*P = reinterpret_cast<ElPermutation>(new Permutation())) }
ElError ElPermutationDestroy(ElConstPermutation P) { EL_TRY(delete Pm(P)) }
ElError ElPermutationMakeIdentity(ElPermutation P, ElInt size) { EL_TRY(Pm(P)->MakeIdentity(size)) }
ElError ElPermutationSwap(ElPermutation P, ElInt o, ElInt d) { EL_TRY(Pm(P)->Swap(o, d)) }
ElError ElPermutationHeight(ElConstPermutation P, ElInt* h) { EL_TRY(*h = Pm(P)->Height()) }
ElError ElPermutationParity(ElConstPermutation P, bool* par) { EL_TRY(*par = Pm(P)->Parity()) }
ElError ElPermutationSwapDestinations(ElConstPermutation P, ElInt* dests, ElInt* numSwaps) {
    EL_TRY({
        const auto& d = Pm(P)->SwapDestinations();
        *numSwaps = (ElInt)d.size();
        if (dests) std::copy(d.begin(), d.end(), dests);
    })
}
ElError ElPermutationExplicitVector(ElConstPermutation P, ElMatrix_i p) { EL_TRY(Pm(P)->ExplicitVector(*Mi(p))) }
ElError ElPermutationPermuteRows_d(ElConstPermutation P, ElMatrix_d A) { EL_TRY(Pm(P)->PermuteRows(*M(A))) }
ElError ElPermutationInversePermuteRows_d(ElConstPermutation P, ElMatrix_d A) {
    EL_TRY(Pm(P)->InversePermuteRows(*M(A)))
}
ElError ElPermutationPermuteRows_s(ElConstPermutation P, ElMatrix_s A) { EL_TRY(Pm(P)->PermuteRows(*M(A))) }
ElError ElPermutationInversePermuteRows_s(ElConstPermutation P, ElMatrix_s A) {
    EL_TRY(Pm(P)->InversePermuteRows(*M(A)))
}

ElError ElBlocksize(ElInt* b) { EL_TRY(*b = El::Blocksize()) }
ElError ElSetBlocksize(ElInt b) { EL_TRY(El::SetBlocksize(b)) }
ElError ElPushBlocksizeStack(ElInt b) { EL_TRY(El::PushBlocksizeStack(b)) }
ElError ElPopBlocksizeStack(void) { EL_TRY(El::PopBlocksizeStack()) }

// The factorization, solve and level-3 entry points for one field (factor-C.cpp:70-135,
// 268-323 and solve-C.cpp:18-46 define them per field through C_PROTO_FIELD).
#define EL_AMD_C_PROTO(SIG, F)                                                                 \
    ElError ElLU_##SIG(ElMatrix_##SIG A) { EL_TRY(El::LU(*M(A))) }                             \
    ElError ElLUPartialPiv_##SIG(ElMatrix_##SIG A, ElPermutation P) {                          \
        EL_TRY(El::LU(*M(A), *Pm(P)))                                                          \
    }                                                                                          \
    ElError ElLUFullPiv_##SIG(ElMatrix_##SIG A, ElPermutation P, ElPermutation Q) {            \
        EL_TRY(El::LU(*M(A), *Pm(P), *Pm(Q)))                                                  \
    }                                                                                          \
    ElError ElSolveAfterLUFullPiv_##SIG(ElOrientation o, ElConstMatrix_##SIG A,                \
                                        ElConstPermutation P, ElConstPermutation Q,            \
                                        ElMatrix_##SIG B) {                                    \
        EL_TRY(El::lu::SolveAfter(O(o), *M(A), *Pm(P), *Pm(Q), *M(B)))                         \
    }                                                                                          \
    ElError ElSolveAfterLU_##SIG(ElOrientation o, ElConstMatrix_##SIG A, ElMatrix_##SIG B) {   \
        EL_TRY(El::lu::SolveAfter(O(o), *M(A), *M(B)))                                         \
    }                                                                                          \
    ElError ElSolveAfterLUPartialPiv_##SIG(ElOrientation o, ElConstMatrix_##SIG A,             \
                                           ElConstPermutation P, ElMatrix_##SIG B) {           \
        EL_TRY(El::lu::SolveAfter(O(o), *M(A), *Pm(P), *M(B)))                                 \
    }                                                                                          \
    ElError ElCholesky_##SIG(ElUpperOrLower u, ElMatrix_##SIG A) {                             \
        EL_TRY(El::Cholesky(U(u), *M(A)))                                                      \
    }                                                                                          \
    ElError ElSolveAfterCholesky_##SIG(ElUpperOrLower u, ElOrientation o,                      \
                                       ElConstMatrix_##SIG A, ElMatrix_##SIG B) {              \
        EL_TRY(El::cholesky::SolveAfter(U(u), O(o), *M(A), *M(B)))                             \
    }                                                                                          \
    ElError ElLinearSolve_##SIG(ElConstMatrix_##SIG A, ElMatrix_##SIG B) {                     \
        EL_TRY(El::LinearSolve(*M(A), *M(B)))                                                  \
    }                                                                                          \
    ElError ElHPDSolve_##SIG(ElUpperOrLower u, ElOrientation o, ElConstMatrix_##SIG A,         \
                             ElMatrix_##SIG B) {                                               \
        EL_TRY(El::HPDSolve(U(u), O(o), *M(A), *M(B)))                                         \
    }                                                                                          \
    ElError ElGemm_##SIG(ElOrientation oA, ElOrientation oB, F alpha, ElConstMatrix_##SIG A,   \
                         ElConstMatrix_##SIG B, F beta, ElMatrix_##SIG C) {                    \
        EL_TRY(El::Gemm(O(oA), O(oB), alpha, *M(A), *M(B), beta, *M(C)))                      \
    }                                                                                          \
    ElError ElTrsm_##SIG(ElLeftOrRight side, ElUpperOrLower u, ElOrientation o,                \
                         ElUnitOrNonUnit diag, F alpha, ElConstMatrix_##SIG A,                 \
                         ElMatrix_##SIG B) {                                                   \
        EL_TRY(El::Trsm(static_cast<El::LeftOrRight>(side), U(u), O(o),                        \
                        static_cast<El::UnitOrNonUnit>(diag), alpha, *M(A), *M(B)))            \
    }                                                                                          \
    ElError ElHerk_##SIG(ElUpperOrLower u, ElOrientation o, F alpha, ElConstMatrix_##SIG A,    \
                         F beta, ElMatrix_##SIG C) {                                           \
        EL_TRY(El::Herk(U(u), O(o), alpha, *M(A), beta, *M(C)))                                \
    }                                                                                          \
    ElError ElSyrk_##SIG(ElUpperOrLower u, ElOrientation o, F alpha, ElConstMatrix_##SIG A,    \
                         F beta, ElMatrix_##SIG C) {                                           \
        EL_TRY(El::Syrk(U(u), O(o), alpha, *M(A), beta, *M(C)))                                \
    }
EL_AMD_C_PROTO(s, float)
EL_AMD_C_PROTO(d, double)
#undef EL_AMD_C_PROTO

}  // extern "C"
