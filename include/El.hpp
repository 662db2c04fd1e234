// El.hpp -- Elemental's C++ API for the sequential dense factor-and-solve path, served by
// the MI355X kernels of libelemental_amd.so.
//
// The declarations keep the reference's names, argument meaning and exceptions so that
// callers of this path compile and link unchanged:
//   Matrix<T>                     include/El/core/Matrix/decl.hpp:17-237
//   Permutation                   include/El/core/Permutation.hpp:14-118
//   Blocksize / SetBlocksize      src/blas_like/blocksizes.cpp:16-67 (default 128)
//   SingularMatrixException,
//   NonHPDMatrixException         include/El/core/environment/decl.hpp:209-231
//   LU, lu::SolveAfter            include/El/lapack_like/factor.hpp:364-392,418-460
//                                 (no, partial and full pivoting; F = float and double)
//   Cholesky, cholesky::SolveAfter include/El/lapack_like/factor.hpp:21-40
//   LinearSolve, HPDSolve         include/El/lapack_like/solve.hpp:19-38,188-212
//   Gemm, Trsm, Herk, Syrk        include/El/blas_like/level3.hpp:34-120,326-334,455-468
// Matrices live in host memory, as in the reference; each call stages the operands in
// device memory, runs the factorization or solve on the GPU and copies the results back.
// Callers that keep data resident in HBM use the device ABI in elemental_amd.h instead.
#ifndef EL_AMD_HPP
#define EL_AMD_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace El {

typedef int Int;  // ElInt is 32-bit unless EL_USE_64BIT_INTS (include/El/core/types.h:28-32)

enum Orientation { NORMAL = 0, TRANSPOSE = 1, ADJOINT = 2 };
enum UpperOrLower { LOWER = 0, UPPER = 1 };
enum LeftOrRight { LEFT = 0, RIGHT = 1 };
enum UnitOrNonUnit { NON_UNIT = 0, UNIT = 1 };

class SingularMatrixException : public std::runtime_error {
   public:
    explicit SingularMatrixException(const char* msg = "Matrix was singular")
        : std::runtime_error(msg) {}
};
class NonHPDMatrixException : public std::runtime_error {
   public:
    explicit NonHPDMatrixException(const char* msg = "Matrix was not HPD")
        : std::runtime_error(msg) {}
};
class ArgException : public std::logic_error {
   public:
    explicit ArgException(const std::string& msg) : std::logic_error(msg) {}
};

void LogicError(const std::string& msg);
void RuntimeError(const std::string& msg);

const Int END = -100;
struct Range {
    Int beg, end;
    Range() : beg(0), end(END) {}
    Range(Int b, Int e) : beg(b), end(e) {}
};
typedef Range IR;
extern const Range ALL;

// Column-major matrix: entry (i,j) at Buffer()[i + j*LDim()], LDim() >= max(Height(),1).
// Owning, or a (locked) view of another buffer; A(IR,IR) returns a view.
template <typename T>
class Matrix {
   public:
    Matrix() : height_(0), width_(0), ldim_(1), buf_(nullptr), viewing_(false), locked_(false) {}
    Matrix(Int height, Int width) : Matrix() { Resize(height, width); }
    Matrix(Int height, Int width, Int ldim) : Matrix() { Resize(height, width, ldim); }
    Matrix(const Matrix<T>& A);
    Matrix<T>& operator=(const Matrix<T>& A);
    ~Matrix() = default;

    Int Height() const { return height_; }
    Int Width() const { return width_; }
    Int LDim() const { return ldim_; }
    bool Viewing() const { return viewing_; }
    bool Locked() const { return locked_; }

    T* Buffer() {
        if (locked_) LogicError("Cannot return non-const buffer of a locked Matrix");
        return buf_;
    }
    T* Buffer(Int i, Int j) { return Buffer() + i + (int64_t)j * ldim_; }
    const T* LockedBuffer() const { return buf_; }
    const T* LockedBuffer(Int i, Int j) const { return buf_ + i + (int64_t)j * ldim_; }

    T Get(Int i, Int j) const { return buf_[i + (int64_t)j * ldim_]; }
    void Set(Int i, Int j, T alpha) { Buffer()[i + (int64_t)j * ldim_] = alpha; }
    T& operator()(Int i, Int j) { return Buffer()[i + (int64_t)j * ldim_]; }
    const T& operator()(Int i, Int j) const { return buf_[i + (int64_t)j * ldim_]; }

    void Empty();
    void Resize(Int height, Int width);
    void Resize(Int height, Int width, Int ldim);
    void Attach(Int height, Int width, T* buffer, Int ldim);
    void LockedAttach(Int height, Int width, const T* buffer, Int ldim);

    // Views sharing this buffer (Range end END means "to the end").
    Matrix<T> operator()(Range I, Range J);
    const Matrix<T> operator()(Range I, Range J) const;

   private:
    Int height_, width_, ldim_;
    T* buf_;
    bool viewing_, locked_;
    std::vector<T> mem_;
};

// A permutation stored as a sequence of row swaps whose origins are 0,1,2,... (the mode
// El::LU produces). SwapDestinations() is LAPACK's ipiv - 1.
class Permutation {
   public:
    Permutation() : size_(0), parity_(false) {}
    void Empty();
    void MakeIdentity(Int size);
    void ReserveSwaps(Int maxSwaps);
    void Swap(Int origin, Int dest);
    Int Height() const { return size_; }
    Int Width() const { return size_; }
    bool Parity() const { return parity_; }
    Int NumSwaps() const { return (Int)dests_.size(); }
    const std::vector<Int>& SwapDestinations() const { return dests_; }
    void ExplicitVector(Matrix<Int>& p) const;
    void ExplicitMatrix(Matrix<Int>& P) const;
    template <typename T>
    void PermuteRows(Matrix<T>& A, Int offset = 0) const;
    template <typename T>
    void InversePermuteRows(Matrix<T>& A, Int offset = 0) const;
    // Used by the factorizations to install a whole swap sequence at once.
    void SetSwapDestinations(Int size, const std::vector<Int>& dests);

   private:
    Int size_;
    bool parity_;
    std::vector<Int> dests_;
};

// Algorithmic blocksize stack (default 128). Any positive blocksize is accepted; the GPU
// factorizations run blocks wider than 128 columns as 128-column panels, which gives the same
// pivots as the wider blocking in exact arithmetic.
Int Blocksize();
void SetBlocksize(Int blocksize);
void PushBlocksizeStack(Int blocksize);
void PopBlocksizeStack();

// ---- level 3 ----
template <typename T>
void Gemm(Orientation orientA, Orientation orientB, T alpha, const Matrix<T>& A,
          const Matrix<T>& B, T beta, Matrix<T>& C);
template <typename F>
void Trsm(LeftOrRight side, UpperOrLower uplo, Orientation orientation, UnitOrNonUnit diag,
          F alpha, const Matrix<F>& A, Matrix<F>& B, bool checkIfSingular = false);
template <typename T>
void Syrk(UpperOrLower uplo, Orientation orientation, T alpha, const Matrix<T>& A, T beta,
          Matrix<T>& C, bool conjugate = false);
template <typename T>
void Herk(UpperOrLower uplo, Orientation orientation, T alpha, const Matrix<T>& A, T beta,
          Matrix<T>& C);

// ---- factorizations and solves (instantiated for F = float and double) ----
template <typename F>
void LU(Matrix<F>& A);
template <typename F>
void LU(Matrix<F>& A, Permutation& P);
template <typename F>
void LU(Matrix<F>& A, Permutation& P, Permutation& Q);
namespace lu {
template <typename F>
void SolveAfter(Orientation orientation, const Matrix<F>& A, Matrix<F>& B);
template <typename F>
void SolveAfter(Orientation orientation, const Matrix<F>& A, const Permutation& P, Matrix<F>& B);
template <typename F>
void SolveAfter(Orientation orientation, const Matrix<F>& A, const Permutation& P,
                const Permutation& Q, Matrix<F>& B);
}  // namespace lu
template <typename F>
void Cholesky(UpperOrLower uplo, Matrix<F>& A);
namespace cholesky {
template <typename F>
void SolveAfter(UpperOrLower uplo, Orientation orientation, const Matrix<F>& A, Matrix<F>& B);
}  // namespace cholesky
template <typename F>
void LinearSolve(const Matrix<F>& A, Matrix<F>& B);
template <typename F>
void HPDSolve(UpperOrLower uplo, Orientation orientation, const Matrix<F>& A, Matrix<F>& B);

// ---- equilibration ----
// A := diag(dRow)^{-1} A diag(dCol)^{-1}, four Ruiz sweeps
// (include/El/lapack_like/equilibrate.hpp:17-21). Defined for double.
void RuizEquil(Matrix<double>& A, Matrix<double>& dRow, Matrix<double>& dCol,
               bool progress = false);
// A := diag(dRow)^{-1} A diag(dCol)^{-1}, geometric scaling (GeomEquil.cpp:26-116).
// Defined for double.
void GeomEquil(Matrix<double>& A, Matrix<double>& dRow, Matrix<double>& dCol,
               bool progress = false);
// A := diag(d)^{-1} A diag(d)^{-1}, maxiter symmetric sweeps
// (include/El/lapack_like/equilibrate.hpp:80-84). Defined for double.
void SymmetricRuizEquil(Matrix<double>& A, Matrix<double>& d, Int maxiter = 3,
                        bool progress = false);

}  // namespace El

#endif  // EL_AMD_HPP
