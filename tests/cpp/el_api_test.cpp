// Drives the El:: C++ API of libelemental_amd.so the way the reference's own tests drive
// Elemental (tests/lapack_like/LU.cpp:20-53, tests/lapack_like/Cholesky.cpp:20-44): factor,
// solve 100 random right-hand sides, and check
//   LU:       ||A Y - X||_oo / (eps max(m,n) max(||A||_1, ||Y||_1)) <= 100
//   Cholesky: ||X - A \ (A X)||_oo / (eps n ||A X||_1) <= 100
// with eps = 2^-53. Unlike the reference tests (which print and exit 0), a failed check
// exits non-zero. Also checks the exception types on singular / non-HPD input.
#include <El.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using El::Int;
using El::Matrix;

static const double kEps = std::ldexp(1.0, -53);
static std::mt19937 gen(1376256);

static void Uniform(Matrix<double>& A, Int m, Int n) {
    A.Resize(m, n);
    std::uniform_real_distribution<double> d(-1.0, 1.0);
    for (Int j = 0; j < n; ++j)
        for (Int i = 0; i < m; ++i) A(i, j) = d(gen);
}
static double OneNorm(const Matrix<double>& A) {
    double best = 0;
    for (Int j = 0; j < A.Width(); ++j) {
        double s = 0;
        for (Int i = 0; i < A.Height(); ++i) s += std::fabs(A.Get(i, j));
        best = std::max(best, s);
    }
    return best;
}
static double InfinityNorm(const Matrix<double>& A) {
    double best = 0;
    for (Int i = 0; i < A.Height(); ++i) {
        double s = 0;
        for (Int j = 0; j < A.Width(); ++j) s += std::fabs(A.Get(i, j));
        best = std::max(best, s);
    }
    return best;
}

static int failures = 0;
static void check(bool ok, const char* what, double v) {
    std::printf("%-40s %s (%.4g)\n", what, ok ? "ok" : "FAILED", v);
    if (!ok) ++failures;
}

static int run() {
    const Int m = 300, numRHS = 100;
    // ---- LU with partial pivoting + SolveAfter (tests/lapack_like/LU.cpp:20-53) ----
    {
        Matrix<double> A, AOrig;
        Uniform(A, m, m);
        AOrig = A;
        El::Permutation P;
        El::LU(A, P);
        Matrix<double> X, Y;
        Uniform(X, m, numRHS);
        Y = X;
        const double oneNormY = OneNorm(Y), oneNormA = OneNorm(AOrig);
        El::lu::SolveAfter(El::NORMAL, A, P, Y);
        El::Gemm(El::NORMAL, El::NORMAL, -1.0, AOrig, Y, 1.0, X);
        const double rel = InfinityNorm(X) / (kEps * m * std::max(oneNormA, oneNormY));
        check(rel <= 100.0, "LU + SolveAfter relError", rel);
        check(P.NumSwaps() == m, "LU permutation length", (double)P.NumSwaps());
    }
    // ---- LinearSolve ----
    {
        Matrix<double> A, B, X;
        Uniform(A, m, m);
        Uniform(B, m, 3);
        X = B;
        El::LinearSolve(A, X);
        Matrix<double> R(B);
        El::Gemm(El::NORMAL, El::NORMAL, -1.0, A, X, 1.0, R);
        const double rel = InfinityNorm(R) / (kEps * m * std::max(OneNorm(A), OneNorm(X)));
        check(rel <= 100.0, "LinearSolve relError", rel);
    }
    // ---- Cholesky + SolveAfter (tests/lapack_like/Cholesky.cpp:20-44) ----
    for (int uplo = 0; uplo < 2; ++uplo) {
        Matrix<double> G, A;
        Uniform(G, m, m);
        A.Resize(m, m);
        El::Gemm(El::NORMAL, El::TRANSPOSE, 1.0 / m, G, G, 0.0, A);
        for (Int i = 0; i < m; ++i) A(i, i) += 1.0;
        Matrix<double> AOrig(A);
        El::Cholesky(uplo == 0 ? El::LOWER : El::UPPER, A);
        Matrix<double> X, Y;
        Uniform(X, m, numRHS);
        Y.Resize(m, numRHS);
        El::Gemm(El::NORMAL, El::NORMAL, 1.0, AOrig, X, 0.0, Y);
        const double oneNormY = OneNorm(Y);
        El::cholesky::SolveAfter(uplo == 0 ? El::LOWER : El::UPPER, El::NORMAL, A, Y);
        for (Int j = 0; j < numRHS; ++j)
            for (Int i = 0; i < m; ++i) X(i, j) -= Y(i, j);
        const double rel = InfinityNorm(X) / (kEps * m * oneNormY);
        check(rel <= 100.0, uplo == 0 ? "Cholesky(LOWER) relErr" : "Cholesky(UPPER) relErr", rel);
        Matrix<double> B, Xs;
        Uniform(B, m, 2);
        Xs = B;
        El::HPDSolve(uplo == 0 ? El::LOWER : El::UPPER, El::NORMAL, AOrig, Xs);
        Matrix<double> R(B);
        El::Gemm(El::NORMAL, El::NORMAL, -1.0, AOrig, Xs, 1.0, R);
        const double rel2 = InfinityNorm(R) / (kEps * m * std::max(OneNorm(AOrig), OneNorm(Xs)));
        check(rel2 <= 100.0, "HPDSolve relError", rel2);
    }
    // ---- exceptions ----
    {
        Matrix<double> Z;
        Uniform(Z, 40, 40);
        for (Int i = 0; i < 40; ++i) Z(i, 7) = 0.0;
        El::Permutation P;
        bool thrown = false;
        try {
            El::LU(Z, P);
        } catch (El::SingularMatrixException&) {
            thrown = true;
        }
        check(thrown, "SingularMatrixException on zero column", 7.0);
        Matrix<double> N(5, 5);
        for (Int i = 0; i < 5; ++i) N(i, i) = (i == 3) ? -1.0 : 1.0;
        thrown = false;
        try {
            El::Cholesky(El::LOWER, N);
        } catch (El::NonHPDMatrixException&) {
            thrown = true;
        }
        check(thrown, "NonHPDMatrixException on negative pivot", 3.0);
        thrown = false;
        try {
            Matrix<double> R(3, 4);
            El::Cholesky(El::LOWER, R);
        } catch (std::logic_error&) {
            thrown = true;
        }
        check(thrown, "logic_error on non-square Cholesky", 0.0);
    }
    {
        // RuizEquil of a rank-one matrix u v^T with power-of-two entries: exactly the sign
        // matrix, dRow = |u| / max|u|, dCol = |v| max|u|
        const double u[4] = {8.0, -1.0 / 32, 1024.0, 0.25};
        const double v[3] = {-16.0, 1.0 / 128, 4.0};
        Matrix<double> A(4, 3), dRow, dCol;
        for (Int j = 0; j < 3; ++j)
            for (Int i = 0; i < 4; ++i) A(i, j) = u[i] * v[j];
        El::RuizEquil(A, dRow, dCol);
        bool ok = dRow.Height() == 4 && dCol.Height() == 3;
        for (Int j = 0; j < 3 && ok; ++j)
            for (Int i = 0; i < 4; ++i)
                ok = ok && A(i, j) == ((u[i] * v[j] < 0) ? -1.0 : 1.0);
        for (Int i = 0; i < 4 && ok; ++i) ok = dRow(i, 0) == std::fabs(u[i]) / 1024.0;
        for (Int j = 0; j < 3 && ok; ++j) ok = dCol(j, 0) == std::fabs(v[j]) * 1024.0;
        check(ok, "RuizEquil rank-one closed form", 0.0);
    }
    {
        // GeomEquil of u v^T with max|u| / min|u| = 64: the sign matrix, dRow = |u| / 2,
        // dCol = 2 |v| (sqrt(min|u| max|u|) = 2)
        const double u[4] = {4.0, -0.25, 16.0, 1.0};
        const double v[3] = {-128.0, 1.0 / 512, 8.0};
        Matrix<double> A(4, 3), dRow, dCol;
        for (Int j = 0; j < 3; ++j)
            for (Int i = 0; i < 4; ++i) A(i, j) = u[i] * v[j];
        El::GeomEquil(A, dRow, dCol);
        bool ok = dRow.Height() == 4 && dCol.Height() == 3;
        for (Int j = 0; j < 3 && ok; ++j)
            for (Int i = 0; i < 4; ++i)
                ok = ok && A(i, j) == ((u[i] * v[j] < 0) ? -1.0 : 1.0);
        for (Int i = 0; i < 4 && ok; ++i) ok = dRow(i, 0) == std::fabs(u[i]) / 2.0;
        for (Int j = 0; j < 3 && ok; ++j) ok = dCol(j, 0) == std::fabs(v[j]) * 2.0;
        check(ok, "GeomEquil rank-one closed form", 0.0);
        // SymmetricRuizEquil of a diagonal of powers of four: the identity, d = sqrt(diag)
        Matrix<double> D(3, 3), d;
        const double dg[3] = {1024.0, 1.0 / 64, 16.0};
        for (Int j = 0; j < 3; ++j)
            for (Int i = 0; i < 3; ++i) D(i, j) = (i == j) ? dg[i] : 0.0;
        El::SymmetricRuizEquil(D, d);
        ok = d.Height() == 3;
        for (Int j = 0; j < 3 && ok; ++j)
            for (Int i = 0; i < 3; ++i) ok = ok && D(i, j) == ((i == j) ? 1.0 : 0.0);
        for (Int i = 0; i < 3 && ok; ++i) ok = d(i, 0) == std::sqrt(dg[i]);
        check(ok, "SymmetricRuizEquil diagonal closed form", 0.0);
    }
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}

int main() {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    try {
        return run();
    } catch (std::exception& e) {
        std::printf("uncaught exception after %d failure(s): %s\n", failures, e.what());
        return 2;
    }
}
