/* El.h -- Elemental's C interface for the sequential dense factor-and-solve path, served by
 * libelemental_amd.so. Names, handle types and error codes follow the reference:
 *   ElError codes            include/El/core/types.h:45-54
 *   ElMatrix_d handles       include/El/core/Matrix.h:30-90
 *   ElPermutation            include/El/core/Permutation.h:49-98
 *   ElLU_, ElLUPartialPiv_, ElSolveAfterLU_, ElSolveAfterLUPartialPiv_, ElCholesky_,
 *   ElSolveAfterCholesky_    src/lapack_like/factor-C.cpp:70-135,268-323
 *   ElLinearSolve_, ElHPDSolve_
 *                            src/lapack_like/solve-C.cpp:18-46
 * each for float (_s) and double (_d), as the reference's C_PROTO_FIELD defines them.
 * Exceptions map to codes as EL_TRY does (include/El/core/CReflect.hpp:12-26): singular /
 * non-HPD matrices -> EL_RUNTIME_ERROR, bad arguments -> EL_LOGIC_ERROR. */
#ifndef EL_AMD_H
#define EL_AMD_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ElInt;

typedef enum {
    EL_SUCCESS,
    EL_ALLOC_ERROR,
    EL_OUT_OF_BOUNDS_ERROR,
    EL_ARG_ERROR,
    EL_LOGIC_ERROR,
    EL_RUNTIME_ERROR,
    EL_ERROR = -1
} ElError;

typedef enum { EL_NORMAL, EL_TRANSPOSE, EL_ADJOINT } ElOrientation;
typedef enum { EL_LOWER, EL_UPPER } ElUpperOrLower;
typedef enum { EL_LEFT, EL_RIGHT } ElLeftOrRight;
typedef enum { EL_NON_UNIT, EL_UNIT } ElUnitOrNonUnit;

typedef struct ElMatrix_sDummy* ElMatrix_s;
typedef const struct ElMatrix_sDummy* ElConstMatrix_s;
typedef struct ElMatrix_dDummy* ElMatrix_d;
typedef const struct ElMatrix_dDummy* ElConstMatrix_d;
typedef struct ElMatrix_iDummy* ElMatrix_i;
typedef const struct ElMatrix_iDummy* ElConstMatrix_i;
typedef struct ElPermutationDummy* ElPermutation;
typedef const struct ElPermutationDummy* ElConstPermutation;

const char* ElErrorString(ElError error);

/* Matrix<double> */
ElError ElMatrixCreate_d(ElMatrix_d* A);
ElError ElMatrixDestroy_d(ElConstMatrix_d A);
ElError ElMatrixResize_d(ElMatrix_d A, ElInt height, ElInt width);
ElError ElMatrixResizeWithLDim_d(ElMatrix_d A, ElInt height, ElInt width, ElInt ldim);
ElError ElMatrixAttach_d(ElMatrix_d A, ElInt height, ElInt width, double* buffer, ElInt ldim);
ElError ElMatrixHeight_d(ElConstMatrix_d A, ElInt* height);
ElError ElMatrixWidth_d(ElConstMatrix_d A, ElInt* width);
ElError ElMatrixLDim_d(ElConstMatrix_d A, ElInt* ldim);
ElError ElMatrixBuffer_d(ElMatrix_d A, double** buffer);
ElError ElMatrixLockedBuffer_d(ElConstMatrix_d A, const double** buffer);

/* Matrix<float> */
ElError ElMatrixCreate_s(ElMatrix_s* A);
ElError ElMatrixDestroy_s(ElConstMatrix_s A);
ElError ElMatrixResize_s(ElMatrix_s A, ElInt height, ElInt width);
ElError ElMatrixResizeWithLDim_s(ElMatrix_s A, ElInt height, ElInt width, ElInt ldim);
ElError ElMatrixAttach_s(ElMatrix_s A, ElInt height, ElInt width, float* buffer, ElInt ldim);
ElError ElMatrixHeight_s(ElConstMatrix_s A, ElInt* height);
ElError ElMatrixWidth_s(ElConstMatrix_s A, ElInt* width);
ElError ElMatrixLDim_s(ElConstMatrix_s A, ElInt* ldim);
ElError ElMatrixBuffer_s(ElMatrix_s A, float** buffer);
ElError ElMatrixLockedBuffer_s(ElConstMatrix_s A, const float** buffer);

/* Matrix<Int> */
ElError ElMatrixCreate_i(ElMatrix_i* A);
ElError ElMatrixDestroy_i(ElConstMatrix_i A);
ElError ElMatrixHeight_i(ElConstMatrix_i A, ElInt* height);
ElError ElMatrixLockedBuffer_i(ElConstMatrix_i A, const ElInt** buffer);

/* Permutation */
ElError ElPermutationCreate(ElPermutation* P);
ElError ElPermutationDestroy(ElConstPermutation P);
ElError ElPermutationMakeIdentity(ElPermutation P, ElInt size);
ElError ElPermutationSwap(ElPermutation P, ElInt origin, ElInt dest);
ElError ElPermutationHeight(ElConstPermutation P, ElInt* height);
ElError ElPermutationParity(ElConstPermutation P, bool* parity);
ElError ElPermutationSwapDestinations(ElConstPermutation P, ElInt* dests, ElInt* numSwaps);
ElError ElPermutationExplicitVector(ElConstPermutation P, ElMatrix_i p);
ElError ElPermutationPermuteRows_d(ElConstPermutation P, ElMatrix_d A);
ElError ElPermutationInversePermuteRows_d(ElConstPermutation P, ElMatrix_d A);
ElError ElPermutationPermuteRows_s(ElConstPermutation P, ElMatrix_s A);
ElError ElPermutationInversePermuteRows_s(ElConstPermutation P, ElMatrix_s A);

/* Blocksize */
ElError ElBlocksize(ElInt* blocksize);
ElError ElSetBlocksize(ElInt blocksize);
ElError ElPushBlocksizeStack(ElInt blocksize);
ElError ElPopBlocksizeStack(void);

/* Factorizations and solves, for float (_s) and double (_d): ElLU_ (no pivoting),
 * ElLUPartialPiv_, ElLUFullPiv_, ElSolveAfterLU_, ElSolveAfterLUPartialPiv_,
 * ElSolveAfterLUFullPiv_ (factor-C.cpp:268-323),
 * ElCholesky_, ElSolveAfterCholesky_ (factor-C.cpp:70-135), ElLinearSolve_, ElHPDSolve_
 * (solve-C.cpp:18-46). */
ElError ElLU_s(ElMatrix_s A);
ElError ElLUPartialPiv_s(ElMatrix_s A, ElPermutation P);
ElError ElSolveAfterLU_s(ElOrientation orientation, ElConstMatrix_s A, ElMatrix_s B);
ElError ElLUFullPiv_s(ElMatrix_s A, ElPermutation P, ElPermutation Q);
ElError ElSolveAfterLUFullPiv_s(ElOrientation orientation, ElConstMatrix_s A,
                                ElConstPermutation P, ElConstPermutation Q, ElMatrix_s B);
ElError ElSolveAfterLUPartialPiv_s(ElOrientation orientation, ElConstMatrix_s A,
                                   ElConstPermutation P, ElMatrix_s B);
ElError ElCholesky_s(ElUpperOrLower uplo, ElMatrix_s A);
ElError ElSolveAfterCholesky_s(ElUpperOrLower uplo, ElOrientation orientation,
                               ElConstMatrix_s A, ElMatrix_s B);
ElError ElLinearSolve_s(ElConstMatrix_s A, ElMatrix_s B);
ElError ElHPDSolve_s(ElUpperOrLower uplo, ElOrientation orientation, ElConstMatrix_s A,
                     ElMatrix_s B);

ElError ElLU_d(ElMatrix_d A);
ElError ElLUPartialPiv_d(ElMatrix_d A, ElPermutation P);
ElError ElSolveAfterLU_d(ElOrientation orientation, ElConstMatrix_d A, ElMatrix_d B);
ElError ElLUFullPiv_d(ElMatrix_d A, ElPermutation P, ElPermutation Q);
ElError ElSolveAfterLUFullPiv_d(ElOrientation orientation, ElConstMatrix_d A,
                                ElConstPermutation P, ElConstPermutation Q, ElMatrix_d B);
ElError ElSolveAfterLUPartialPiv_d(ElOrientation orientation, ElConstMatrix_d A,
                                   ElConstPermutation P, ElMatrix_d B);
ElError ElCholesky_d(ElUpperOrLower uplo, ElMatrix_d A);
ElError ElSolveAfterCholesky_d(ElUpperOrLower uplo, ElOrientation orientation,
                               ElConstMatrix_d A, ElMatrix_d B);
ElError ElLinearSolve_d(ElConstMatrix_d A, ElMatrix_d B);
ElError ElHPDSolve_d(ElUpperOrLower uplo, ElOrientation orientation, ElConstMatrix_d A,
                     ElMatrix_d B);

/* Level 3 (src/blas_like/level3-C.cpp) */
ElError ElGemm_s(ElOrientation orientA, ElOrientation orientB, float alpha, ElConstMatrix_s A,
                 ElConstMatrix_s B, float beta, ElMatrix_s C);
ElError ElTrsm_s(ElLeftOrRight side, ElUpperOrLower uplo, ElOrientation orientation,
                 ElUnitOrNonUnit diag, float alpha, ElConstMatrix_s A, ElMatrix_s B);
ElError ElHerk_s(ElUpperOrLower uplo, ElOrientation orientation, float alpha,
                 ElConstMatrix_s A, float beta, ElMatrix_s C);
ElError ElSyrk_s(ElUpperOrLower uplo, ElOrientation orientation, float alpha,
                 ElConstMatrix_s A, float beta, ElMatrix_s C);

ElError ElGemm_d(ElOrientation orientA, ElOrientation orientB, double alpha, ElConstMatrix_d A,
                 ElConstMatrix_d B, double beta, ElMatrix_d C);
ElError ElTrsm_d(ElLeftOrRight side, ElUpperOrLower uplo, ElOrientation orientation,
                 ElUnitOrNonUnit diag, double alpha, ElConstMatrix_d A, ElMatrix_d B);
ElError ElHerk_d(ElUpperOrLower uplo, ElOrientation orientation, double alpha,
                 ElConstMatrix_d A, double beta, ElMatrix_d C);
ElError ElSyrk_d(ElUpperOrLower uplo, ElOrientation orientation, double alpha,
                 ElConstMatrix_d A, double beta, ElMatrix_d C);

#ifdef __cplusplus
}
#endif

#endif /* EL_AMD_H */
