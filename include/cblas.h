/*
 * cblas.h — CBLAS interface of the MI355X (gfx950) BLIS library.
 *
 * The C-language BLAS layer that BLIS ships (frame/compat/cblas/src/cblas.h,
 * netlib CBLAS plus BLIS's extra/ routines) for the operations this library
 * implements. Row-major callers are served the way that layer does it: a
 * row-major problem is the column-major problem on the transposed operands
 * (for GEMM, C^T = B^T A^T: the operands and m/n swap; for the structured
 * operations, side and uplo flip), handed to the Fortran-style symbols
 * (?gemm_, ?symm_, ...). Enum values are netlib's.
 *
 * Reference: frame/compat/cblas/src/cblas.h (enums :31-35, prototypes),
 * cblas_?gemm.c, cblas_?symm.c, cblas_?syrk.c, cblas_?syr2k.c, cblas_?trmm.c,
 * cblas_?trsm.c, cblas_?gemv.c, cblas_?axpy.c, cblas_?dot.c,
 * extra/cblas_?gemm_batch.c, extra/cblas_?gemmt.c, cblas_xerbla.c.
 *
 * Errors in an enum argument are reported by cblas_xerbla ("Parameter %d to
 * routine %s was incorrect", then the detail) which ends the process with
 * exit(-1), as the reference's does. Matrix and vector arguments may be
 * device or host pointers (see blis.h); every call is synchronous.
 */
#ifndef CBLAS_H
#define CBLAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef BLIS_H
typedef int32_t f77_int;
#endif

enum CBLAS_ORDER { CblasRowMajor = 101, CblasColMajor = 102 };
enum CBLAS_TRANSPOSE { CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113 };
enum CBLAS_UPLO { CblasUpper = 121, CblasLower = 122 };
enum CBLAS_DIAG { CblasNonUnit = 131, CblasUnit = 132 };
enum CBLAS_SIDE { CblasLeft = 141, CblasRight = 142 };
typedef enum CBLAS_ORDER CBLAS_LAYOUT;

/* ---- level 1 ---------------------------------------------------------- */
float  cblas_sdot( f77_int N, const float* X, f77_int incX, const float* Y, f77_int incY );
double cblas_ddot( f77_int N, const double* X, f77_int incX, const double* Y, f77_int incY );
void   cblas_saxpy( f77_int N, float alpha, const float* X, f77_int incX,
                    float* Y, f77_int incY );
void   cblas_daxpy( f77_int N, double alpha, const double* X, f77_int incX,
                    double* Y, f77_int incY );

/* ---- level 2 ---------------------------------------------------------- */
void cblas_sgemv( enum CBLAS_ORDER order, enum CBLAS_TRANSPOSE TransA, f77_int M, f77_int N,
                  float alpha, const float* A, f77_int lda, const float* X, f77_int incX,
                  float beta, float* Y, f77_int incY );
void cblas_dgemv( enum CBLAS_ORDER order, enum CBLAS_TRANSPOSE TransA, f77_int M, f77_int N,
                  double alpha, const double* A, f77_int lda, const double* X, f77_int incX,
                  double beta, double* Y, f77_int incY );

/* ---- level 3 ---------------------------------------------------------- */
#define BLIS_AMD_CBLAS_L3( ch, T ) \
void cblas_##ch##gemm( enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, \
                       enum CBLAS_TRANSPOSE TransB, f77_int M, f77_int N, f77_int K, \
                       T alpha, const T* A, f77_int lda, const T* B, f77_int ldb, \
                       T beta, T* C, f77_int ldc ); \
void cblas_##ch##symm( enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, \
                       f77_int M, f77_int N, T alpha, const T* A, f77_int lda, \
                       const T* B, f77_int ldb, T beta, T* C, f77_int ldc ); \
void cblas_##ch##syrk( enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, \
                       enum CBLAS_TRANSPOSE Trans, f77_int N, f77_int K, T alpha, \
                       const T* A, f77_int lda, T beta, T* C, f77_int ldc ); \
void cblas_##ch##syr2k( enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, \
                        enum CBLAS_TRANSPOSE Trans, f77_int N, f77_int K, T alpha, \
                        const T* A, f77_int lda, const T* B, f77_int ldb, T beta, \
                        T* C, f77_int ldc ); \
void cblas_##ch##trmm( enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, \
                       enum CBLAS_TRANSPOSE TransA, enum CBLAS_DIAG Diag, f77_int M, \
                       f77_int N, T alpha, const T* A, f77_int lda, T* B, f77_int ldb ); \
void cblas_##ch##trsm( enum CBLAS_ORDER Order, enum CBLAS_SIDE Side, enum CBLAS_UPLO Uplo, \
                       enum CBLAS_TRANSPOSE TransA, enum CBLAS_DIAG Diag, f77_int M, \
                       f77_int N, T alpha, const T* A, f77_int lda, T* B, f77_int ldb ); \
void cblas_##ch##gemmt( enum CBLAS_ORDER Order, enum CBLAS_UPLO Uplo, \
                        enum CBLAS_TRANSPOSE TransA, enum CBLAS_TRANSPOSE TransB, \
                        f77_int M, f77_int K, T alpha, const T* A, f77_int lda, \
                        const T* B, f77_int ldb, T beta, T* C, f77_int ldc ); \
void cblas_##ch##gemm_batch( enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE* TransA_array, \
                             enum CBLAS_TRANSPOSE* TransB_array, f77_int* M_array, \
                             f77_int* N_array, f77_int* K_array, const T* alpha_array, \
                             const T** A_array, f77_int* lda_array, \
                             const T** B_array, f77_int* ldb_array, \
                             const T* beta_array, T** C_array, f77_int* ldc_array, \
                             f77_int group_count, f77_int* group_size );
BLIS_AMD_CBLAS_L3( s, float )
BLIS_AMD_CBLAS_L3( d, double )

/* Complex GEMM (single and double complex; alpha and beta by address). */
#define BLIS_AMD_CBLAS_L3Z( ch ) \
void cblas_##ch##gemm( enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE TransA, \
                       enum CBLAS_TRANSPOSE TransB, f77_int M, f77_int N, f77_int K, \
                       const void* alpha, const void* A, f77_int lda, const void* B, \
                       f77_int ldb, const void* beta, void* C, f77_int ldc ); \
void cblas_##ch##gemm_batch( enum CBLAS_ORDER Order, enum CBLAS_TRANSPOSE* TransA_array, \
                             enum CBLAS_TRANSPOSE* TransB_array, f77_int* M_array, \
                             f77_int* N_array, f77_int* K_array, const void* alpha_array, \
                             const void** A_array, f77_int* lda_array, \
                             const void** B_array, f77_int* ldb_array, \
                             const void* beta_array, void** C_array, f77_int* ldc_array, \
                             f77_int group_count, f77_int* group_size );
BLIS_AMD_CBLAS_L3Z( c )
BLIS_AMD_CBLAS_L3Z( z )

void cblas_xerbla( f77_int p, const char* rout, const char* form, ... );

#ifdef __cplusplus
}
#endif

#endif /* CBLAS_H */
