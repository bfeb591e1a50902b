/*
 * blis_oracle.h — CPU restatement of BLIS's level-3 hot path, used ONLY as
 * test infrastructure: the parity checker for the GPU library in tests/,
 * in __graft_entry__.smoke(), and as the CPU baseline timed by bench.py.
 * Nothing in the product library links, loads or calls it.
 *
 * Each function names the reference code it restates (paths relative to the
 * BLIS source tree). Strides follow BLIS: element (i, j) of a matrix is at
 * p[i*rs + j*cs]; transposition is given by a flag as in the typed API.
 * Enum arguments use BLIS's encodings (trans: 0 = N, 8 = T; uplo: 0xC0 =
 * lower, 0x60 = upper; side: 0 = left, 1 = right; diag: 0 = non-unit,
 * 0x100 = unit).
 */
#ifndef BLIS_ORACLE_H
#define BLIS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Five-loop GEMM with packm and the reference microkernel. */
void oracle_dgemm(int transa, int transb, int64_t m, int64_t n, int64_t k,
                  double alpha, const double* a, int64_t rs_a, int64_t cs_a,
                  const double* b, int64_t rs_b, int64_t cs_b,
                  double beta, double* c, int64_t rs_c, int64_t cs_c);
void oracle_sgemm(int transa, int transb, int64_t m, int64_t n, int64_t k,
                  float alpha, const float* a, int64_t rs_a, int64_t cs_a,
                  const float* b, int64_t rs_b, int64_t cs_b,
                  float beta, float* c, int64_t rs_c, int64_t cs_c);

/* Same algorithm with the ic loop spread over OpenMP threads (CPU baseline). */
void oracle_dgemm_mt(int transa, int transb, int64_t m, int64_t n, int64_t k,
                     double alpha, const double* a, int64_t rs_a, int64_t cs_a,
                     const double* b, int64_t rs_b, int64_t cs_b,
                     double beta, double* c, int64_t rs_c, int64_t cs_c, int nthreads);
void oracle_sgemm_mt(int transa, int transb, int64_t m, int64_t n, int64_t k,
                     float alpha, const float* a, int64_t rs_a, int64_t cs_a,
                     const float* b, int64_t rs_b, int64_t cs_b,
                     float beta, float* c, int64_t rs_c, int64_t cs_c, int nthreads);

/* TRSM: B := alpha * inv(op(A)) * B (left) or alpha * B * inv(op(A)) (right). */
void oracle_dtrsm(int side, int uplo, int transa, int diag, int64_t m, int64_t n,
                  double alpha, const double* a, int64_t rs_a, int64_t cs_a,
                  double* b, int64_t rs_b, int64_t cs_b);
void oracle_strsm(int side, int uplo, int transa, int diag, int64_t m, int64_t n,
                  float alpha, const float* a, int64_t rs_a, int64_t cs_a,
                  float* b, int64_t rs_b, int64_t cs_b);

/* Level-1v / level-2. */
void oracle_daxpyv(int64_t n, double alpha, const double* x, int64_t incx,
                   double* y, int64_t incy);
void oracle_saxpyv(int64_t n, float alpha, const float* x, int64_t incx,
                   float* y, int64_t incy);
double oracle_ddotv(int64_t n, const double* x, int64_t incx, const double* y, int64_t incy);
float  oracle_sdotv(int64_t n, const float* x, int64_t incx, const float* y, int64_t incy);
void oracle_dgemv(int transa, int64_t m, int64_t n, double alpha,
                  const double* a, int64_t rs_a, int64_t cs_a,
                  const double* x, int64_t incx, double beta, double* y, int64_t incy);
void oracle_sgemv(int transa, int64_t m, int64_t n, float alpha,
                  const float* a, int64_t rs_a, int64_t cs_a,
                  const float* x, int64_t incx, float beta, float* y, int64_t incy);

/* Batched GEMM, group semantics of ?gemm_batch_ (one call per problem). */
void oracle_dgemm_batch(const int* transa, const int* transb,
                        const int64_t* m, const int64_t* n, const int64_t* k,
                        const double* alpha, const double* const* a,
                        const int64_t* lda, const double* const* b, const int64_t* ldb,
                        const double* beta, double* const* c, const int64_t* ldc,
                        int64_t group_count, const int64_t* group_size);
void oracle_sgemm_batch(const int* transa, const int* transb,
                        const int64_t* m, const int64_t* n, const int64_t* k,
                        const float* alpha, const float* const* a,
                        const int64_t* lda, const float* const* b, const int64_t* ldb,
                        const float* beta, float* const* c, const int64_t* ldc,
                        int64_t group_count, const int64_t* group_size);

/* Double-double ("FP64x2") GEMM: C := alpha*op(A)*op(B) + beta*C with every
   product and sum carried in double-double (Dekker/Knuth error-free
   transformations); the rounded high part is returned in c. Used as the
   accuracy reference for the split-precision GEMM. Rows [r0, r1) only. */
void oracle_dgemm_dd_rows(int transa, int transb, int64_t m, int64_t n, int64_t k,
                          double alpha, const double* a, int64_t rs_a, int64_t cs_a,
                          const double* b, int64_t rs_b, int64_t cs_b,
                          double beta, double* c, int64_t rs_c, int64_t cs_c,
                          int64_t r0, int64_t r1, int nthreads);

/* Structured level-3 (blis_oracle_struc.c). uplo/side/diag/trans use the
   encodings above; C of gemmt/syrk/syr2k is m x m. */
#define ORACLE_STRUC_DECL(T, ch)                                                       \
void oracle_##ch##gemmt(int uploc, int transa, int transb, int64_t m, int64_t k,       \
                        T alpha, const T* a, int64_t rs_a, int64_t cs_a,               \
                        const T* b, int64_t rs_b, int64_t cs_b, T beta,                \
                        T* c, int64_t rs_c, int64_t cs_c);                             \
void oracle_##ch##syrk(int uploc, int transa, int64_t m, int64_t k, T alpha,           \
                       const T* a, int64_t rs_a, int64_t cs_a, T beta,                 \
                       T* c, int64_t rs_c, int64_t cs_c);                              \
void oracle_##ch##syr2k(int uploc, int transa, int transb, int64_t m, int64_t k,       \
                        T alpha, const T* a, int64_t rs_a, int64_t cs_a,               \
                        const T* b, int64_t rs_b, int64_t cs_b, T beta,                \
                        T* c, int64_t rs_c, int64_t cs_c);                             \
void oracle_##ch##symm(int side, int uploa, int transb, int64_t m, int64_t n,         \
                       T alpha, const T* a, int64_t rs_a, int64_t cs_a,                \
                       const T* b, int64_t rs_b, int64_t cs_b, T beta,                 \
                       T* c, int64_t rs_c, int64_t cs_c);                              \
void oracle_##ch##trmm3(int side, int uploa, int transa, int diaga, int transb,       \
                        int64_t m, int64_t n, T alpha, const T* a, int64_t rs_a,       \
                        int64_t cs_a, const T* b, int64_t rs_b, int64_t cs_b, T beta,  \
                        T* c, int64_t rs_c, int64_t cs_c);                             \
void oracle_##ch##trmm(int side, int uploa, int transa, int diaga, int64_t m,         \
                       int64_t n, T alpha, const T* a, int64_t rs_a, int64_t cs_a,     \
                       T* b, int64_t rs_b, int64_t cs_b);
ORACLE_STRUC_DECL(double, d)
ORACLE_STRUC_DECL(float, s)

/* Complex GEMM by the 1m method (blis_oracle_cx.c). Matrices are arrays of
   interleaved (re, im) pairs; strides count complex elements; transa/transb
   may carry the conjugate bit 0x10 besides the transpose bit 0x8. */
void oracle_zgemm(int transa, int transb, int64_t m, int64_t n, int64_t k,
                  double alpha_re, double alpha_im, const double* a, int64_t rs_a,
                  int64_t cs_a, const double* b, int64_t rs_b, int64_t cs_b,
                  double beta_re, double beta_im, double* c, int64_t rs_c, int64_t cs_c);
void oracle_cgemm(int transa, int transb, int64_t m, int64_t n, int64_t k,
                  float alpha_re, float alpha_im, const float* a, int64_t rs_a,
                  int64_t cs_a, const float* b, int64_t rs_b, int64_t cs_b,
                  float beta_re, float beta_im, float* c, int64_t rs_c, int64_t cs_c);

#ifdef __cplusplus
}
#endif

#endif
