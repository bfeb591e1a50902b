/*
 * blas_f77.c — BLAS-symbol wrappers around the oracle (TEST INFRASTRUCTURE
 * ONLY). They let the reference's own netlib test drivers (blastest/src/
 * dblat2.c, dblat3.c, sblat2.c, sblat3.c, compiled from the reference tree by
 * oracle/Makefile) exercise the oracle, which pins the oracle against the
 * reference's known-answer checks (blastest/src/dblat3.c:4109-4317, dmmch_).
 *
 * The structured level-3 wrappers (symm, trmm, syrk, syr2k) are at the end.
 * Argument checks and INFO codes restate frame/compat/check/bla_gemm_check.h:
 * 39-82, bla_trmm_check.h (used for trsm) and bla_gemv_check.h; the m == 1 /
 * n == 1 redirection to gemv restates frame/compat/bla_gemm.c:181-222; the
 * n_x == 0 early return restates frame/compat/bla_gemv.c:92-111; negative
 * increments restate frame/include/bli_blas_macro_defs.h:50-74.
 */
#include "blis_oracle.h"

#include <ctype.h>
#include <stdio.h>
#include <stdint.h>

typedef int32_t f77_int;

#define TR_N 0x0
#define TR_T 0x8

__attribute__((weak)) int xerbla_(const char* srname, const f77_int* info, f77_int len)
{
    printf("** On entry to %6.*s, parameter number %2i had an illegal value\n",
           (int)(len < 6 ? len : 6), srname, (int)*info);
    return 0;
}

static char up(char c) { return (char)toupper((unsigned char)c); }
static int maxi(int a, int b) { return a > b ? a : b; }

static void err(const char* name, f77_int info)
{
    xerbla_(name, &info, 6);
}

#define DEFINE_F77(T, ch, CH)                                                          \
void ch##gemv_(const char* transa, const f77_int* m, const f77_int* n, const T* alpha, \
               const T* a, const f77_int* lda, const T* x, const f77_int* incx,       \
               const T* beta, T* y, const f77_int* incy)                               \
{                                                                                      \
    const char t = up(*transa);                                                        \
    f77_int info = 0;                                                                  \
    if (t != 'N' && t != 'T' && t != 'C') info = 1;                                    \
    else if (*m < 0) info = 2;                                                         \
    else if (*n < 0) info = 3;                                                         \
    else if (*lda < maxi(1, *m)) info = 6;                                             \
    else if (*incx == 0) info = 8;                                                     \
    else if (*incy == 0) info = 11;                                                    \
    if (info) { err(CH "GEMV ", info); return; }                                       \
    const int tr = (t == 'N') ? TR_N : TR_T;                                           \
    const int64_t my = tr ? *n : *m, nx = tr ? *m : *n;                                \
    if (my > 0 && nx == 0) return;                                                     \
    const T* x0 = (*incx < 0) ? x + (nx - 1) * (int64_t)(-*incx) : x;                  \
    T* y0 = (*incy < 0) ? y + (my - 1) * (int64_t)(-*incy) : y;                        \
    oracle_##ch##gemv(tr, *m, *n, *alpha, a, 1, *lda, x0, *incx, *beta, y0, *incy);    \
}                                                                                      \
                                                                                       \
void ch##gemm_(const char* transa, const char* transb, const f77_int* m,              \
               const f77_int* n, const f77_int* k, const T* alpha, const T* a,        \
               const f77_int* lda, const T* b, const f77_int* ldb, const T* beta,     \
               T* c, const f77_int* ldc)                                               \
{                                                                                      \
    const char ta = up(*transa), tb = up(*transb);                                     \
    const int nota = ta == 'N', notb = tb == 'N';                                      \
    const f77_int nrowa = nota ? *m : *k, nrowb = notb ? *k : *n;                      \
    f77_int info = 0;                                                                  \
    if (!nota && ta != 'C' && ta != 'T') info = 1;                                     \
    else if (!notb && tb != 'C' && tb != 'T') info = 2;                                \
    else if (*m < 0) info = 3;                                                         \
    else if (*n < 0) info = 4;                                                         \
    else if (*k < 0) info = 5;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 8;                                          \
    else if (*ldb < maxi(1, nrowb)) info = 10;                                         \
    else if (*ldc < maxi(1, *m)) info = 13;                                            \
    if (info) { err(CH "GEMM ", info); return; }                                       \
    const int tra = nota ? TR_N : TR_T, trb = notb ? TR_N : TR_T;                      \
    const int64_t m0 = *m, n0 = *n, k0 = *k;                                           \
    if (n0 == 1) {                                                                     \
        const int64_t m0t = tra ? k0 : m0, k0t = tra ? m0 : k0;                        \
        oracle_##ch##gemv(tra, m0t, k0t, *alpha, a, 1, *lda, b, trb ? *ldb : 1,        \
                          *beta, c, 1);                                                \
        return;                                                                        \
    }                                                                                  \
    if (m0 == 1) {                                                                     \
        const int64_t n0t = trb ? k0 : n0, k0t = trb ? n0 : k0;                        \
        oracle_##ch##gemv(trb, n0t, k0t, *alpha, b, *ldb, 1, a, tra ? 1 : *lda,        \
                          *beta, c, *ldc);                                             \
        return;                                                                        \
    }                                                                                  \
    oracle_##ch##gemm(tra, trb, m0, n0, k0, *alpha, a, 1, *lda, b, 1, *ldb, *beta,     \
                      c, 1, *ldc);                                                     \
}                                                                                      \
                                                                                       \
void ch##trsm_(const char* side, const char* uploa, const char* transa,               \
               const char* diaga, const f77_int* m, const f77_int* n, const T* alpha,  \
               const T* a, const f77_int* lda, T* b, const f77_int* ldb)               \
{                                                                                      \
    const char s = up(*side), u = up(*uploa), t = up(*transa), d = up(*diaga);         \
    const f77_int nrowa = (s == 'L') ? *m : *n;                                        \
    f77_int info = 0;                                                                  \
    if (s != 'L' && s != 'R') info = 1;                                                \
    else if (u != 'L' && u != 'U') info = 2;                                           \
    else if (t != 'N' && t != 'T' && t != 'C') info = 3;                               \
    else if (d != 'U' && d != 'N') info = 4;                                           \
    else if (*m < 0) info = 5;                                                         \
    else if (*n < 0) info = 6;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 9;                                          \
    else if (*ldb < maxi(1, *m)) info = 11;                                            \
    if (info) { err(CH "TRSM ", info); return; }                                       \
    oracle_##ch##trsm(s == 'L' ? 0 : 1, u == 'L' ? 0xC0 : 0x60, t == 'N' ? TR_N : TR_T, \
                      d == 'U' ? 0x100 : 0, *m, *n, *alpha, a, 1, *lda, b, 1, *ldb);   \
}                                                                                      \
                                                                                       \
void ch##axpy_(const f77_int* n, const T* alpha, const T* x, const f77_int* incx,      \
               T* y, const f77_int* incy)                                              \
{                                                                                      \
    const int64_t n0 = *n < 0 ? 0 : *n;                                                \
    const T* x0 = (*incx < 0) ? x + (n0 - 1) * (int64_t)(-*incx) : x;                  \
    T* y0 = (*incy < 0) ? y + (n0 - 1) * (int64_t)(-*incy) : y;                        \
    oracle_##ch##axpyv(n0, *alpha, x0, *incx, y0, *incy);                              \
}                                                                                      \
                                                                                       \
T ch##dot_(const f77_int* n, const T* x, const f77_int* incx, const T* y,             \
           const f77_int* incy)                                                        \
{                                                                                      \
    const int64_t n0 = *n < 0 ? 0 : *n;                                                \
    const T* x0 = (*incx < 0) ? x + (n0 - 1) * (int64_t)(-*incx) : x;                  \
    const T* y0 = (*incy < 0) ? y + (n0 - 1) * (int64_t)(-*incy) : y;                  \
    return oracle_##ch##dotv(n0, x0, *incx, y0, *incy);                                \
}

DEFINE_F77(double, d, "D")
DEFINE_F77(float, s, "S")

/* Structured level-3 BLAS over the oracle. Argument checks restate
   frame/compat/check/bla_hemm_check.h (symm), bla_trmm_check.h,
   bla_syrk_check.h, bla_syr2k_check.h and bla_gemmt_check.h. */
#define DEFINE_F77_STRUC(T, ch, CH)                                                    \
void ch##symm_(const char* side, const char* uploa, const f77_int* m, const f77_int* n, \
               const T* alpha, const T* a, const f77_int* lda, const T* b,            \
               const f77_int* ldb, const T* beta, T* c, const f77_int* ldc)            \
{                                                                                      \
    const char s = up(*side), u = up(*uploa);                                          \
    const f77_int nrowa = (s == 'L') ? *m : *n;                                        \
    f77_int info = 0;                                                                  \
    if (s != 'L' && s != 'R') info = 1;                                                \
    else if (u != 'L' && u != 'U') info = 2;                                           \
    else if (*m < 0) info = 3;                                                         \
    else if (*n < 0) info = 4;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 7;                                          \
    else if (*ldb < maxi(1, *m)) info = 9;                                             \
    else if (*ldc < maxi(1, *m)) info = 12;                                            \
    if (info) { err(CH "SYMM ", info); return; }                                       \
    oracle_##ch##symm(s == 'L' ? 0 : 1, u == 'L' ? 0xC0 : 0x60, TR_N, *m, *n, *alpha,  \
                      a, 1, *lda, b, 1, *ldb, *beta, c, 1, *ldc);                      \
}                                                                                      \
                                                                                       \
void ch##trmm_(const char* side, const char* uploa, const char* transa,               \
               const char* diaga, const f77_int* m, const f77_int* n, const T* alpha,  \
               const T* a, const f77_int* lda, T* b, const f77_int* ldb)               \
{                                                                                      \
    const char s = up(*side), u = up(*uploa), t = up(*transa), d = up(*diaga);         \
    const f77_int nrowa = (s == 'L') ? *m : *n;                                        \
    f77_int info = 0;                                                                  \
    if (s != 'L' && s != 'R') info = 1;                                                \
    else if (u != 'L' && u != 'U') info = 2;                                           \
    else if (t != 'N' && t != 'T' && t != 'C') info = 3;                               \
    else if (d != 'U' && d != 'N') info = 4;                                           \
    else if (*m < 0) info = 5;                                                         \
    else if (*n < 0) info = 6;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 9;                                          \
    else if (*ldb < maxi(1, *m)) info = 11;                                            \
    if (info) { err(CH "TRMM ", info); return; }                                       \
    oracle_##ch##trmm(s == 'L' ? 0 : 1, u == 'L' ? 0xC0 : 0x60, t == 'N' ? TR_N : TR_T, \
                      d == 'U' ? 0x100 : 0, *m, *n, *alpha, a, 1, *lda, b, 1, *ldb);   \
}                                                                                      \
                                                                                       \
void ch##syrk_(const char* uploc, const char* transa, const f77_int* m, const f77_int* k, \
               const T* alpha, const T* a, const f77_int* lda, const T* beta, T* c,    \
               const f77_int* ldc)                                                     \
{                                                                                      \
    const char u = up(*uploc), t = up(*transa);                                        \
    const f77_int nrowa = (t == 'N') ? *m : *k;                                        \
    f77_int info = 0;                                                                  \
    if (u != 'L' && u != 'U') info = 1;                                                \
    else if (t != 'N' && t != 'T' && t != 'C') info = 2;                               \
    else if (*m < 0) info = 3;                                                         \
    else if (*k < 0) info = 4;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 7;                                          \
    else if (*ldc < maxi(1, *m)) info = 10;                                            \
    if (info) { err(CH "SYRK ", info); return; }                                       \
    oracle_##ch##syrk(u == 'L' ? 0xC0 : 0x60, t == 'N' ? TR_N : TR_T, *m, *k, *alpha,  \
                      a, 1, *lda, *beta, c, 1, *ldc);                                  \
}                                                                                      \
                                                                                       \
void ch##syr2k_(const char* uploc, const char* trans, const f77_int* m,               \
                const f77_int* k, const T* alpha, const T* a, const f77_int* lda,      \
                const T* b, const f77_int* ldb, const T* beta, T* c,                   \
                const f77_int* ldc)                                                    \
{                                                                                      \
    const char u = up(*uploc), t = up(*trans);                                         \
    const f77_int nrowa = (t == 'N') ? *m : *k;                                        \
    f77_int info = 0;                                                                  \
    if (u != 'L' && u != 'U') info = 1;                                                \
    else if (t != 'N' && t != 'T' && t != 'C') info = 2;                               \
    else if (*m < 0) info = 3;                                                         \
    else if (*k < 0) info = 4;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 7;                                          \
    else if (*ldb < maxi(1, nrowa)) info = 9;                                          \
    else if (*ldc < maxi(1, *m)) info = 12;                                            \
    if (info) { err(CH "SYR2K", info); return; }                                       \
    const int tr = t == 'N' ? TR_N : TR_T;                                             \
    oracle_##ch##syr2k(u == 'L' ? 0xC0 : 0x60, tr, tr, *m, *k, *alpha, a, 1, *lda,     \
                       b, 1, *ldb, *beta, c, 1, *ldc);                                 \
}

DEFINE_F77_STRUC(double, d, "D")
DEFINE_F77_STRUC(float, s, "S")

/* Complex GEMM over the 1m oracle; the checks restate bla_gemm_check.h:39-82
   (a 'C' transpose conjugates). The complex types are (re, im) pairs. */
#define DEFINE_F77_ZGEMM(R, ch, CH)                                                    \
void ch##gemm_(const char* transa, const char* transb, const f77_int* m,              \
               const f77_int* n, const f77_int* k, const R* alpha, const R* a,        \
               const f77_int* lda, const R* b, const f77_int* ldb, const R* beta,     \
               R* c, const f77_int* ldc)                                               \
{                                                                                      \
    const char ta = up(*transa), tb = up(*transb);                                     \
    const int nota = ta == 'N', notb = tb == 'N';                                      \
    const f77_int nrowa = nota ? *m : *k, nrowb = notb ? *k : *n;                      \
    f77_int info = 0;                                                                  \
    if (!nota && ta != 'C' && ta != 'T') info = 1;                                     \
    else if (!notb && tb != 'C' && tb != 'T') info = 2;                                \
    else if (*m < 0) info = 3;                                                         \
    else if (*n < 0) info = 4;                                                         \
    else if (*k < 0) info = 5;                                                         \
    else if (*lda < maxi(1, nrowa)) info = 8;                                          \
    else if (*ldb < maxi(1, nrowb)) info = 10;                                         \
    else if (*ldc < maxi(1, *m)) info = 13;                                            \
    if (info) { err(CH "GEMM ", info); return; }                                       \
    const int tra = nota ? 0 : (ta == 'T' ? 0x8 : 0x18);                               \
    const int trb = notb ? 0 : (tb == 'T' ? 0x8 : 0x18);                               \
    oracle_##ch##gemm(tra, trb, *m, *n, *k, alpha[0], alpha[1], a, 1, *lda, b, 1, *ldb, \
                      beta[0], beta[1], c, 1, *ldc);                                   \
}

DEFINE_F77_ZGEMM(double, z, "Z")
DEFINE_F77_ZGEMM(float, c, "C")
