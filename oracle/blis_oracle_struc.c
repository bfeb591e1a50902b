/*
 * blis_oracle_struc.c — CPU restatement of BLIS's structured level-3
 * operations: gemmt, syrk, syr2k, symm, trmm, trmm3.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the GPU library in
 * tests/ and the oracle behind the netlib ?blat3 drivers (blas_f77.c). The
 * product library never links or calls it.
 *
 * Restated reference code (paths relative to the BLIS source tree):
 *   gemmt        frame/3/gemmt/bli_gemmt_l_ker_var2.c:260-360 and
 *                bli_gemmt_u_ker_var2.c: micro-tiles wholly in the stored
 *                triangle take the gemm microkernel with beta; micro-tiles
 *                the diagonal crosses are computed into a temporary
 *                (ct := alpha*A*B) and only their stored elements are
 *                updated, c := ct + beta*c (bli_txpbys_mxn_uplo). Both
 *                forms give c := alpha*ab + beta*c per element, which is what
 *                computing the full product into a copy of C and writing
 *                back its stored triangle does here.
 *   syrk, syr2k  frame/3/bli_l3_oapi_ex.c:260-330: gemmt(A, A^T), and
 *                gemmt(A, B^T) followed by gemmt(B, A^T) with beta = 1.
 *   symm         frame/1m/packm/bli_packm_struc_cxk.c:38-306: the symmetric
 *                operand is packed densely, its unstored triangle mirrored
 *                from the stored one; the rest is GEMM
 *                (frame/3/bli_l3_oapi_ex.c:400-470 swaps A and B for the
 *                right side).
 *   trmm, trmm3  bli_packm_struc_cxk.c zero-fills the unstored triangle and
 *                writes ones on a unit diagonal (the diagonal is not read);
 *                frame/3/trmm/bli_trmm_ll_ker_var2.c skips k ranges that
 *                are zero, which only drops exact zero terms. trmm writes the
 *                product over B (computed from a copy of B).
 * Triangle convention (frame/include/bli_param_macro_defs.h:638-673):
 * element (i, j) is stored when j - i <= doff (lower) or j - i >= doff
 * (upper); the typed interfaces use doff = 0.
 */
#include "blis_oracle.h"

#include <stdlib.h>
#include <string.h>

#define TRANS_BIT 0x8
#define UPLO_LOWER 0xC0
#define DIAG_UNIT 0x100
#define SIDE_LEFT 0

static int stored(int lower, int64_t doff, int64_t i, int64_t j)
{
    return lower ? (j - i <= doff) : (j - i >= doff);
}

#define DEFINE_STRUC(T, ch)                                                            \
void oracle_##ch##gemmt(int uploc, int transa, int transb, int64_t m, int64_t k,       \
                        T alpha, const T* a, int64_t rs_a, int64_t cs_a,               \
                        const T* b, int64_t rs_b, int64_t cs_b, T beta,                \
                        T* c, int64_t rs_c, int64_t cs_c)                              \
{                                                                                      \
    if (m <= 0) return;                                                                \
    const int lower = uploc == UPLO_LOWER;                                             \
    T* t = (T*)malloc(sizeof(T) * (size_t)m * (size_t)m);                              \
    for (int64_t j = 0; j < m; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i)                                                \
            t[i + j * m] = stored(lower, 0, i, j) ? c[i * rs_c + j * cs_c] : (T)0;     \
    oracle_##ch##gemm(transa, transb, m, m, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,    \
                      beta, t, 1, m);                                                  \
    for (int64_t j = 0; j < m; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i)                                                \
            if (stored(lower, 0, i, j)) c[i * rs_c + j * cs_c] = t[i + j * m];         \
    free(t);                                                                           \
}                                                                                      \
                                                                                       \
void oracle_##ch##syrk(int uploc, int transa, int64_t m, int64_t k, T alpha,           \
                       const T* a, int64_t rs_a, int64_t cs_a, T beta,                 \
                       T* c, int64_t rs_c, int64_t cs_c)                               \
{                                                                                      \
    oracle_##ch##gemmt(uploc, transa, transa ^ TRANS_BIT, m, k, alpha, a, rs_a, cs_a,  \
                       a, rs_a, cs_a, beta, c, rs_c, cs_c);                            \
}                                                                                      \
                                                                                       \
void oracle_##ch##syr2k(int uploc, int transa, int transb, int64_t m, int64_t k,       \
                        T alpha, const T* a, int64_t rs_a, int64_t cs_a,               \
                        const T* b, int64_t rs_b, int64_t cs_b, T beta,                \
                        T* c, int64_t rs_c, int64_t cs_c)                              \
{                                                                                      \
    oracle_##ch##gemmt(uploc, transa, transb ^ TRANS_BIT, m, k, alpha, a, rs_a, cs_a,  \
                       b, rs_b, cs_b, beta, c, rs_c, cs_c);                            \
    oracle_##ch##gemmt(uploc, transb, transa ^ TRANS_BIT, m, k, alpha, b, rs_b, cs_b,  \
                       a, rs_a, cs_a, (T)1, c, rs_c, cs_c);                            \
}                                                                                      \
                                                                                       \
/* Dense n x n copy (column-major) of a structured matrix. */                          \
static T* ch##_densify(int sym, int lower, int unit, int64_t doff, int64_t n,         \
                       const T* a, int64_t rs_a, int64_t cs_a)                         \
{                                                                                      \
    T* d = (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1) * (size_t)(n > 0 ? n : 1));  \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t i = 0; i < n; ++i) {                                              \
            T v;                                                                       \
            if (stored(lower, doff, i, j))                                             \
                v = (unit && j - i == doff) ? (T)1 : a[i * rs_a + j * cs_a];           \
            else                                                                       \
                v = sym ? a[j * rs_a + i * cs_a] : (T)0;                               \
            d[i + j * n] = v;                                                          \
        }                                                                              \
    return d;                                                                          \
}                                                                                      \
                                                                                       \
void oracle_##ch##symm(int side, int uploa, int transb, int64_t m, int64_t n,         \
                       T alpha, const T* a, int64_t rs_a, int64_t cs_a,                \
                       const T* b, int64_t rs_b, int64_t cs_b, T beta,                 \
                       T* c, int64_t rs_c, int64_t cs_c)                               \
{                                                                                      \
    if (m <= 0 || n <= 0) return;                                                      \
    const int64_t na = (side == SIDE_LEFT) ? m : n;                                    \
    T* d = ch##_densify(1, uploa == UPLO_LOWER, 0, 0, na, a, rs_a, cs_a);              \
    if (side == SIDE_LEFT)                                                             \
        oracle_##ch##gemm(0, transb, m, n, m, alpha, d, 1, na, b, rs_b, cs_b, beta,    \
                          c, rs_c, cs_c);                                              \
    else                                                                               \
        oracle_##ch##gemm(transb, 0, m, n, n, alpha, b, rs_b, cs_b, d, 1, na, beta,    \
                          c, rs_c, cs_c);                                              \
    free(d);                                                                           \
}                                                                                      \
                                                                                       \
void oracle_##ch##trmm3(int side, int uploa, int transa, int diaga, int transb,       \
                        int64_t m, int64_t n, T alpha, const T* a, int64_t rs_a,       \
                        int64_t cs_a, const T* b, int64_t rs_b, int64_t cs_b, T beta,  \
                        T* c, int64_t rs_c, int64_t cs_c)                              \
{                                                                                      \
    if (m <= 0 || n <= 0) return;                                                      \
    const int64_t na = (side == SIDE_LEFT) ? m : n;                                    \
    T* d = ch##_densify(0, uploa == UPLO_LOWER, diaga == DIAG_UNIT, 0, na, a, rs_a, cs_a); \
    if (side == SIDE_LEFT)                                                             \
        oracle_##ch##gemm(transa, transb, m, n, m, alpha, d, 1, na, b, rs_b, cs_b,     \
                          beta, c, rs_c, cs_c);                                        \
    else                                                                               \
        oracle_##ch##gemm(transb, transa, m, n, n, alpha, b, rs_b, cs_b, d, 1, na,     \
                          beta, c, rs_c, cs_c);                                        \
    free(d);                                                                           \
}                                                                                      \
                                                                                       \
void oracle_##ch##trmm(int side, int uploa, int transa, int diaga, int64_t m,         \
                       int64_t n, T alpha, const T* a, int64_t rs_a, int64_t cs_a,     \
                       T* b, int64_t rs_b, int64_t cs_b)                               \
{                                                                                      \
    if (m <= 0 || n <= 0) return;                                                      \
    T* w = (T*)malloc(sizeof(T) * (size_t)m * (size_t)n);                              \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i) w[i + j * m] = b[i * rs_b + j * cs_b];         \
    oracle_##ch##trmm3(side, uploa, transa, diaga, 0, m, n, alpha, a, rs_a, cs_a,      \
                       w, 1, m, (T)0, b, rs_b, cs_b);                                  \
    free(w);                                                                           \
}

DEFINE_STRUC(double, d)
DEFINE_STRUC(float, s)
