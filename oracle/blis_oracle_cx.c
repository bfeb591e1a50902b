/*
 * blis_oracle_cx.c — CPU restatement of BLIS's complex GEMM by the 1m
 * induced method.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the GPU library's complex
 * GEMM in tests/ and the oracle behind the netlib ?blat3 complex drivers
 * (blas_f77.c). The product library never links or calls it.
 *
 * Restated reference code (paths relative to the BLIS source tree):
 *   1m method   frame/3/bli_l3_ind.c (induced-method dispatch for complex
 *               level-3), frame/1m/packm/bli_packm_cxk_1er.c (packing
 *               formats): A is packed "1e" — element a becomes the real
 *               block [[ar, -ai], [ai, ar]] — and B "1r" — element b becomes
 *               (br, bi) down a column; the real five-loop GEMM of the
 *               2m x 2k and 2k x n packed operands yields C with real and
 *               imaginary parts interleaved down its columns.
 *   scalars     alpha scales A while it is packed; beta scales C first
 *               (C := beta*C), after which the real product accumulates
 *               with beta = 1; beta == 0 overwrites C without reading it
 *               (ref_kernels/3/bli_gemm_ref.c:257-310).
 * Transposition flags: 0x8 transposes, 0x10 conjugates (BLIS encodings).
 */
#include "blis_oracle.h"

#include <stdlib.h>
#include <string.h>

#define TRANS_BIT 0x8
#define CONJ_BIT 0x10

#define DEFINE_ZGEMM(R, ch, rch)                                                       \
void oracle_##ch##gemm(int transa, int transb, int64_t m, int64_t n, int64_t k,        \
                       R alpha_re, R alpha_im, const R* a, int64_t rs_a, int64_t cs_a, \
                       const R* b, int64_t rs_b, int64_t cs_b, R beta_re, R beta_im,   \
                       R* c, int64_t rs_c, int64_t cs_c)                               \
{                                                                                      \
    if (m <= 0 || n <= 0) return;                                                      \
    /* C := beta*C (complex; beta == 0 writes zeros without reading C). */             \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i) {                                              \
            R* p = c + 2 * (i * rs_c + j * cs_c);                                      \
            if (beta_re == (R)0 && beta_im == (R)0) { p[0] = (R)0; p[1] = (R)0; }      \
            else {                                                                     \
                const R cr = p[0], ci = p[1];                                          \
                p[0] = beta_re * cr - beta_im * ci;                                    \
                p[1] = beta_re * ci + beta_im * cr;                                    \
            }                                                                          \
        }                                                                              \
    if (k <= 0 || (alpha_re == (R)0 && alpha_im == (R)0)) return;                      \
    const int ta = (transa & TRANS_BIT) != 0, ca = (transa & CONJ_BIT) != 0;           \
    const int tb = (transb & TRANS_BIT) != 0, cb = (transb & CONJ_BIT) != 0;           \
    const int64_t sa_i = ta ? cs_a : rs_a, sa_p = ta ? rs_a : cs_a;                    \
    const int64_t sb_p = tb ? cs_b : rs_b, sb_j = tb ? rs_b : cs_b;                    \
    /* 1e: 2m x 2k column-major; 1r: 2k x n column-major; C: 2m x n. */                \
    R* ae = (R*)malloc(sizeof(R) * (size_t)(4 * m * k));                               \
    R* br = (R*)malloc(sizeof(R) * (size_t)(2 * k * n));                               \
    R* ct = (R*)malloc(sizeof(R) * (size_t)(2 * m * n));                               \
    for (int64_t p = 0; p < k; ++p)                                                    \
        for (int64_t i = 0; i < m; ++i) {                                              \
            const R* x = a + 2 * (i * sa_i + p * sa_p);                                \
            const R xr = x[0], xi = ca ? -x[1] : x[1];                                 \
            const R vr = alpha_re * xr - alpha_im * xi;                                \
            const R vi = alpha_re * xi + alpha_im * xr;                                \
            ae[(2 * i) + (2 * p) * (2 * m)] = vr;                                      \
            ae[(2 * i + 1) + (2 * p) * (2 * m)] = vi;                                  \
            ae[(2 * i) + (2 * p + 1) * (2 * m)] = -vi;                                 \
            ae[(2 * i + 1) + (2 * p + 1) * (2 * m)] = vr;                              \
        }                                                                              \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t p = 0; p < k; ++p) {                                              \
            const R* y = b + 2 * (p * sb_p + j * sb_j);                                \
            br[2 * p + j * (2 * k)] = y[0];                                            \
            br[2 * p + 1 + j * (2 * k)] = cb ? -y[1] : y[1];                           \
        }                                                                              \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i) {                                              \
            const R* p = c + 2 * (i * rs_c + j * cs_c);                                \
            ct[2 * i + j * (2 * m)] = p[0];                                            \
            ct[2 * i + 1 + j * (2 * m)] = p[1];                                        \
        }                                                                              \
    oracle_##rch##gemm(0, 0, 2 * m, n, 2 * k, (R)1, ae, 1, 2 * m, br, 1, 2 * k, (R)1,  \
                       ct, 1, 2 * m);                                                  \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i) {                                              \
            R* p = c + 2 * (i * rs_c + j * cs_c);                                      \
            p[0] = ct[2 * i + j * (2 * m)];                                            \
            p[1] = ct[2 * i + 1 + j * (2 * m)];                                        \
        }                                                                              \
    free(ae); free(br); free(ct);                                                      \
}

DEFINE_ZGEMM(double, z, d)
DEFINE_ZGEMM(float, c, s)
