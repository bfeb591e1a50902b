/*
 * blis_oracle.c — CPU restatement of BLIS's level-3 hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle against which the GPU
 * library is checked (tests/, __graft_entry__.smoke) and the CPU baseline
 * timed by bench.py. The product library never links or calls it.
 *
 * Restated reference code (paths relative to the BLIS source tree):
 *   five-loop GEMM        sandbox/gemmlike/bls_gemm_bp_var1.c:37-318
 *                         (jc over NC, pc over KC with beta applied only in
 *                         the first pc iteration as frame/3/gemm/
 *                         bli_gemm_blk_var3.c:110-112, ic over MC, jr/ir
 *                         over NR/MR)
 *   panel packing         ref_kernels/1m/bli_packm_cxk_ref.c:78-147
 *                         (MR-row / NR-column micro-panels, zero padding)
 *   microkernel           ref_kernels/3/bli_gemm_ref.c:163-317
 *                         (k rank-1 updates into ab, scale by alpha,
 *                         beta == 0 overwrites C without reading it)
 *   blocksizes            config/zen3/bli_cntx_init_zen3.c:204-208
 *                         (d: MR 6, NR 8, MC 72, KC 256, NC 4080;
 *                          s: MR 6, NR 16, MC 144, KC 256, NC 4080)
 *   trivial cases         frame/3/bli_l3_util.c:41-65
 *   TRSM                  frame/3/bli_l3_oapi_ex.c:753-758 (right side by
 *                         transposition), frame/3/trsm/bli_trsm_blk_var1.c:
 *                         37-189 (diagonal KC block, then GEMM update of the
 *                         rows below), frame/3/trsm/bli_trsm_ll_ker_var2.c:
 *                         240-335 (gemmtrsm on diagonal micro-panels),
 *                         ref_kernels/3/bli_gemmtrsm_ref.c:41-199,
 *                         ref_kernels/3/bli_trsm_ref.c:42-231 (multiply by
 *                         the reciprocal stored at pack time,
 *                         ref_kernels/1m/bli_packm_cxc_diag_ref.c:99-107)
 *   axpyv / dotv          ref_kernels/1/bli_axpyv_ref.c:38-121,
 *                         ref_kernels/1/bli_dotv_ref.c:38-122
 *   gemv                  frame/2/bli_l2_tapi.c:43-123 (variant choice),
 *                         frame/2/gemv/bli_gemv_unf_var1.c (dotxf rows),
 *                         frame/2/gemv/bli_gemv_unf_var2.c (axpyf columns)
 *   gemm_batch            frame/compat/extra/bla_gemm_batch.c:172-241
 */
#include "blis_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TRANS_BIT 0x8
#define UPLO_LOWER 0xC0
#define DIAG_UNIT 0x100
#define SIDE_LEFT 0

/* ---------------------------------------------------------------- GEMM -- */

#define DEFINE_GEMM(T, ch, MR, NR, MC, KC, NC)                                         \
static void ch##_ukr(int64_t k, T alpha, const T* ap, const T* bp, T beta,             \
                     T* c, int64_t rs_c, int64_t cs_c, int64_t m_cur, int64_t n_cur) \
{                                                                                      \
    T ab[MR * NR];                                                                     \
    for (int i = 0; i < MR * NR; ++i) ab[i] = (T)0;                                    \
    for (int64_t l = 0; l < k; ++l) {                                                  \
        for (int i = 0; i < MR; ++i)                                                   \
            for (int j = 0; j < NR; ++j) ab[i * NR + j] += ap[i] * bp[j];              \
        ap += MR;                                                                      \
        bp += NR;                                                                      \
    }                                                                                  \
    for (int i = 0; i < MR * NR; ++i) ab[i] *= alpha;                                  \
    if (beta == (T)0) {                                                                \
        for (int64_t j = 0; j < n_cur; ++j)                                            \
            for (int64_t i = 0; i < m_cur; ++i) c[i * rs_c + j * cs_c] = ab[i * NR + j]; \
    } else {                                                                           \
        for (int64_t j = 0; j < n_cur; ++j)                                            \
            for (int64_t i = 0; i < m_cur; ++i)                                        \
                c[i * rs_c + j * cs_c] = ab[i * NR + j] + beta * c[i * rs_c + j * cs_c]; \
    }                                                                                  \
}                                                                                      \
                                                                                       \
/* A_eff(i, p) = a[i*sa_i + p*sa_p] for rows [i0, i0+mc), cols [p0, p0+kc). */         \
static void ch##_packa(const T* a, int64_t sa_i, int64_t sa_p, int64_t i0, int64_t mc, \
                       int64_t p0, int64_t kc, T* ap)                                  \
{                                                                                      \
    for (int64_t ir = 0; ir < mc; ir += MR) {                                          \
        const int64_t mr = (mc - ir < MR) ? mc - ir : MR;                              \
        for (int64_t l = 0; l < kc; ++l) {                                             \
            for (int64_t i = 0; i < MR; ++i)                                           \
                ap[l * MR + i] = (i < mr) ? a[(i0 + ir + i) * sa_i + (p0 + l) * sa_p] : (T)0; \
        }                                                                              \
        ap += MR * kc;                                                                 \
    }                                                                                  \
}                                                                                      \
                                                                                       \
static void ch##_packb(const T* b, int64_t sb_p, int64_t sb_j, int64_t p0, int64_t kc, \
                       int64_t j0, int64_t nc, T* bp)                                  \
{                                                                                      \
    for (int64_t jr = 0; jr < nc; jr += NR) {                                          \
        const int64_t nr = (nc - jr < NR) ? nc - jr : NR;                              \
        for (int64_t l = 0; l < kc; ++l) {                                             \
            for (int64_t j = 0; j < NR; ++j)                                           \
                bp[l * NR + j] = (j < nr) ? b[(p0 + l) * sb_p + (j0 + jr + j) * sb_j] : (T)0; \
        }                                                                              \
        bp += NR * kc;                                                                 \
    }                                                                                  \
}                                                                                      \
                                                                                       \
static void ch##_scalm(int64_t m, int64_t n, T beta, T* c, int64_t rs_c, int64_t cs_c) \
{                                                                                      \
    for (int64_t j = 0; j < n; ++j)                                                    \
        for (int64_t i = 0; i < m; ++i) {                                              \
            T* p = c + i * rs_c + j * cs_c;                                            \
            *p = (beta == (T)0) ? (T)0 : beta * *p;                                    \
        }                                                                              \
}                                                                                      \
                                                                                       \
static void ch##_gemm_impl(int transa, int transb, int64_t m, int64_t n, int64_t k,    \
                           T alpha, const T* a, int64_t rs_a, int64_t cs_a,            \
                           const T* b, int64_t rs_b, int64_t cs_b, T beta,             \
                           T* c, int64_t rs_c, int64_t cs_c, int nthreads)             \
{                                                                                      \
    if (m <= 0 || n <= 0) return;                                                      \
    if (alpha == (T)0 || k <= 0) { ch##_scalm(m, n, beta, c, rs_c, cs_c); return; }    \
    const int ta = (transa & TRANS_BIT) != 0, tb = (transb & TRANS_BIT) != 0;          \
    const int64_t sa_i = ta ? cs_a : rs_a, sa_p = ta ? rs_a : cs_a;                    \
    const int64_t sb_p = tb ? cs_b : rs_b, sb_j = tb ? rs_b : cs_b;                    \
    T* bp = (T*)malloc(sizeof(T) * (size_t)KC * (size_t)(NC + NR));                    \
    for (int64_t jc = 0; jc < n; jc += NC) {                                           \
        const int64_t nc = (n - jc < NC) ? n - jc : NC;                                \
        for (int64_t pc = 0; pc < k; pc += KC) {                                       \
            const int64_t kc = (k - pc < KC) ? k - pc : KC;                            \
            const T beta_use = (pc == 0) ? beta : (T)1;                                \
            ch##_packb(b, sb_p, sb_j, pc, kc, jc, nc, bp);                             \
            const int64_t n_ic = (m + MC - 1) / MC;                                    \
            _Pragma("omp parallel num_threads(nthreads) if(nthreads > 1)")             \
            {                                                                          \
                T* ap = (T*)malloc(sizeof(T) * (size_t)KC * (size_t)(MC + MR));        \
                _Pragma("omp for schedule(static)")                                    \
                for (int64_t icb = 0; icb < n_ic; ++icb) {                             \
                    const int64_t ic = icb * MC;                                       \
                    const int64_t mc = (m - ic < MC) ? m - ic : MC;                    \
                    ch##_packa(a, sa_i, sa_p, ic, mc, pc, kc, ap);                     \
                    for (int64_t jr = 0; jr < nc; jr += NR) {                          \
                        const int64_t nr = (nc - jr < NR) ? nc - jr : NR;              \
                        for (int64_t ir = 0; ir < mc; ir += MR) {                      \
                            const int64_t mr = (mc - ir < MR) ? mc - ir : MR;          \
                            T* cij = c + (ic + ir) * rs_c + (jc + jr) * cs_c;          \
                            ch##_ukr(kc, alpha, ap + (ir / MR) * MR * kc,              \
                                     bp + (jr / NR) * NR * kc, beta_use, cij, rs_c,    \
                                     cs_c, mr, nr);                                    \
                        }                                                              \
                    }                                                                  \
                }                                                                      \
                free(ap);                                                              \
            }                                                                          \
        }                                                                              \
    }                                                                                  \
    free(bp);                                                                          \
}                                                                                      \
                                                                                       \
void oracle_##ch##gemm(int transa, int transb, int64_t m, int64_t n, int64_t k,        \
                       T alpha, const T* a, int64_t rs_a, int64_t cs_a,                \
                       const T* b, int64_t rs_b, int64_t cs_b, T beta,                 \
                       T* c, int64_t rs_c, int64_t cs_c)                               \
{                                                                                      \
    ch##_gemm_impl(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,       \
                   beta, c, rs_c, cs_c, 1);                                            \
}                                                                                      \
                                                                                       \
void oracle_##ch##gemm_mt(int transa, int transb, int64_t m, int64_t n, int64_t k,     \
                          T alpha, const T* a, int64_t rs_a, int64_t cs_a,             \
                          const T* b, int64_t rs_b, int64_t cs_b, T beta,              \
                          T* c, int64_t rs_c, int64_t cs_c, int nthreads)              \
{                                                                                      \
    ch##_gemm_impl(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,       \
                   beta, c, rs_c, cs_c, nthreads < 1 ? 1 : nthreads);                  \
}

DEFINE_GEMM(double, d, 6, 8, 72, 256, 4080)
DEFINE_GEMM(float, s, 6, 16, 144, 256, 4080)

/* ---------------------------------------------------------------- TRSM -- */

/* Left-side solve L_eff * X = alpha * B' in place, L_eff(i,j) =
   a[i*ars + j*acs], lower or upper, B'(i,j) = b[i*brs + j*bcs].
   Right-looking over KC diagonal blocks (bli_trsm_blk_var1); inside a block,
   left-looking over MR micro-panels with the gemmtrsm microkernel. */
#define DEFINE_TRSM(T, ch, MR, KC)                                                     \
static void ch##_trsm_left(int lower, int unit, int64_t m, int64_t n, T alpha,         \
                           const T* a, int64_t ars, int64_t acs,                       \
                           T* b, int64_t brs, int64_t bcs)                             \
{                                                                                      \
    const int64_t nblk = (m + KC - 1) / KC;                                            \
    for (int64_t bi = 0; bi < nblk; ++bi) {                                            \
        /* Diagonal block rows [k0, k0+kc) in solve order. */                          \
        const int64_t blk = lower ? bi : (nblk - 1 - bi);                              \
        const int64_t k0 = blk * KC;                                                   \
        const int64_t kc = (m - k0 < KC) ? m - k0 : KC;                                \
        const T alpha_blk = (bi == 0) ? alpha : (T)1;                                  \
        const int64_t npan = (kc + MR - 1) / MR;                                       \
        for (int64_t pi = 0; pi < npan; ++pi) {                                        \
            const int64_t pan = lower ? pi : (npan - 1 - pi);                          \
            const int64_t r0 = k0 + pan * MR;                                          \
            const int64_t mr = (k0 + kc - r0 < MR) ? k0 + kc - r0 : MR;                \
            /* gemm part: B11 := alpha*B11 - A10*B01 (A10 within this block). */       \
            const int64_t s0 = lower ? k0 : r0 + mr;                                   \
            const int64_t s1 = lower ? r0 : k0 + kc;                                   \
            for (int64_t j = 0; j < n; ++j)                                            \
                for (int64_t i = 0; i < mr; ++i) {                                     \
                    T acc = (T)0;                                                      \
                    for (int64_t l = s0; l < s1; ++l)                                  \
                        acc += a[(r0 + i) * ars + l * acs] * b[l * brs + j * bcs];     \
                    T* p = b + (r0 + i) * brs + j * bcs;                               \
                    *p = alpha_blk * *p - acc;                                         \
                }                                                                      \
            /* trsm part: multiply by the pre-inverted diagonal. */                    \
            for (int64_t j = 0; j < n; ++j) {                                          \
                for (int64_t t = 0; t < mr; ++t) {                                     \
                    const int64_t i = lower ? t : (mr - 1 - t);                        \
                    T beta11 = b[(r0 + i) * brs + j * bcs];                            \
                    T rho = (T)0;                                                      \
                    if (lower) {                                                       \
                        for (int64_t l = 0; l < i; ++l)                                \
                            rho += a[(r0 + i) * ars + (r0 + l) * acs] * b[(r0 + l) * brs + j * bcs]; \
                    } else {                                                           \
                        for (int64_t l = i + 1; l < mr; ++l)                           \
                            rho += a[(r0 + i) * ars + (r0 + l) * acs] * b[(r0 + l) * brs + j * bcs]; \
                    }                                                                  \
                    beta11 -= rho;                                                     \
                    const T inv = unit ? (T)1 : (T)1 / a[(r0 + i) * ars + (r0 + i) * acs]; \
                    beta11 *= inv;                                                     \
                    b[(r0 + i) * brs + j * bcs] = beta11;                              \
                }                                                                      \
            }                                                                          \
        }                                                                              \
        /* GEMM update of the rows not yet solved: beta = alpha_blk. */                \
        const int64_t u0 = lower ? k0 + kc : 0;                                        \
        const int64_t u1 = lower ? m : k0;                                             \
        if (u1 > u0) {                                                                 \
            oracle_##ch##gemm(0, 0, u1 - u0, n, kc, (T)-1, a + u0 * ars + k0 * acs, ars, acs, \
                              b + k0 * brs, brs, bcs, alpha_blk, b + u0 * brs, brs, bcs); \
        }                                                                              \
    }                                                                                  \
}                                                                                      \
                                                                                       \
void oracle_##ch##trsm(int side, int uplo, int transa, int diag, int64_t m, int64_t n, \
                       T alpha, const T* a, int64_t rs_a, int64_t cs_a,                \
                       T* b, int64_t rs_b, int64_t cs_b)                               \
{                                                                                      \
    if (m <= 0 || n <= 0) return;                                                      \
    if (alpha == (T)0) { ch##_scalm(m, n, (T)0, b, rs_b, cs_b); return; }              \
    const int t = (transa & TRANS_BIT) != 0;                                           \
    const int lower = (uplo == UPLO_LOWER);                                            \
    const int unit = (diag == DIAG_UNIT);                                              \
    if (side == SIDE_LEFT) {                                                           \
        ch##_trsm_left(t ? !lower : lower, unit, m, n, alpha, a,                       \
                       t ? cs_a : rs_a, t ? rs_a : cs_a, b, rs_b, cs_b);               \
    } else {                                                                           \
        ch##_trsm_left(t ? lower : !lower, unit, n, m, alpha, a,                       \
                       t ? rs_a : cs_a, t ? cs_a : rs_a, b, cs_b, rs_b);               \
    }                                                                                  \
}

DEFINE_TRSM(double, d, 6, 256)
DEFINE_TRSM(float, s, 6, 256)

/* ------------------------------------------------------ level-1v / 2 -- */

#define DEFINE_L12(T, ch)                                                              \
void oracle_##ch##axpyv(int64_t n, T alpha, const T* x, int64_t incx, T* y, int64_t incy) \
{                                                                                      \
    if (n <= 0 || alpha == (T)0) return;                                               \
    for (int64_t i = 0; i < n; ++i) y[i * incy] += alpha * x[i * incx];                \
}                                                                                      \
                                                                                       \
T oracle_##ch##dotv(int64_t n, const T* x, int64_t incx, const T* y, int64_t incy)     \
{                                                                                      \
    T rho = (T)0;                                                                      \
    for (int64_t i = 0; i < n; ++i) rho += x[i * incx] * y[i * incy];                  \
    return rho;                                                                        \
}                                                                                      \
                                                                                       \
void oracle_##ch##gemv(int transa, int64_t m, int64_t n, T alpha,                      \
                       const T* a, int64_t rs_a, int64_t cs_a,                         \
                       const T* x, int64_t incx, T beta, T* y, int64_t incy)           \
{                                                                                      \
    const int t = (transa & TRANS_BIT) != 0;                                           \
    const int64_t my = t ? n : m, nx = t ? m : n;                                      \
    if (my <= 0) return;                                                               \
    if (nx <= 0 || alpha == (T)0) {                                                    \
        for (int64_t i = 0; i < my; ++i)                                               \
            y[i * incy] = (beta == (T)0) ? (T)0 : beta * y[i * incy];                  \
        return;                                                                        \
    }                                                                                  \
    /* op(A)(i, j) = a[i*so + j*si] */                                                 \
    const int64_t so = t ? cs_a : rs_a, si = t ? rs_a : cs_a;                          \
    const int row_stored = (cs_a == 1 || cs_a == -1) && !(rs_a == 1 || rs_a == -1);    \
    const int use_var1 = t ? !row_stored : row_stored;                                 \
    if (use_var1) {                                                                    \
        /* dotxf over rows: y_i := beta*y_i + alpha * sum_p op(A)(i,p) x_p */          \
        for (int64_t i = 0; i < my; ++i) {                                             \
            T yi = (beta == (T)0) ? (T)0 : beta * y[i * incy];                         \
            T r = (T)0;                                                                \
            for (int64_t p = 0; p < nx; ++p) r += a[i * so + p * si] * x[p * incx];    \
            y[i * incy] = yi + alpha * r;                                              \
        }                                                                              \
    } else {                                                                           \
        /* axpyf over columns: y := beta*y; y += (alpha*x_j) * op(A)(:, j) */          \
        for (int64_t i = 0; i < my; ++i)                                               \
            y[i * incy] = (beta == (T)0) ? (T)0 : beta * y[i * incy];                  \
        for (int64_t j = 0; j < nx; ++j) {                                             \
            const T axj = alpha * x[j * incx];                                         \
            for (int64_t i = 0; i < my; ++i) y[i * incy] += axj * a[i * so + j * si];  \
        }                                                                              \
    }                                                                                  \
}                                                                                      \
                                                                                       \
void oracle_##ch##gemm_batch(const int* transa, const int* transb,                     \
                             const int64_t* m, const int64_t* n, const int64_t* k,     \
                             const T* alpha, const T* const* a, const int64_t* lda,    \
                             const T* const* b, const int64_t* ldb, const T* beta,     \
                             T* const* c, const int64_t* ldc,                          \
                             int64_t group_count, const int64_t* group_size)           \
{                                                                                      \
    int64_t idx = 0;                                                                   \
    for (int64_t g = 0; g < group_count; ++g)                                          \
        for (int64_t s = 0; s < group_size[g]; ++s, ++idx)                             \
            oracle_##ch##gemm(transa[g], transb[g], m[g], n[g], k[g], alpha[g],        \
                              a[idx], 1, lda[g], b[idx], 1, ldb[g], beta[g],           \
                              c[idx], 1, ldc[g]);                                      \
}

DEFINE_L12(double, d)
DEFINE_L12(float, s)

/* ------------------------------------------------------ double-double -- */

/* Error-free transformations: Knuth's TwoSum and an FMA-based TwoProd
   (Ogita, Rump, Oishi, "Accurate sum and dot product", SISC 2005). */
static inline void two_sum(double a, double b, double* s, double* e)
{
    const double x = a + b;
    const double z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

static inline void two_prod(double a, double b, double* p, double* e)
{
    const double x = a * b;
    *e = fma(a, b, -x);
    *p = x;
}

/* (hi, lo) += (p_hi, p_lo) */
static inline void dd_add(double* hi, double* lo, double p_hi, double p_lo)
{
    double s, e;
    two_sum(*hi, p_hi, &s, &e);
    e += *lo + p_lo;
    two_sum(s, e, hi, lo);
}

void oracle_dgemm_dd_rows(int transa, int transb, int64_t m, int64_t n, int64_t k,
                          double alpha, const double* a, int64_t rs_a, int64_t cs_a,
                          const double* b, int64_t rs_b, int64_t cs_b,
                          double beta, double* c, int64_t rs_c, int64_t cs_c,
                          int64_t r0, int64_t r1, int nthreads)
{
    const int ta = (transa & TRANS_BIT) != 0, tb = (transb & TRANS_BIT) != 0;
    const int64_t sa_i = ta ? cs_a : rs_a, sa_p = ta ? rs_a : cs_a;
    const int64_t sb_p = tb ? cs_b : rs_b, sb_j = tb ? rs_b : cs_b;
    if (r1 > m) r1 = m;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) if(nthreads > 1)
    for (int64_t i = r0; i < r1; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            double hi = 0.0, lo = 0.0;
            for (int64_t p = 0; p < k; ++p) {
                double ph, pl;
                two_prod(a[i * sa_i + p * sa_p], b[p * sb_p + j * sb_j], &ph, &pl);
                dd_add(&hi, &lo, ph, pl);
            }
            /* alpha * (hi, lo) */
            double ah, al;
            two_prod(alpha, hi, &ah, &al);
            al += alpha * lo;
            double* cp = c + i * rs_c + j * cs_c;
            if (beta != 0.0) {
                double bh, bl;
                two_prod(beta, *cp, &bh, &bl);
                dd_add(&ah, &al, bh, bl);
            }
            *cp = ah + al;
        }
    }
}
