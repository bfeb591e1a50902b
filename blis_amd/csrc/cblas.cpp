// cblas.cpp — the CBLAS layer over this library's BLAS symbols.
//
// Same mapping as the reference's CBLAS (frame/compat/cblas/src/
// cblas_?gemm.c, cblas_?symm.c, cblas_?syrk.c, cblas_?syr2k.c,
// cblas_?trmm.c, cblas_?trsm.c, cblas_?gemv.c, cblas_?axpy.c, cblas_?dot.c,
// extra/cblas_?gemm_batch.c, extra/cblas_?gemmt.c): a column-major call
// passes straight through to the Fortran-style symbol; a row-major call is
// the column-major call on the transposed problem:
//   gemm, gemmt, gemm_batch   C^T = op(B)^T op(A)^T: A <-> B, M <-> N
//                             (gemmt: uplo flips);
//   symm                      side and uplo flip, M <-> N;
//   syrk, syr2k               uplo flips, trans toggles (N <-> T; C -> N);
//   trmm, trsm                side and uplo flip, M <-> N, trans kept;
//   gemv                      trans toggles, M <-> N.
// An enum argument outside its range goes to cblas_xerbla with the
// parameter position the reference reports (cblas_xerbla.c), which prints
// the message and ends the process with exit(-1).
#include "api_common.hpp"

#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cblas.h"

#define BLIS_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

// Set while a row-major call runs, as RowMajorStrg in cblas_globals.c:
// cblas_xerbla renumbers some parameters for row-major callers.
thread_local int g_row_major = 0;

struct RowMajorScope {
    explicit RowMajorScope(bool rm) { g_row_major = rm ? 1 : 0; }
    ~RowMajorScope() { g_row_major = 0; }
};

bool trans_char(CBLAS_TRANSPOSE t, char* c) {
    switch (t) {
    case CblasNoTrans: *c = 'N'; return true;
    case CblasTrans: *c = 'T'; return true;
    case CblasConjTrans: *c = 'C'; return true;
    default: return false;
    }
}

// Row-major syrk/syr2k on real types: N <-> T, and C (= T for real) -> N.
bool trans_char_toggled(CBLAS_TRANSPOSE t, char* c) {
    switch (t) {
    case CblasNoTrans: *c = 'T'; return true;
    case CblasTrans:
    case CblasConjTrans: *c = 'N'; return true;
    default: return false;
    }
}

bool uplo_char(CBLAS_UPLO u, bool flip, char* c) {
    if (u == CblasUpper) { *c = flip ? 'L' : 'U'; return true; }
    if (u == CblasLower) { *c = flip ? 'U' : 'L'; return true; }
    return false;
}

bool side_char(CBLAS_SIDE s, bool flip, char* c) {
    if (s == CblasLeft) { *c = flip ? 'R' : 'L'; return true; }
    if (s == CblasRight) { *c = flip ? 'L' : 'R'; return true; }
    return false;
}

bool diag_char(CBLAS_DIAG d, char* c) {
    if (d == CblasUnit) { *c = 'U'; return true; }
    if (d == CblasNonUnit) { *c = 'N'; return true; }
    return false;
}

bool order_ok(CBLAS_ORDER o) { return o == CblasRowMajor || o == CblasColMajor; }

} // namespace

BLIS_EXPORT void cblas_xerbla(f77_int info, const char* rout, const char* form, ...) {
    if (g_row_major) {
        // Parameter positions that differ between the row-major CBLAS call
        // and the column-major Fortran call it becomes (cblas_xerbla.c).
        auto swap = [&](f77_int x, f77_int y) {
            if (info == x) info = y;
            else if (info == y) info = x;
        };
        if (std::strstr(rout, "gemm")) { swap(4, 5); swap(9, 11); }
        else if (std::strstr(rout, "symm") || std::strstr(rout, "hemm")) swap(4, 5);
        else if (std::strstr(rout, "trmm") || std::strstr(rout, "trsm")) swap(6, 7);
        else if (std::strstr(rout, "gemv")) swap(3, 4);
    }
    if (info) std::fprintf(stderr, "Parameter %d to routine %s was incorrect\n", (int)info, rout);
    va_list ap;
    va_start(ap, form);
    std::vfprintf(stderr, form, ap);
    va_end(ap);
    std::fflush(stderr);
    std::exit(-1);
}

// ---- level 1 --------------------------------------------------------------

BLIS_EXPORT float cblas_sdot(f77_int N, const float* X, f77_int incX, const float* Y,
                             f77_int incY) {
    return sdot_(&N, X, &incX, Y, &incY);
}
BLIS_EXPORT double cblas_ddot(f77_int N, const double* X, f77_int incX, const double* Y,
                              f77_int incY) {
    return ddot_(&N, X, &incX, Y, &incY);
}
BLIS_EXPORT void cblas_saxpy(f77_int N, float alpha, const float* X, f77_int incX, float* Y,
                             f77_int incY) {
    saxpy_(&N, &alpha, X, &incX, Y, &incY);
}
BLIS_EXPORT void cblas_daxpy(f77_int N, double alpha, const double* X, f77_int incX, double* Y,
                             f77_int incY) {
    daxpy_(&N, &alpha, X, &incX, Y, &incY);
}

// ---- level 2 --------------------------------------------------------------

#define CBLAS_GEMV(ch, T)                                                              \
BLIS_EXPORT void cblas_##ch##gemv(CBLAS_ORDER order, CBLAS_TRANSPOSE TransA, f77_int M, \
        f77_int N, T alpha, const T* A, f77_int lda, const T* X, f77_int incX, T beta,  \
        T* Y, f77_int incY) {                                                          \
    const char* rout = "cblas_" #ch "gemv";                                            \
    if (!order_ok(order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", order); return; } \
    RowMajorScope rm(order == CblasRowMajor);                                          \
    char ta;                                                                           \
    const bool ok = order == CblasColMajor ? trans_char(TransA, &ta)                   \
                                           : trans_char_toggled(TransA, &ta);          \
    if (!ok) { cblas_xerbla(2, rout, "Illegal TransA setting, %d\n", TransA); return; } \
    if (order == CblasColMajor)                                                        \
        ch##gemv_(&ta, &M, &N, &alpha, A, &lda, X, &incX, &beta, Y, &incY);            \
    else                                                                               \
        ch##gemv_(&ta, &N, &M, &alpha, A, &lda, X, &incX, &beta, Y, &incY);            \
}
CBLAS_GEMV(s, float)
CBLAS_GEMV(d, double)

// ---- level 3 --------------------------------------------------------------

#define CBLAS_L3(ch, T)                                                                \
BLIS_EXPORT void cblas_##ch##gemm(CBLAS_ORDER Order, CBLAS_TRANSPOSE TransA,           \
        CBLAS_TRANSPOSE TransB, f77_int M, f77_int N, f77_int K, T alpha, const T* A,  \
        f77_int lda, const T* B, f77_int ldb, T beta, T* C, f77_int ldc) {             \
    const char* rout = "cblas_" #ch "gemm";                                            \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    RowMajorScope rm(Order == CblasRowMajor);                                          \
    char ta, tb;                                                                       \
    if (!trans_char(TransA, &ta)) { cblas_xerbla(2, rout, "Illegal TransA setting, %d\n", TransA); return; } \
    if (!trans_char(TransB, &tb)) {                                                    \
        cblas_xerbla(Order == CblasColMajor ? 3 : 2, rout, "Illegal TransB setting, %d\n", TransB); \
        return;                                                                        \
    }                                                                                  \
    if (Order == CblasColMajor)                                                        \
        ch##gemm_(&ta, &tb, &M, &N, &K, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);     \
    else                                                                               \
        ch##gemm_(&tb, &ta, &N, &M, &K, &alpha, B, &ldb, A, &lda, &beta, C, &ldc);     \
}                                                                                      \
BLIS_EXPORT void cblas_##ch##gemmt(CBLAS_ORDER Order, CBLAS_UPLO Uplo,                 \
        CBLAS_TRANSPOSE TransA, CBLAS_TRANSPOSE TransB, f77_int M, f77_int K, T alpha, \
        const T* A, f77_int lda, const T* B, f77_int ldb, T beta, T* C, f77_int ldc) { \
    const char* rout = "cblas_" #ch "gemmt";                                           \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    const bool row = Order == CblasRowMajor;                                           \
    RowMajorScope rm(row);                                                             \
    char ul, ta, tb;                                                                   \
    if (!uplo_char(Uplo, row, &ul)) { cblas_xerbla(2, rout, "Illegal Uplo setting, %d\n", Uplo); return; } \
    if (!trans_char(TransA, &ta)) { cblas_xerbla(3, rout, "Illegal TransA setting, %d\n", TransA); return; } \
    if (!trans_char(TransB, &tb)) { cblas_xerbla(4, rout, "Illegal TransB setting, %d\n", TransB); return; } \
    if (!row)                                                                          \
        ch##gemmt_(&ul, &ta, &tb, &M, &K, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);   \
    else                                                                               \
        ch##gemmt_(&ul, &tb, &ta, &M, &K, &alpha, B, &ldb, A, &lda, &beta, C, &ldc);   \
}                                                                                      \
BLIS_EXPORT void cblas_##ch##symm(CBLAS_ORDER Order, CBLAS_SIDE Side, CBLAS_UPLO Uplo, \
        f77_int M, f77_int N, T alpha, const T* A, f77_int lda, const T* B,            \
        f77_int ldb, T beta, T* C, f77_int ldc) {                                      \
    const char* rout = "cblas_" #ch "symm";                                            \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    const bool row = Order == CblasRowMajor;                                           \
    RowMajorScope rm(row);                                                             \
    char sd, ul;                                                                       \
    if (!side_char(Side, row, &sd)) { cblas_xerbla(2, rout, "Illegal Side setting, %d\n", Side); return; } \
    if (!uplo_char(Uplo, row, &ul)) { cblas_xerbla(3, rout, "Illegal Uplo setting, %d\n", Uplo); return; } \
    if (!row)                                                                          \
        ch##symm_(&sd, &ul, &M, &N, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);         \
    else                                                                               \
        ch##symm_(&sd, &ul, &N, &M, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);         \
}                                                                                      \
BLIS_EXPORT void cblas_##ch##syrk(CBLAS_ORDER Order, CBLAS_UPLO Uplo,                  \
        CBLAS_TRANSPOSE Trans, f77_int N, f77_int K, T alpha, const T* A, f77_int lda, \
        T beta, T* C, f77_int ldc) {                                                   \
    const char* rout = "cblas_" #ch "syrk";                                            \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    const bool row = Order == CblasRowMajor;                                           \
    RowMajorScope rm(row);                                                             \
    char ul, tr;                                                                       \
    if (!uplo_char(Uplo, row, &ul)) {                                                  \
        cblas_xerbla(row ? 3 : 2, rout, "Illegal Uplo setting, %d\n", Uplo);           \
        return;                                                                        \
    }                                                                                  \
    if (!(row ? trans_char_toggled(Trans, &tr) : trans_char(Trans, &tr))) {            \
        cblas_xerbla(3, rout, "Illegal Trans setting, %d\n", Trans);                   \
        return;                                                                        \
    }                                                                                  \
    ch##syrk_(&ul, &tr, &N, &K, &alpha, A, &lda, &beta, C, &ldc);                      \
}                                                                                      \
BLIS_EXPORT void cblas_##ch##syr2k(CBLAS_ORDER Order, CBLAS_UPLO Uplo,                 \
        CBLAS_TRANSPOSE Trans, f77_int N, f77_int K, T alpha, const T* A, f77_int lda, \
        const T* B, f77_int ldb, T beta, T* C, f77_int ldc) {                          \
    const char* rout = "cblas_" #ch "syr2k";                                           \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    const bool row = Order == CblasRowMajor;                                           \
    RowMajorScope rm(row);                                                             \
    char ul, tr;                                                                       \
    if (!uplo_char(Uplo, row, &ul)) {                                                  \
        cblas_xerbla(row ? 3 : 2, rout, "Illegal Uplo setting, %d\n", Uplo);           \
        return;                                                                        \
    }                                                                                  \
    if (!(row ? trans_char_toggled(Trans, &tr) : trans_char(Trans, &tr))) {            \
        cblas_xerbla(3, rout, "Illegal Trans setting, %d\n", Trans);                   \
        return;                                                                        \
    }                                                                                  \
    ch##syr2k_(&ul, &tr, &N, &K, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);            \
}                                                                                      \
BLIS_EXPORT void cblas_##ch##trmm(CBLAS_ORDER Order, CBLAS_SIDE Side, CBLAS_UPLO Uplo, \
        CBLAS_TRANSPOSE TransA, CBLAS_DIAG Diag, f77_int M, f77_int N, T alpha,        \
        const T* A, f77_int lda, T* B, f77_int ldb) {                                  \
    const char* rout = "cblas_" #ch "trmm";                                            \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    const bool row = Order == CblasRowMajor;                                           \
    RowMajorScope rm(row);                                                             \
    char sd, ul, ta, di;                                                               \
    if (!side_char(Side, row, &sd)) { cblas_xerbla(2, rout, "Illegal Side setting, %d\n", Side); return; } \
    if (!uplo_char(Uplo, row, &ul)) { cblas_xerbla(3, rout, "Illegal Uplo setting, %d\n", Uplo); return; } \
    if (!trans_char(TransA, &ta)) { cblas_xerbla(4, rout, "Illegal Trans setting, %d\n", TransA); return; } \
    if (!diag_char(Diag, &di)) { cblas_xerbla(5, rout, "Illegal Diag setting, %d\n", Diag); return; } \
    if (!row) ch##trmm_(&sd, &ul, &ta, &di, &M, &N, &alpha, A, &lda, B, &ldb);         \
    else ch##trmm_(&sd, &ul, &ta, &di, &N, &M, &alpha, A, &lda, B, &ldb);              \
}                                                                                      \
BLIS_EXPORT void cblas_##ch##trsm(CBLAS_ORDER Order, CBLAS_SIDE Side, CBLAS_UPLO Uplo, \
        CBLAS_TRANSPOSE TransA, CBLAS_DIAG Diag, f77_int M, f77_int N, T alpha,        \
        const T* A, f77_int lda, T* B, f77_int ldb) {                                  \
    const char* rout = "cblas_" #ch "trsm";                                            \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    const bool row = Order == CblasRowMajor;                                           \
    RowMajorScope rm(row);                                                             \
    char sd, ul, ta, di;                                                               \
    if (!side_char(Side, row, &sd)) { cblas_xerbla(2, rout, "Illegal Side setting, %d\n", Side); return; } \
    if (!uplo_char(Uplo, row, &ul)) { cblas_xerbla(3, rout, "Illegal Uplo setting, %d\n", Uplo); return; } \
    if (!trans_char(TransA, &ta)) { cblas_xerbla(4, rout, "Illegal Trans setting, %d\n", TransA); return; } \
    if (!diag_char(Diag, &di)) { cblas_xerbla(5, rout, "Illegal Diag setting, %d\n", Diag); return; } \
    if (!row) ch##trsm_(&sd, &ul, &ta, &di, &M, &N, &alpha, A, &lda, B, &ldb);         \
    else ch##trsm_(&sd, &ul, &ta, &di, &N, &M, &alpha, A, &lda, B, &ldb);              \
}
CBLAS_L3(s, float)
CBLAS_L3(d, double)

// gemm_batch (extra/cblas_?gemm_batch.c): per group, the same mapping as
// cblas_?gemm, then one ?gemm_batch_ call over all groups.
namespace {

template <typename T, typename Fn>
void cblas_batch(const char* rout, CBLAS_ORDER Order, CBLAS_TRANSPOSE* TransA,
                 CBLAS_TRANSPOSE* TransB, f77_int* M, f77_int* N, f77_int* K,
                 const T* alpha, const T** A, f77_int* lda, const T** B, f77_int* ldb,
                 const T* beta, T** C, f77_int* ldc, f77_int group_count, f77_int* group_size,
                 Fn f77_batch) {
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; }
    const bool row = Order == CblasRowMajor;
    RowMajorScope rm(row);
    const f77_int G = group_count > 0 ? group_count : 0;
    std::vector<char> ta(G), tb(G);
    for (f77_int g = 0; g < G; ++g) {
        if (!trans_char(TransA[g], &ta[g])) {
            cblas_xerbla(2, rout, "Illegal TransA setting %d for group %d\n", TransA[g], g);
            return;
        }
        if (!trans_char(TransB[g], &tb[g])) {
            cblas_xerbla(row ? 2 : 3, rout, "Illegal TransB setting %d for group %d\n",
                         TransB[g], g);
            return;
        }
    }
    if (!row) {
        f77_batch(ta.data(), tb.data(), M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                  &group_count, group_size);
    } else {
        f77_batch(tb.data(), ta.data(), N, M, K, alpha, B, ldb, A, lda, beta, C, ldc,
                  &group_count, group_size);
    }
}

} // namespace

#define CBLAS_BATCH(ch, T)                                                             \
BLIS_EXPORT void cblas_##ch##gemm_batch(CBLAS_ORDER Order, CBLAS_TRANSPOSE* TransA_array, \
        CBLAS_TRANSPOSE* TransB_array, f77_int* M_array, f77_int* N_array,             \
        f77_int* K_array, const T* alpha_array, const T** A_array, f77_int* lda_array, \
        const T** B_array, f77_int* ldb_array, const T* beta_array, T** C_array,       \
        f77_int* ldc_array, f77_int group_count, f77_int* group_size) {                \
    cblas_batch<T>("cblas_" #ch "gemm_batch", Order, TransA_array, TransB_array,       \
                   M_array, N_array, K_array, alpha_array, A_array, lda_array, B_array, \
                   ldb_array, beta_array, C_array, ldc_array, group_count, group_size, \
                   ch##gemm_batch_);                                                   \
}
CBLAS_BATCH(s, float)
CBLAS_BATCH(d, double)

// The complex forms take their scalars and pointer arrays as void.
#define CBLAS_ZBATCH(ch, T)                                                            \
BLIS_EXPORT void cblas_##ch##gemm_batch(CBLAS_ORDER Order, CBLAS_TRANSPOSE* TransA_array, \
        CBLAS_TRANSPOSE* TransB_array, f77_int* M_array, f77_int* N_array,             \
        f77_int* K_array, const void* alpha_array, const void** A_array,               \
        f77_int* lda_array, const void** B_array, f77_int* ldb_array,                  \
        const void* beta_array, void** C_array, f77_int* ldc_array, f77_int group_count, \
        f77_int* group_size) {                                                         \
    cblas_batch<T>("cblas_" #ch "gemm_batch", Order, TransA_array, TransB_array,       \
                   M_array, N_array, K_array, (const T*)alpha_array,                   \
                   (const T**)A_array, lda_array, (const T**)B_array, ldb_array,       \
                   (const T*)beta_array, (T**)C_array, ldc_array, group_count,         \
                   group_size, ch##gemm_batch_);                                       \
}
CBLAS_ZBATCH(c, scomplex)
CBLAS_ZBATCH(z, dcomplex)

#define CBLAS_ZGEMM(ch, T)                                                             \
BLIS_EXPORT void cblas_##ch##gemm(CBLAS_ORDER Order, CBLAS_TRANSPOSE TransA,           \
        CBLAS_TRANSPOSE TransB, f77_int M, f77_int N, f77_int K, const void* alpha,    \
        const void* A, f77_int lda, const void* B, f77_int ldb, const void* beta,      \
        void* C, f77_int ldc) {                                                        \
    const char* rout = "cblas_" #ch "gemm";                                            \
    if (!order_ok(Order)) { cblas_xerbla(1, rout, "Illegal Order setting, %d\n", Order); return; } \
    RowMajorScope rm(Order == CblasRowMajor);                                          \
    char ta, tb;                                                                       \
    if (!trans_char(TransA, &ta)) { cblas_xerbla(2, rout, "Illegal TransA setting, %d\n", TransA); return; } \
    if (!trans_char(TransB, &tb)) {                                                    \
        cblas_xerbla(Order == CblasColMajor ? 3 : 2, rout, "Illegal TransB setting, %d\n", TransB); \
        return;                                                                        \
    }                                                                                  \
    const T* al = (const T*)alpha; const T* be = (const T*)beta;                       \
    if (Order == CblasColMajor)                                                        \
        ch##gemm_(&ta, &tb, &M, &N, &K, al, (const T*)A, &lda, (const T*)B, &ldb, be,  \
                  (T*)C, &ldc);                                                        \
    else                                                                               \
        ch##gemm_(&tb, &ta, &N, &M, &K, al, (const T*)B, &ldb, (const T*)A, &lda, be,  \
                  (T*)C, &ldc);                                                        \
}
CBLAS_ZGEMM(c, scomplex)
CBLAS_ZGEMM(z, dcomplex)
