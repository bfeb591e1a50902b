// api_struc.cpp — typed, object and BLAS interfaces of the structured
// level-3 operations: gemmt, syrk, syr2k, symm, trmm, trmm3.
//
// Reference interfaces (paths relative to the BLIS source tree):
//   typed                 frame/3/bli_l3_tapi_ex.h:60-225 (prototypes),
//                         frame/3/bli_l3_tapi_ex.c (object wrapping)
//   object                frame/3/bli_l3_oapi_ex.c:151-690
//                         (syrk = gemmt(A, A^T); syr2k = two gemmt calls, beta
//                         only in the first; symm swaps A and B on the right)
//   BLAS                  frame/compat/bla_symm.c, bla_trmm.c, bla_syrk.c,
//                         bla_syr2k.c, extra/bla_gemmt.c
//   BLAS argument checks  frame/compat/check/bla_hemm_check.h (symm),
//                         bla_trmm_check.h, bla_syrk_check.h,
//                         bla_syr2k_check.h, bla_gemmt_check.h
// Every entry point checks its arguments as the reference does, stages host
// operands, folds transposition into strides and hands the problem to the
// structured planners of struc.hip.
#include "api_common.hpp"

#include <algorithm>

using namespace blis_amd;
using namespace blis_amd::api;

#define BLIS_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

inline bool valid_uplo(uplo_t u) { return u == BLIS_LOWER || u == BLIS_UPPER; }
inline bool valid_side(side_t s) { return s == BLIS_LEFT || s == BLIS_RIGHT; }
inline bool valid_diag(diag_t d) { return d == BLIS_NONUNIT_DIAG || d == BLIS_UNIT_DIAG; }
inline trans_t toggle(trans_t t) { return (trans_t)((unsigned)t ^ BLIS_TRANS_BIT); }

// View of op(X) for X with strides (rs, cs): transposition swaps them.
inline View view_of(void* p, trans_t t, inc_t rs, inc_t cs) {
    return does_trans(t) ? View{ p, cs, rs } : View{ p, rs, cs };
}

// ---- typed cores ----------------------------------------------------------

template <typename T>
void gemmt_core(uplo_t uploc, trans_t transa, trans_t transb, dim_t m, dim_t k,
                const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
                const T* b, inc_t rs_b, inc_t cs_b,
                const T* beta, T* c, inc_t rs_c, inc_t cs_c, doff_t doff = 0) {
    init_once();
    require(valid_uplo(uploc), BLIS_INVALID_UPLO, "gemmt: uploc");
    require(valid_trans(transa) && valid_trans(transb), BLIS_INVALID_TRANS, "gemmt: trans");
    require(m >= 0 && k >= 0, BLIS_NEGATIVE_DIMENSION, "gemmt: dimensions");
    if (m <= 0) return;
    const double al = (double)*alpha, be = (double)*beta;
    const bool ta = does_trans(transa), tb = does_trans(transb);
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb, sc;
    if (k > 0 && al != 0.0) {
        stage(sa, a, dt, ta ? k : m, ta ? m : k, rs_a, cs_a, false);
        stage(sb, b, dt, tb ? m : k, tb ? k : m, rs_b, cs_b, false);
    }
    stage(sc, c, dt, m, m, rs_c, cs_c, true);
    GemmDesc d;
    d.dt = dt; d.m = m; d.n = m; d.k = k;
    d.a = view_of(sa.dev, transa, rs_a, cs_a);
    d.b = view_of(sb.dev, transb, rs_b, cs_b);
    d.c = { sc.dev, rs_c, cs_c };
    d.alpha = al; d.beta = be;
    gemmt_run(d, uploc, doff, stream());
    sync_if_needed();
}

template <typename T>
void syr2k_core(uplo_t uploc, trans_t transa, trans_t transb, dim_t m, dim_t k,
                const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
                const T* b, inc_t rs_b, inc_t cs_b,
                const T* beta, T* c, inc_t rs_c, inc_t cs_c) {
    // C := alpha*op(A)*op(B)^T + beta*C, then C += alpha*op(B)*op(A)^T on
    // the same triangle (bli_syr2k_ex, frame/3/bli_l3_oapi_ex.c:260-285).
    const T one = T(1);
    gemmt_core<T>(uploc, transa, toggle(transb), m, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                  beta, c, rs_c, cs_c);
    gemmt_core<T>(uploc, transb, toggle(transa), m, k, alpha, b, rs_b, cs_b, a, rs_a, cs_a,
                  &one, c, rs_c, cs_c);
}

template <typename T>
void symm_core(side_t side, uplo_t uploa, trans_t transb, dim_t m, dim_t n,
               const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
               const T* b, inc_t rs_b, inc_t cs_b,
               const T* beta, T* c, inc_t rs_c, inc_t cs_c) {
    init_once();
    require(valid_side(side), BLIS_INVALID_SIDE, "symm: side");
    require(valid_uplo(uploa), BLIS_INVALID_UPLO, "symm: uploa");
    require(valid_trans(transb), BLIS_INVALID_TRANS, "symm: transb");
    require(m >= 0 && n >= 0, BLIS_NEGATIVE_DIMENSION, "symm: dimensions");
    if (m <= 0 || n <= 0) return;
    const double al = (double)*alpha, be = (double)*beta;
    const bool tb = does_trans(transb);
    const dim_t na = side == BLIS_LEFT ? m : n;
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb, sc;
    if (al != 0.0) {
        stage(sa, a, dt, na, na, rs_a, cs_a, false);
        stage(sb, b, dt, tb ? n : m, tb ? m : n, rs_b, cs_b, false);
    }
    stage(sc, c, dt, m, n, rs_c, cs_c, true);
    symm_run(dt, side, uploa, m, n, al, View{ sa.dev, rs_a, cs_a },
             view_of(sb.dev, transb, rs_b, cs_b), be, View{ sc.dev, rs_c, cs_c }, stream());
    sync_if_needed();
}

template <typename T>
void trmm3_core(side_t side, uplo_t uploa, trans_t transa, diag_t diaga, trans_t transb,
                dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
                const T* b, inc_t rs_b, inc_t cs_b,
                const T* beta, T* c, inc_t rs_c, inc_t cs_c, doff_t doff = 0) {
    init_once();
    require(valid_side(side), BLIS_INVALID_SIDE, "trmm3: side");
    require(valid_uplo(uploa), BLIS_INVALID_UPLO, "trmm3: uploa");
    require(valid_trans(transa) && valid_trans(transb), BLIS_INVALID_TRANS, "trmm3: trans");
    require(valid_diag(diaga), BLIS_INVALID_DIAG, "trmm3: diaga");
    require(m >= 0 && n >= 0, BLIS_NEGATIVE_DIMENSION, "trmm3: dimensions");
    if (m <= 0 || n <= 0) return;
    const double al = (double)*alpha, be = (double)*beta;
    const bool tb = does_trans(transb);
    const dim_t na = side == BLIS_LEFT ? m : n;
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb, sc;
    if (al != 0.0) {
        stage(sa, a, dt, na, na, rs_a, cs_a, false);
        stage(sb, b, dt, tb ? n : m, tb ? m : n, rs_b, cs_b, false);
    }
    stage(sc, c, dt, m, n, rs_c, cs_c, true);
    trmm3_run(dt, side, uploa, transa, diaga, doff, m, n, al, View{ sa.dev, rs_a, cs_a },
              view_of(sb.dev, transb, rs_b, cs_b), be, View{ sc.dev, rs_c, cs_c }, stream());
    sync_if_needed();
}

template <typename T>
void trmm_core(side_t side, uplo_t uploa, trans_t transa, diag_t diaga, dim_t m, dim_t n,
               const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
               T* b, inc_t rs_b, inc_t cs_b, doff_t doff = 0) {
    init_once();
    require(valid_side(side), BLIS_INVALID_SIDE, "trmm: side");
    require(valid_uplo(uploa), BLIS_INVALID_UPLO, "trmm: uploa");
    require(valid_trans(transa), BLIS_INVALID_TRANS, "trmm: transa");
    require(valid_diag(diaga), BLIS_INVALID_DIAG, "trmm: diaga");
    require(m >= 0 && n >= 0, BLIS_NEGATIVE_DIMENSION, "trmm: dimensions");
    if (m <= 0 || n <= 0) return;
    const double al = (double)*alpha;
    const dim_t na = side == BLIS_LEFT ? m : n;
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb;
    if (al != 0.0) stage(sa, a, dt, na, na, rs_a, cs_a, false);
    stage(sb, b, dt, m, n, rs_b, cs_b, true);
    trmm_run(dt, side, uploa, transa, diaga, doff, m, n, al, View{ sa.dev, rs_a, cs_a },
             View{ sb.dev, rs_b, cs_b }, stream());
    sync_if_needed();
}

// ---- object helpers ---------------------------------------------------------

template <typename T>
const T* cptr(const obj_t* o) { return (const T*)obj_origin(o); }
template <typename T>
T* mptr(const obj_t* o) { return (T*)obj_origin(o); }

// A scalar object's value in the operation's datatype.
template <typename T>
T scal(const obj_t* o) { return (T)obj_scalar(o); }

bool obj_is_double(const obj_t* o) { return bli_obj_dt(o) == BLIS_DOUBLE; }

void check_same_dt(const obj_t* a, const obj_t* b, const char* where) {
    require(bli_obj_dt(a) == bli_obj_dt(b), BLIS_INCONSISTENT_DATATYPES, where);
}

} // namespace

// ===========================================================================
// Typed API
// ===========================================================================

#define STRUC_TYPED(ch, T)                                                             \
BLIS_EXPORT void bli_##ch##gemmt_ex(uplo_t uploc, trans_t transa, trans_t transb,      \
        dim_t m, dim_t k, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,          \
        const T* b, inc_t rs_b, inc_t cs_b, const T* beta, T* c, inc_t rs_c,           \
        inc_t cs_c, const cntx_t* cntx, const rntm_t* rntm) {                          \
    (void)cntx; (void)rntm;                                                            \
    gemmt_core<T>(uploc, transa, transb, m, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,    \
                  beta, c, rs_c, cs_c);                                                \
}                                                                                      \
BLIS_EXPORT void bli_##ch##gemmt(uplo_t uploc, trans_t transa, trans_t transb,         \
        dim_t m, dim_t k, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,          \
        const T* b, inc_t rs_b, inc_t cs_b, const T* beta, T* c, inc_t rs_c,           \
        inc_t cs_c) {                                                                  \
    gemmt_core<T>(uploc, transa, transb, m, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,    \
                  beta, c, rs_c, cs_c);                                                \
}                                                                                      \
BLIS_EXPORT void bli_##ch##syrk_ex(uplo_t uploc, trans_t transa, dim_t m, dim_t k,     \
        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, const T* beta, T* c,       \
        inc_t rs_c, inc_t cs_c, const cntx_t* cntx, const rntm_t* rntm) {              \
    (void)cntx; (void)rntm;                                                            \
    gemmt_core<T>(uploc, transa, toggle(transa), m, k, alpha, a, rs_a, cs_a, a, rs_a,  \
                  cs_a, beta, c, rs_c, cs_c);                                          \
}                                                                                      \
BLIS_EXPORT void bli_##ch##syrk(uplo_t uploc, trans_t transa, dim_t m, dim_t k,        \
        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, const T* beta, T* c,       \
        inc_t rs_c, inc_t cs_c) {                                                      \
    gemmt_core<T>(uploc, transa, toggle(transa), m, k, alpha, a, rs_a, cs_a, a, rs_a,  \
                  cs_a, beta, c, rs_c, cs_c);                                          \
}                                                                                      \
BLIS_EXPORT void bli_##ch##syr2k_ex(uplo_t uploc, trans_t transa, trans_t transb,      \
        dim_t m, dim_t k, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,          \
        const T* b, inc_t rs_b, inc_t cs_b, const T* beta, T* c, inc_t rs_c,           \
        inc_t cs_c, const cntx_t* cntx, const rntm_t* rntm) {                          \
    (void)cntx; (void)rntm;                                                            \
    syr2k_core<T>(uploc, transa, transb, m, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,    \
                  beta, c, rs_c, cs_c);                                                \
}                                                                                      \
BLIS_EXPORT void bli_##ch##syr2k(uplo_t uploc, trans_t transa, trans_t transb,         \
        dim_t m, dim_t k, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,          \
        const T* b, inc_t rs_b, inc_t cs_b, const T* beta, T* c, inc_t rs_c,           \
        inc_t cs_c) {                                                                  \
    syr2k_core<T>(uploc, transa, transb, m, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,    \
                  beta, c, rs_c, cs_c);                                                \
}                                                                                      \
BLIS_EXPORT void bli_##ch##symm_ex(side_t side, uplo_t uploa, conj_t conja,            \
        trans_t transb, dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a,      \
        inc_t cs_a, const T* b, inc_t rs_b, inc_t cs_b, const T* beta, T* c,           \
        inc_t rs_c, inc_t cs_c, const cntx_t* cntx, const rntm_t* rntm) {              \
    (void)conja; (void)cntx; (void)rntm;                                               \
    symm_core<T>(side, uploa, transb, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, \
                 c, rs_c, cs_c);                                                       \
}                                                                                      \
BLIS_EXPORT void bli_##ch##symm(side_t side, uplo_t uploa, conj_t conja,               \
        trans_t transb, dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a,      \
        inc_t cs_a, const T* b, inc_t rs_b, inc_t cs_b, const T* beta, T* c,           \
        inc_t rs_c, inc_t cs_c) {                                                      \
    (void)conja;                                                                       \
    symm_core<T>(side, uploa, transb, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, \
                 c, rs_c, cs_c);                                                       \
}                                                                                      \
BLIS_EXPORT void bli_##ch##trmm3_ex(side_t side, uplo_t uploa, trans_t transa,         \
        diag_t diaga, trans_t transb, dim_t m, dim_t n, const T* alpha, const T* a,    \
        inc_t rs_a, inc_t cs_a, const T* b, inc_t rs_b, inc_t cs_b, const T* beta,     \
        T* c, inc_t rs_c, inc_t cs_c, const cntx_t* cntx, const rntm_t* rntm) {        \
    (void)cntx; (void)rntm;                                                            \
    trmm3_core<T>(side, uploa, transa, diaga, transb, m, n, alpha, a, rs_a, cs_a, b,   \
                  rs_b, cs_b, beta, c, rs_c, cs_c);                                    \
}                                                                                      \
BLIS_EXPORT void bli_##ch##trmm3(side_t side, uplo_t uploa, trans_t transa,            \
        diag_t diaga, trans_t transb, dim_t m, dim_t n, const T* alpha, const T* a,    \
        inc_t rs_a, inc_t cs_a, const T* b, inc_t rs_b, inc_t cs_b, const T* beta,     \
        T* c, inc_t rs_c, inc_t cs_c) {                                                \
    trmm3_core<T>(side, uploa, transa, diaga, transb, m, n, alpha, a, rs_a, cs_a, b,   \
                  rs_b, cs_b, beta, c, rs_c, cs_c);                                    \
}                                                                                      \
BLIS_EXPORT void bli_##ch##trmm_ex(side_t side, uplo_t uploa, trans_t transa,          \
        diag_t diaga, dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a,        \
        inc_t cs_a, T* b, inc_t rs_b, inc_t cs_b, const cntx_t* cntx,                  \
        const rntm_t* rntm) {                                                          \
    (void)cntx; (void)rntm;                                                            \
    trmm_core<T>(side, uploa, transa, diaga, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b); \
}                                                                                      \
BLIS_EXPORT void bli_##ch##trmm(side_t side, uplo_t uploa, trans_t transa,             \
        diag_t diaga, dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a,        \
        inc_t cs_a, T* b, inc_t rs_b, inc_t cs_b) {                                    \
    trmm_core<T>(side, uploa, transa, diaga, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b); \
}
STRUC_TYPED(s, float)
STRUC_TYPED(d, double)

// ===========================================================================
// Object API
// ===========================================================================

BLIS_EXPORT void bli_gemmt_ex(const obj_t* alpha, const obj_t* a, const obj_t* b,
                              const obj_t* beta, const obj_t* c,
                              const cntx_t* cntx, const rntm_t* rntm) {
    (void)cntx; (void)rntm;
    init_once();
    check_matrix_obj(a, "bli_gemmt: a");
    check_matrix_obj(b, "bli_gemmt: b");
    check_matrix_obj(c, "bli_gemmt: c");
    check_same_dt(a, c, "bli_gemmt");
    check_same_dt(b, c, "bli_gemmt");
    require(c->dim[0] == c->dim[1], BLIS_EXPECTED_SQUARE_OBJECT, "bli_gemmt: c");
    const dim_t m = c->dim[0], k = bli_obj_width_after_trans(a);
    require(bli_obj_length_after_trans(a) == m && bli_obj_width_after_trans(b) == m &&
            bli_obj_length_after_trans(b) == k, BLIS_NONCONFORMAL_DIMENSIONS, "bli_gemmt");
    require(valid_uplo(bli_obj_uplo(c)), BLIS_EXPECTED_UPPER_OR_LOWER_OBJECT, "bli_gemmt: c");
    const trans_t ta = bli_obj_onlytrans_status(a), tb = bli_obj_onlytrans_status(b);
    if (obj_is_double(c)) {
        const double al = scal<double>(alpha), be = scal<double>(beta);
        gemmt_core<double>(bli_obj_uplo(c), ta, tb, m, k, &al, cptr<double>(a), a->rs, a->cs,
                           cptr<double>(b), b->rs, b->cs, &be, mptr<double>(c), c->rs, c->cs,
                           bli_obj_diag_offset(c));
    } else {
        const float al = scal<float>(alpha), be = scal<float>(beta);
        gemmt_core<float>(bli_obj_uplo(c), ta, tb, m, k, &al, cptr<float>(a), a->rs, a->cs,
                          cptr<float>(b), b->rs, b->cs, &be, mptr<float>(c), c->rs, c->cs,
                          bli_obj_diag_offset(c));
    }
}

BLIS_EXPORT void bli_gemmt(const obj_t* alpha, const obj_t* a, const obj_t* b,
                           const obj_t* beta, const obj_t* c) {
    bli_gemmt_ex(alpha, a, b, beta, c, nullptr, nullptr);
}

// syrk: gemmt(A, A^T) (bli_l3_oapi_ex.c:312-330).
BLIS_EXPORT void bli_syrk_ex(const obj_t* alpha, const obj_t* a, const obj_t* beta,
                             const obj_t* c, const cntx_t* cntx, const rntm_t* rntm) {
    require(a != nullptr, BLIS_NULL_POINTER, "bli_syrk: a");
    obj_t at = *a;
    bli_obj_toggle_trans(&at);
    bli_gemmt_ex(alpha, a, &at, beta, c, cntx, rntm);
}

BLIS_EXPORT void bli_syrk(const obj_t* alpha, const obj_t* a, const obj_t* beta,
                          const obj_t* c) {
    bli_syrk_ex(alpha, a, beta, c, nullptr, nullptr);
}

// syr2k: gemmt(A, B^T, beta) then gemmt(B, A^T, 1) (bli_l3_oapi_ex.c:260-285).
BLIS_EXPORT void bli_syr2k_ex(const obj_t* alpha, const obj_t* a, const obj_t* b,
                              const obj_t* beta, const obj_t* c,
                              const cntx_t* cntx, const rntm_t* rntm) {
    require(a != nullptr && b != nullptr, BLIS_NULL_POINTER, "bli_syr2k");
    obj_t at = *a, bt = *b;
    bli_obj_toggle_trans(&at);
    bli_obj_toggle_trans(&bt);
    bli_gemmt_ex(alpha, a, &bt, beta, c, cntx, rntm);
    bli_gemmt_ex(alpha, b, &at, &BLIS_ONE, c, cntx, rntm);
}

BLIS_EXPORT void bli_syr2k(const obj_t* alpha, const obj_t* a, const obj_t* b,
                           const obj_t* beta, const obj_t* c) {
    bli_syr2k_ex(alpha, a, b, beta, c, nullptr, nullptr);
}

BLIS_EXPORT void bli_symm_ex(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                             const obj_t* beta, const obj_t* c,
                             const cntx_t* cntx, const rntm_t* rntm) {
    (void)cntx; (void)rntm;
    init_once();
    check_matrix_obj(a, "bli_symm: a");
    check_matrix_obj(b, "bli_symm: b");
    check_matrix_obj(c, "bli_symm: c");
    check_same_dt(a, c, "bli_symm");
    check_same_dt(b, c, "bli_symm");
    require(valid_side(side), BLIS_INVALID_SIDE, "bli_symm: side");
    require(bli_obj_struc(a) == BLIS_SYMMETRIC, BLIS_EXPECTED_SYMMETRIC_OBJECT, "bli_symm: a");
    require(a->dim[0] == a->dim[1], BLIS_EXPECTED_SQUARE_OBJECT, "bli_symm: a");
    require(valid_uplo(bli_obj_uplo(a)), BLIS_EXPECTED_UPPER_OR_LOWER_OBJECT, "bli_symm: a");
    require(bli_obj_diag_offset(a) == 0, BLIS_UNEXPECTED_DIAG_OFFSET, "bli_symm: a");
    const dim_t m = c->dim[0], n = c->dim[1];
    require(a->dim[0] == (side == BLIS_LEFT ? m : n) && bli_obj_length_after_trans(b) == m &&
            bli_obj_width_after_trans(b) == n, BLIS_NONCONFORMAL_DIMENSIONS, "bli_symm");
    const trans_t tb = bli_obj_onlytrans_status(b);
    if (obj_is_double(c)) {
        const double al = scal<double>(alpha), be = scal<double>(beta);
        symm_core<double>(side, bli_obj_uplo(a), tb, m, n, &al, cptr<double>(a), a->rs, a->cs,
                          cptr<double>(b), b->rs, b->cs, &be, mptr<double>(c), c->rs, c->cs);
    } else {
        const float al = scal<float>(alpha), be = scal<float>(beta);
        symm_core<float>(side, bli_obj_uplo(a), tb, m, n, &al, cptr<float>(a), a->rs, a->cs,
                         cptr<float>(b), b->rs, b->cs, &be, mptr<float>(c), c->rs, c->cs);
    }
}

BLIS_EXPORT void bli_symm(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                          const obj_t* beta, const obj_t* c) {
    bli_symm_ex(side, alpha, a, b, beta, c, nullptr, nullptr);
}

namespace {

void check_triangular(const obj_t* a, const char* where) {
    require(bli_obj_struc(a) == BLIS_TRIANGULAR, BLIS_EXPECTED_TRIANGULAR_OBJECT, where);
    require(a->dim[0] == a->dim[1], BLIS_EXPECTED_SQUARE_OBJECT, where);
    require(valid_uplo(bli_obj_uplo(a)), BLIS_EXPECTED_UPPER_OR_LOWER_OBJECT, where);
}

} // namespace

BLIS_EXPORT void bli_trmm3_ex(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                              const obj_t* beta, const obj_t* c,
                              const cntx_t* cntx, const rntm_t* rntm) {
    (void)cntx; (void)rntm;
    init_once();
    check_matrix_obj(a, "bli_trmm3: a");
    check_matrix_obj(b, "bli_trmm3: b");
    check_matrix_obj(c, "bli_trmm3: c");
    check_same_dt(a, c, "bli_trmm3");
    check_same_dt(b, c, "bli_trmm3");
    require(valid_side(side), BLIS_INVALID_SIDE, "bli_trmm3: side");
    check_triangular(a, "bli_trmm3: a");
    const dim_t m = c->dim[0], n = c->dim[1];
    require(a->dim[0] == (side == BLIS_LEFT ? m : n) && bli_obj_length_after_trans(b) == m &&
            bli_obj_width_after_trans(b) == n, BLIS_NONCONFORMAL_DIMENSIONS, "bli_trmm3");
    const trans_t ta = bli_obj_onlytrans_status(a), tb = bli_obj_onlytrans_status(b);
    const diag_t dg = bli_obj_diag(a);
    const doff_t doff = bli_obj_diag_offset(a);
    if (obj_is_double(c)) {
        const double al = scal<double>(alpha), be = scal<double>(beta);
        trmm3_core<double>(side, bli_obj_uplo(a), ta, dg, tb, m, n, &al, cptr<double>(a), a->rs,
                           a->cs, cptr<double>(b), b->rs, b->cs, &be, mptr<double>(c), c->rs,
                           c->cs, doff);
    } else {
        const float al = scal<float>(alpha), be = scal<float>(beta);
        trmm3_core<float>(side, bli_obj_uplo(a), ta, dg, tb, m, n, &al, cptr<float>(a), a->rs,
                          a->cs, cptr<float>(b), b->rs, b->cs, &be, mptr<float>(c), c->rs,
                          c->cs, doff);
    }
}

BLIS_EXPORT void bli_trmm3(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                           const obj_t* beta, const obj_t* c) {
    bli_trmm3_ex(side, alpha, a, b, beta, c, nullptr, nullptr);
}

BLIS_EXPORT void bli_trmm_ex(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                             const cntx_t* cntx, const rntm_t* rntm) {
    (void)cntx; (void)rntm;
    init_once();
    check_matrix_obj(a, "bli_trmm: a");
    check_matrix_obj(b, "bli_trmm: b");
    check_same_dt(a, b, "bli_trmm");
    require(valid_side(side), BLIS_INVALID_SIDE, "bli_trmm: side");
    check_triangular(a, "bli_trmm: a");
    const dim_t m = b->dim[0], n = b->dim[1];
    require(a->dim[0] == (side == BLIS_LEFT ? m : n), BLIS_NONCONFORMAL_DIMENSIONS, "bli_trmm");
    const trans_t ta = bli_obj_onlytrans_status(a);
    const diag_t dg = bli_obj_diag(a);
    const doff_t doff = bli_obj_diag_offset(a);
    if (obj_is_double(b)) {
        const double al = scal<double>(alpha);
        trmm_core<double>(side, bli_obj_uplo(a), ta, dg, m, n, &al, cptr<double>(a), a->rs,
                          a->cs, mptr<double>(b), b->rs, b->cs, doff);
    } else {
        const float al = scal<float>(alpha);
        trmm_core<float>(side, bli_obj_uplo(a), ta, dg, m, n, &al, cptr<float>(a), a->rs,
                         a->cs, mptr<float>(b), b->rs, b->cs, doff);
    }
}

BLIS_EXPORT void bli_trmm(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b) {
    bli_trmm_ex(side, alpha, a, b, nullptr, nullptr);
}

// ===========================================================================
// BLAS compatibility
// ===========================================================================

namespace {

inline uplo_t uplo_of(char ch) { return up(ch) == 'L' ? BLIS_LOWER : BLIS_UPPER; }

template <typename T>
void blas_symm(const char* dtc, const f77_char* side, const f77_char* uploa,
               const f77_int* m, const f77_int* n, const T* alpha, const T* a,
               const f77_int* lda, const T* b, const f77_int* ldb, const T* beta, T* c,
               const f77_int* ldc) {
    // frame/compat/check/bla_hemm_check.h (bla_symm_check is bla_hemm_check)
    const bool left = up(*side) == 'L', right = up(*side) == 'R';
    const bool lower = up(*uploa) == 'L', upper = up(*uploa) == 'U';
    const f77_int nrowa = left ? *m : *n;
    f77_int info = 0;
    if (!left && !right) info = 1;
    else if (!lower && !upper) info = 2;
    else if (*m < 0) info = 3;
    else if (*n < 0) info = 4;
    else if (*lda < std::max<f77_int>(1, nrowa)) info = 7;
    else if (*ldb < std::max<f77_int>(1, *m)) info = 9;
    else if (*ldc < std::max<f77_int>(1, *m)) info = 12;
    if (info != 0) { blas_error(dtc, "symm", info); return; }
    symm_core<T>(left ? BLIS_LEFT : BLIS_RIGHT, uplo_of(*uploa), BLIS_NO_TRANSPOSE, *m, *n,
                 alpha, a, 1, *lda, b, 1, *ldb, beta, c, 1, *ldc);
}

template <typename T>
void blas_trmm(const char* dtc, const f77_char* side, const f77_char* uploa,
               const f77_char* transa, const f77_char* diaga, const f77_int* m,
               const f77_int* n, const T* alpha, const T* a, const f77_int* lda, T* b,
               const f77_int* ldb) {
    // frame/compat/check/bla_trmm_check.h
    const bool left = up(*side) == 'L', right = up(*side) == 'R';
    const bool lower = up(*uploa) == 'L', upper = up(*uploa) == 'U';
    const bool nota = up(*transa) == 'N', ta = up(*transa) == 'T', conja = up(*transa) == 'C';
    const bool unita = up(*diaga) == 'U', nonua = up(*diaga) == 'N';
    const f77_int nrowa = left ? *m : *n;
    f77_int info = 0;
    if (!left && !right) info = 1;
    else if (!lower && !upper) info = 2;
    else if (!nota && !ta && !conja) info = 3;
    else if (!unita && !nonua) info = 4;
    else if (*m < 0) info = 5;
    else if (*n < 0) info = 6;
    else if (*lda < std::max<f77_int>(1, nrowa)) info = 9;
    else if (*ldb < std::max<f77_int>(1, *m)) info = 11;
    if (info != 0) { blas_error(dtc, "trmm", info); return; }
    trans_t bt;
    map_trans(*transa, &bt);
    trmm_core<T>(left ? BLIS_LEFT : BLIS_RIGHT, uplo_of(*uploa), bt,
                 unita ? BLIS_UNIT_DIAG : BLIS_NONUNIT_DIAG, *m, *n, alpha, a, 1, *lda, b, 1,
                 *ldb);
}

template <typename T>
void blas_syrk(const char* dtc, const f77_char* uploc, const f77_char* transa,
               const f77_int* m, const f77_int* k, const T* alpha, const T* a,
               const f77_int* lda, const T* beta, T* c, const f77_int* ldc) {
    // frame/compat/check/bla_syrk_check.h ('C' is accepted for real types)
    const bool lower = up(*uploc) == 'L', upper = up(*uploc) == 'U';
    const bool nota = up(*transa) == 'N', ta = up(*transa) == 'T', cta = up(*transa) == 'C';
    const f77_int nrowa = nota ? *m : *k;
    f77_int info = 0;
    if (!lower && !upper) info = 1;
    else if (!nota && !ta && !cta) info = 2;
    else if (*m < 0) info = 3;
    else if (*k < 0) info = 4;
    else if (*lda < std::max<f77_int>(1, nrowa)) info = 7;
    else if (*ldc < std::max<f77_int>(1, *m)) info = 10;
    if (info != 0) { blas_error(dtc, "syrk", info); return; }
    const trans_t bt = nota ? BLIS_NO_TRANSPOSE : BLIS_TRANSPOSE;
    gemmt_core<T>(uplo_of(*uploc), bt, toggle(bt), *m, *k, alpha, a, 1, *lda, a, 1, *lda,
                  beta, c, 1, *ldc);
}

template <typename T>
void blas_syr2k(const char* dtc, const f77_char* uploc, const f77_char* trans,
                const f77_int* m, const f77_int* k, const T* alpha, const T* a,
                const f77_int* lda, const T* b, const f77_int* ldb, const T* beta, T* c,
                const f77_int* ldc) {
    // frame/compat/check/bla_syr2k_check.h
    const bool lower = up(*uploc) == 'L', upper = up(*uploc) == 'U';
    const bool nota = up(*trans) == 'N', ta = up(*trans) == 'T', cta = up(*trans) == 'C';
    const f77_int nrowa = nota ? *m : *k;
    f77_int info = 0;
    if (!lower && !upper) info = 1;
    else if (!nota && !ta && !cta) info = 2;
    else if (*m < 0) info = 3;
    else if (*k < 0) info = 4;
    else if (*lda < std::max<f77_int>(1, nrowa)) info = 7;
    else if (*ldb < std::max<f77_int>(1, nrowa)) info = 9;
    else if (*ldc < std::max<f77_int>(1, *m)) info = 12;
    if (info != 0) { blas_error(dtc, "syr2k", info); return; }
    const trans_t bt = nota ? BLIS_NO_TRANSPOSE : BLIS_TRANSPOSE;
    syr2k_core<T>(uplo_of(*uploc), bt, bt, *m, *k, alpha, a, 1, *lda, b, 1, *ldb, beta, c, 1,
                  *ldc);
}

template <typename T>
void blas_gemmt(const char* dtc, const f77_char* uploc, const f77_char* transa,
                const f77_char* transb, const f77_int* m, const f77_int* k, const T* alpha,
                const T* a, const f77_int* lda, const T* b, const f77_int* ldb, const T* beta,
                T* c, const f77_int* ldc) {
    // frame/compat/check/bla_gemmt_check.h
    const bool nota = up(*transa) == 'N', notb = up(*transb) == 'N';
    const bool conja = up(*transa) == 'C', conjb = up(*transb) == 'C';
    const bool ta = up(*transa) == 'T', tb = up(*transb) == 'T';
    const bool lower = up(*uploc) == 'L', upper = up(*uploc) == 'U';
    const f77_int nrowa = nota ? *m : *k, nrowb = notb ? *k : *m;
    f77_int info = 0;
    if (!lower && !upper) info = 1;
    else if (!nota && !conja && !ta) info = 2;
    else if (!notb && !conjb && !tb) info = 3;
    else if (*m < 0) info = 4;
    else if (*k < 0) info = 5;
    else if (*lda < std::max<f77_int>(1, nrowa)) info = 8;
    else if (*ldb < std::max<f77_int>(1, nrowb)) info = 10;
    else if (*ldc < std::max<f77_int>(1, *m)) info = 13;
    if (info != 0) { blas_error(dtc, "gemmt", info); return; }
    trans_t bta, btb;
    map_trans(*transa, &bta);
    map_trans(*transb, &btb);
    gemmt_core<T>(uplo_of(*uploc), bta, btb, *m, *k, alpha, a, 1, *lda, b, 1, *ldb, beta, c, 1,
                  *ldc);
}

} // namespace

#define STRUC_BLAS(ch, T, CH)                                                          \
BLIS_EXPORT void ch##symm_(const f77_char* side, const f77_char* uploa,                \
        const f77_int* m, const f77_int* n, const T* alpha, const T* a,                \
        const f77_int* lda, const T* b, const f77_int* ldb, const T* beta, T* c,       \
        const f77_int* ldc) {                                                          \
    blas_symm<T>(#ch, side, uploa, m, n, alpha, a, lda, b, ldb, beta, c, ldc);         \
}                                                                                      \
BLIS_EXPORT void ch##trmm_(const f77_char* side, const f77_char* uploa,                \
        const f77_char* transa, const f77_char* diaga, const f77_int* m,               \
        const f77_int* n, const T* alpha, const T* a, const f77_int* lda, T* b,        \
        const f77_int* ldb) {                                                          \
    blas_trmm<T>(#ch, side, uploa, transa, diaga, m, n, alpha, a, lda, b, ldb);        \
}                                                                                      \
BLIS_EXPORT void ch##syrk_(const f77_char* uploc, const f77_char* transa,              \
        const f77_int* m, const f77_int* k, const T* alpha, const T* a,                \
        const f77_int* lda, const T* beta, T* c, const f77_int* ldc) {                 \
    blas_syrk<T>(#ch, uploc, transa, m, k, alpha, a, lda, beta, c, ldc);               \
}                                                                                      \
BLIS_EXPORT void ch##syr2k_(const f77_char* uploc, const f77_char* trans,              \
        const f77_int* m, const f77_int* k, const T* alpha, const T* a,                \
        const f77_int* lda, const T* b, const f77_int* ldb, const T* beta, T* c,       \
        const f77_int* ldc) {                                                          \
    blas_syr2k<T>(#ch, uploc, trans, m, k, alpha, a, lda, b, ldb, beta, c, ldc);       \
}                                                                                      \
BLIS_EXPORT void ch##gemmt_(const f77_char* uploc, const f77_char* transa,             \
        const f77_char* transb, const f77_int* m, const f77_int* k, const T* alpha,    \
        const T* a, const f77_int* lda, const T* b, const f77_int* ldb, const T* beta, \
        T* c, const f77_int* ldc) {                                                    \
    blas_gemmt<T>(#ch, uploc, transa, transb, m, k, alpha, a, lda, b, ldb, beta, c,    \
                  ldc);                                                                \
}
STRUC_BLAS(s, float, "S")
STRUC_BLAS(d, double, "D")
