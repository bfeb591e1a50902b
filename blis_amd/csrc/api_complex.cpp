// api_complex.cpp — complex GEMM (typed, object and BLAS interfaces) and
// mixed-datatype GEMM through the object API.
//
// Reference interfaces and behaviour (paths relative to the BLIS tree):
//   bli_cgemm/bli_zgemm[_ex]  frame/3/bli_l3_tapi_ex.c:43-99 (scomplex and
//                             dcomplex instances of the typed GEMM)
//   cgemm_/zgemm_             frame/compat/bla_gemm.c:122-260, argument checks
//                             frame/compat/check/bla_gemm_check.h:39-82
//   cgemm_batch_/zgemm_batch_ frame/compat/extra/bla_gemm_batch.c:135-246
//   complex object GEMM       frame/3/bli_l3_oapi_ex.c:48-148 with the 1m
//                             induced method (frame/3/bli_l3_ind.c)
//   mixed datatypes           docs/MixedDatatypes.md: the computation
//                             precision is C's comp_prec field (default: C's
//                             storage precision); the computation domain is
//                             complex when A and B are complex, or when C and
//                             one of A, B are complex, else real; a complex
//                             operand in a real computation contributes its
//                             real part; a real C keeps only the real part of
//                             a complex product; a complex C updated by a real
//                             product keeps its imaginary parts.
// The complex product runs on the real GEMM tiles in 1m form (complex.hip).
// A mixed-datatype call converts A and B to the computation datatype where
// they differ from it (the typecast BLIS does while packing), forms the
// product in that datatype, and accumulates it into C with C's type.
#include "api_common.hpp"

#include <algorithm>
#include <vector>

using namespace blis_amd;
using namespace blis_amd::api;

#define BLIS_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

template <typename C> struct Real;
template <> struct Real<scomplex> { using type = float; static constexpr Dt dt = Dt::F32; };
template <> struct Real<dcomplex> { using type = double; static constexpr Dt dt = Dt::F64; };

template <typename C>
C read_cx(const C* p) { return read_scalar(p); }

inline bool does_conj(trans_t t) { return ((unsigned)t & BLIS_CONJ_BIT) != 0; }

template <typename C>
void zgemm_core(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k, const C* alpha,
                const C* a, inc_t rs_a, inc_t cs_a, const C* b, inc_t rs_b, inc_t cs_b,
                const C* beta, C* c, inc_t rs_c, inc_t cs_c) {
    init_once();
    require(valid_trans(transa), BLIS_INVALID_TRANS, "gemm: transa");
    require(valid_trans(transb), BLIS_INVALID_TRANS, "gemm: transb");
    require(m >= 0 && n >= 0 && k >= 0, BLIS_NEGATIVE_DIMENSION, "gemm: dimensions");
    if (m <= 0 || n <= 0) return;
    const C al = read_cx(alpha), be = read_cx(beta);
    const Dt dt = Real<C>::dt;
    const bool ta = does_trans(transa), tb = does_trans(transb);
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb, sc;
    const bool touch_ab = k > 0 && (al.real != 0 || al.imag != 0);
    if (touch_ab) {
        stage_complex(sa, a, dt, ta ? k : m, ta ? m : k, rs_a, cs_a, false);
        stage_complex(sb, b, dt, tb ? n : k, tb ? k : n, rs_b, cs_b, false);
    }
    stage_complex(sc, c, dt, m, n, rs_c, cs_c, true);
    ZGemmDesc z;
    z.dt = dt;
    z.m = m; z.n = n; z.k = k;
    z.a = ta ? View{ sa.dev, cs_a, rs_a } : View{ sa.dev, rs_a, cs_a };
    z.b = tb ? View{ sb.dev, cs_b, rs_b } : View{ sb.dev, rs_b, cs_b };
    z.c = View{ sc.dev, rs_c, cs_c };
    z.conja = does_conj(transa);
    z.conjb = does_conj(transb);
    z.alpha_re = al.real; z.alpha_im = al.imag;
    z.beta_re = be.real; z.beta_im = be.imag;
    zgemm_run(z, stream());
    sync_if_needed();
}

// ---- object helpers ---------------------------------------------------------

inline bool dt_is_fp(num_t d) {
    return d == BLIS_FLOAT || d == BLIS_DOUBLE || d == BLIS_SCOMPLEX || d == BLIS_DCOMPLEX;
}
inline bool dt_is_cx(num_t d) { return ((unsigned)d & BLIS_DOMAIN_BIT) != 0; }
inline Dt dt_prec(num_t d) { return ((unsigned)d & BLIS_BITVAL_DOUBLE_PREC) ? Dt::F64 : Dt::F32; }

// Value of a 1x1 scalar object as a complex double.
void obj_scalar_cx(const obj_t* o, double* re, double* im) {
    require(o != nullptr, BLIS_NULL_POINTER, "scalar object");
    const num_t dt = bli_obj_dt(o);
    if (dt == BLIS_CONSTANT) {
        const ConstData* k = (const ConstData*)o->buffer;
        *re = k->z.real; *im = k->z.imag;
        return;
    }
    const void* p = obj_origin(o);
    if (dt == BLIS_DCOMPLEX) {
        const dcomplex v = read_scalar((const dcomplex*)p);
        *re = v.real; *im = v.imag;
    } else if (dt == BLIS_SCOMPLEX) {
        const scomplex v = read_scalar((const scomplex*)p);
        *re = v.real; *im = v.imag;
    } else {
        *re = obj_scalar(o);
        *im = 0.0;
    }
}

void check_obj_any_fp(const obj_t* o, const char* where) {
    require(o != nullptr, BLIS_NULL_POINTER, where);
    require(dt_is_fp(bli_obj_dt(o)), BLIS_EXPECTED_FLOATING_POINT_DATATYPE, where);
    require(o->dim[0] >= 0 && o->dim[1] >= 0, BLIS_NEGATIVE_DIMENSION, where);
    require(o->buffer != nullptr || o->dim[0] == 0 || o->dim[1] == 0,
            BLIS_EXPECTED_NONNULL_OBJECT_BUFFER, where);
}

void stage_obj(Staged& s, const obj_t* o, dim_t m, dim_t n, bool out) {
    const num_t d = bli_obj_dt(o);
    if (dt_is_cx(d)) stage_complex(s, obj_origin(o), dt_prec(d), m, n, o->rs, o->cs, out);
    else stage(s, obj_origin(o), dt_prec(d), m, n, o->rs, o->cs, out);
}

} // namespace

namespace blis_amd {
namespace api {

void gemm_obj_md(const obj_t* alpha, const obj_t* a, const obj_t* b, const obj_t* beta,
                 const obj_t* c) {
    init_once();
    check_obj_any_fp(a, "bli_gemm: a");
    check_obj_any_fp(b, "bli_gemm: b");
    check_obj_any_fp(c, "bli_gemm: c");
    const dim_t m = c->dim[0], n = c->dim[1], k = bli_obj_width_after_trans(a);
    require(bli_obj_length_after_trans(a) == m && bli_obj_width_after_trans(b) == n &&
            bli_obj_length_after_trans(b) == k, BLIS_NONCONFORMAL_DIMENSIONS, "bli_gemm");
    require(bli_obj_struc(c) == BLIS_GENERAL, BLIS_EXPECTED_GENERAL_OBJECT, "bli_gemm: c");
    if (m <= 0 || n <= 0) return;
    double ar, ai, br, bi;
    obj_scalar_cx(alpha, &ar, &ai);
    obj_scalar_cx(beta, &br, &bi);
    const num_t da = bli_obj_dt(a), db = bli_obj_dt(b), dc = bli_obj_dt(c);
    const bool acx = dt_is_cx(da), bcx = dt_is_cx(db), ccx = dt_is_cx(dc);
    const trans_t ta = bli_obj_conjtrans_status(a), tb = bli_obj_conjtrans_status(b);
    // Computation datatype (docs/MixedDatatypes.md).
    const unsigned cp = (c->info & BLIS_COMP_PREC_BIT) >> BLIS_COMP_PREC_SHIFT;
    const Dt pdt = (cp & BLIS_BITVAL_DOUBLE_PREC) ? Dt::F64 : Dt::F32;
    const bool comp_cx = (acx && bcx) || (ccx && (acx || bcx));
    const num_t comp_dt = (num_t)((comp_cx ? BLIS_BITVAL_COMPLEX : 0u) |
                                  (pdt == Dt::F64 ? BLIS_BITVAL_DOUBLE_PREC : 0u));
    const Dt cdt = dt_prec(dc);
    const bool touch_ab = k > 0 && (ar != 0.0 || (comp_cx && ai != 0.0));

    std::lock_guard<std::mutex> lock(api_mutex());
    hipStream_t st = stream();
    Staged sa, sb, sc;
    stage_obj(sc, c, m, n, true);
    const View cv{ sc.dev, c->rs, c->cs };
    if (!touch_ab) {
        // C := beta*C (bli_l3_return_early_if_trivial).
        if (ccx) zscalm_run(cdt, m, n, br, bi, cv, st);
        else scalm_run(cdt, m, n, br, cv, st);
        sync_if_needed();
        return;
    }
    stage_obj(sa, a, does_trans(ta) ? k : m, does_trans(ta) ? m : k, false);
    stage_obj(sb, b, does_trans(tb) ? n : k, does_trans(tb) ? k : n, false);
    // op(A) and op(B) views (conjugation is applied in the product).
    View av = does_trans(ta) ? View{ sa.dev, a->cs, a->rs } : View{ sa.dev, a->rs, a->cs };
    View bv = does_trans(tb) ? View{ sb.dev, b->cs, b->rs } : View{ sb.dev, b->rs, b->cs };
    const size_t es = dt_size(pdt) * (comp_cx ? 2 : 1);
    // Operands not already in the computation datatype are converted into
    // column-major copies (slots 25, 26).
    if (da != comp_dt) {
        void* w = workspace((size_t)std::max<dim_t>(m * k, 1) * es, 25);
        md_convert_run(dt_prec(da), acx, av.ptr, m, k, av.rs, av.cs, pdt, comp_cx, w, m, st);
        av = View{ w, 1, m };
    }
    if (db != comp_dt) {
        void* w = workspace((size_t)std::max<dim_t>(k * n, 1) * es, 26);
        md_convert_run(dt_prec(db), bcx, bv.ptr, k, n, bv.rs, bv.cs, pdt, comp_cx, w, k, st);
        bv = View{ w, 1, k };
    }
    // The product, straight into C when C has the computation datatype,
    // else into a column-major copy (slot 27) accumulated into C after.
    const bool direct = dc == comp_dt;
    View pv = direct ? cv : View{ workspace((size_t)m * n * es, 27), 1, m };
    if (comp_cx) {
        ZGemmDesc z;
        z.dt = pdt;
        z.m = m; z.n = n; z.k = k;
        z.a = av; z.b = bv; z.c = pv;
        z.conja = ((unsigned)ta & BLIS_CONJ_BIT) && acx;
        z.conjb = ((unsigned)tb & BLIS_CONJ_BIT) && bcx;
        z.alpha_re = ar; z.alpha_im = ai;
        z.beta_re = direct ? br : 0.0;
        z.beta_im = direct ? bi : 0.0;
        zgemm_run(z, st);
    } else {
        GemmDesc d;
        d.dt = pdt;
        d.m = m; d.n = n; d.k = k;
        d.a = av; d.b = bv; d.c = pv;
        d.alpha = ar;
        d.beta = direct ? br : 0.0;
        gemm_run(d, nullptr, st);
    }
    if (!direct) md_accum_run(pdt, comp_cx, pv.ptr, m, m, n, br, bi, cdt, ccx, cv, st);
    sync_if_needed();
}

} // namespace api
} // namespace blis_amd

// ===========================================================================
// Typed complex GEMM
// ===========================================================================

#define ZGEMM_TYPED(ch, C)                                                             \
BLIS_EXPORT void bli_##ch##gemm_ex(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k, \
        const C* alpha, const C* a, inc_t rs_a, inc_t cs_a, const C* b, inc_t rs_b,    \
        inc_t cs_b, const C* beta, C* c, inc_t rs_c, inc_t cs_c, const cntx_t* cntx,   \
        const rntm_t* rntm) {                                                          \
    (void)cntx; (void)rntm;                                                            \
    zgemm_core<C>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta,  \
                  c, rs_c, cs_c);                                                      \
}                                                                                      \
BLIS_EXPORT void bli_##ch##gemm(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k, \
        const C* alpha, const C* a, inc_t rs_a, inc_t cs_a, const C* b, inc_t rs_b,    \
        inc_t cs_b, const C* beta, C* c, inc_t rs_c, inc_t cs_c) {                     \
    zgemm_core<C>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta,  \
                  c, rs_c, cs_c);                                                      \
}
ZGEMM_TYPED(c, scomplex)
ZGEMM_TYPED(z, dcomplex)

// ===========================================================================
// BLAS
// ===========================================================================

namespace {

template <typename C>
bool blas_gemm_args_ok(const char* dtc, const f77_char* transa, const f77_char* transb,
                       f77_int m, f77_int n, f77_int k, f77_int lda, f77_int ldb, f77_int ldc) {
    // frame/compat/check/bla_gemm_check.h:39-82
    const bool nota = up(*transa) == 'N', notb = up(*transb) == 'N';
    const bool conja = up(*transa) == 'C', conjb = up(*transb) == 'C';
    const bool ta = up(*transa) == 'T', tb = up(*transb) == 'T';
    const f77_int nrowa = nota ? m : k, nrowb = notb ? k : n;
    f77_int info = 0;
    if (!nota && !conja && !ta) info = 1;
    else if (!notb && !conjb && !tb) info = 2;
    else if (m < 0) info = 3;
    else if (n < 0) info = 4;
    else if (k < 0) info = 5;
    else if (lda < std::max<f77_int>(1, nrowa)) info = 8;
    else if (ldb < std::max<f77_int>(1, nrowb)) info = 10;
    else if (ldc < std::max<f77_int>(1, m)) info = 13;
    if (info != 0) { blas_error(dtc, "gemm", info); return false; }
    return true;
}

template <typename C>
void blas_zgemm(const char* dtc, const f77_char* transa, const f77_char* transb,
                const f77_int* m, const f77_int* n, const f77_int* k, const C* alpha,
                const C* a, const f77_int* lda, const C* b, const f77_int* ldb, const C* beta,
                C* c, const f77_int* ldc) {
    if (!blas_gemm_args_ok<C>(dtc, transa, transb, *m, *n, *k, *lda, *ldb, *ldc)) return;
    trans_t bta, btb;
    map_trans(*transa, &bta);
    map_trans(*transb, &btb);
    zgemm_core<C>(bta, btb, *m, *n, *k, alpha, a, 1, *lda, b, 1, *ldb, beta, c, 1, *ldc);
}

template <typename C>
void blas_zgemm_batch(const char* dtc, const f77_char* transa_array,
                      const f77_char* transb_array, const f77_int* m_array,
                      const f77_int* n_array, const f77_int* k_array, const C* alpha_array,
                      const C** a_array, const f77_int* lda_array, const C** b_array,
                      const f77_int* ldb_array, const C* beta_array, C** c_array,
                      const f77_int* ldc_array, const f77_int* group_count,
                      const f77_int* group_size) {
    // Every group is checked before any problem runs (bla_gemm_batch.c:152-168),
    // then the problems run in order, one GEMM each (:172-241).
    for (f77_int g = 0; g < *group_count; ++g)
        if (!blas_gemm_args_ok<C>(dtc, transa_array + g, transb_array + g, m_array[g],
                                  n_array[g], k_array[g], lda_array[g], ldb_array[g],
                                  ldc_array[g]))
            return;
    f77_int idx = 0;
    for (f77_int g = 0; g < *group_count; ++g) {
        trans_t ta, tb;
        map_trans(transa_array[g], &ta);
        map_trans(transb_array[g], &tb);
        for (f77_int s = 0; s < group_size[g]; ++s, ++idx)
            zgemm_core<C>(ta, tb, m_array[g], n_array[g], k_array[g], &alpha_array[g],
                          a_array[idx], 1, lda_array[g], b_array[idx], 1, ldb_array[g],
                          &beta_array[g], c_array[idx], 1, ldc_array[g]);
    }
}

} // namespace

#define ZGEMM_BLAS(ch, C)                                                              \
BLIS_EXPORT void ch##gemm_(const f77_char* transa, const f77_char* transb,             \
        const f77_int* m, const f77_int* n, const f77_int* k, const C* alpha,          \
        const C* a, const f77_int* lda, const C* b, const f77_int* ldb, const C* beta, \
        C* c, const f77_int* ldc) {                                                    \
    blas_zgemm<C>(#ch, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);  \
}                                                                                      \
BLIS_EXPORT void ch##gemm_batch_(const f77_char* transa_array,                         \
        const f77_char* transb_array, const f77_int* m_array, const f77_int* n_array,  \
        const f77_int* k_array, const C* alpha_array, const C** a_array,               \
        const f77_int* lda_array, const C** b_array, const f77_int* ldb_array,         \
        const C* beta_array, C** c_array, const f77_int* ldc_array,                    \
        const f77_int* group_count, const f77_int* group_size) {                       \
    blas_zgemm_batch<C>(#ch, transa_array, transb_array, m_array, n_array, k_array,    \
                        alpha_array, a_array, lda_array, b_array, ldb_array,           \
                        beta_array, c_array, ldc_array, group_count, group_size);      \
}
ZGEMM_BLAS(c, scomplex)
ZGEMM_BLAS(z, dcomplex)
