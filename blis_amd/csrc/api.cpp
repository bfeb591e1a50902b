// api.cpp — the public BLIS interfaces of this library: initialisation and
// information queries, obj_t management, the typed and object APIs for
// gemm / trsm / gemm_batch / axpyv / dotv / gemv, and the BLAS compatibility
// symbols. Every entry point validates its arguments the way the reference
// does, stages host operands (runtime.cpp), builds a descriptor and calls the
// launch planner of the corresponding kernel family.
//
// Reference behaviour reproduced here (paths relative to the BLIS tree):
//   trivial-case return        frame/3/bli_l3_util.c:41-65
//   typed -> object wrapping    frame/3/bli_l3_tapi_ex.c:43-99, 532-586
//   right-side trsm             frame/3/bli_l3_oapi_ex.c:753-758 (by transposition)
//   gemv early exits            frame/2/bli_l2_tapi.c:43-123
//   axpyv / dotv early exits    ref_kernels/1/bli_axpyv_ref.c:38-121,
//                               ref_kernels/1/bli_dotv_ref.c:38-122
//   BLAS gemm (m==1/n==1->gemv) frame/compat/bla_gemm.c:122-260
//   BLAS argument checks        frame/compat/check/bla_gemm_check.h:39-82,
//                               bla_trmm_check.h (trsm), bla_gemv_check.h
//   BLAS gemv n_x == 0 quirk    frame/compat/bla_gemv.c:92-111
//   negative BLAS increments    frame/include/bli_blas_macro_defs.h:50-74
//   gemm_batch group indexing   frame/compat/extra/bla_gemm_batch.c:135-246
//   xerbla message              frame/compat/f2c/bla_xerbla.c
#include "api_common.hpp"

#include <cctype>
#include <cstring>
#include <deque>
#include <set>

using namespace blis_amd;
using namespace blis_amd::api;

#define BLIS_EXPORT extern "C" __attribute__((visibility("default")))

namespace blis_amd {
namespace api {
errlev_t& errlev() {
    static errlev_t level = BLIS_FULL_ERROR_CHECKING;
    return level;
}
} // namespace api
} // namespace blis_amd

namespace {

ConstData k_two  = { 2.0f,  2.0, { 2.0f, 0.0f }, { 2.0, 0.0 },  2 };
ConstData k_one   = { 1.0f,  1.0, { 1.0f, 0.0f }, { 1.0, 0.0 },  1 };
ConstData k_zero  = { 0.0f,  0.0, { 0.0f, 0.0f }, { 0.0, 0.0 },  0 };
ConstData k_mone  = {-1.0f, -1.0, {-1.0f, 0.0f }, {-1.0, 0.0 }, -1 };
ConstData k_mtwo  = {-2.0f, -2.0, {-2.0f, 0.0f }, {-2.0, 0.0 }, -2 };

obj_t make_const(ConstData* p) {
    obj_t o = BLIS_OBJECT_INITIALIZER_1X1;
    o.info = ( o.info & ~BLIS_DATATYPE_BITS ) | BLIS_CONSTANT;
    o.elem_size = sizeof( ConstData );
    o.buffer = p;
    o.rs = 1; o.cs = 1; o.is = 1;
    return o;
}


std::mutex g_alloc_mutex;
std::set<void*> g_owned;   // buffers allocated by bli_obj_create

// ---- level-3 cores ------------------------------------------------------

template <typename T>
void gemm_core(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
               const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
               const T* b, inc_t rs_b, inc_t cs_b,
               const T* beta, T* c, inc_t rs_c, inc_t cs_c) {
    init_once();
    require(valid_trans(transa), BLIS_INVALID_TRANS, "gemm: transa");
    require(valid_trans(transb), BLIS_INVALID_TRANS, "gemm: transb");
    require(m >= 0 && n >= 0 && k >= 0, BLIS_NEGATIVE_DIMENSION, "gemm: dimensions");
    if (m <= 0 || n <= 0) return;
    const double al = (double)*alpha, be = (double)*beta;
    const bool ta = does_trans(transa), tb = does_trans(transb);
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb, sc;
    const bool touch_ab = (k > 0 && al != 0.0);
    if (touch_ab) {
        stage(sa, a, dt, ta ? k : m, ta ? m : k, rs_a, cs_a, false);
        stage(sb, b, dt, tb ? n : k, tb ? k : n, rs_b, cs_b, false);
    }
    stage(sc, c, dt, m, n, rs_c, cs_c, true);
    GemmDesc d;
    d.dt = dt; d.m = m; d.n = n; d.k = k;
    d.a = { sa.dev, ta ? cs_a : rs_a, ta ? rs_a : cs_a };
    d.b = { sb.dev, tb ? cs_b : rs_b, tb ? rs_b : cs_b };
    d.c = { sc.dev, rs_c, cs_c };
    d.alpha = al; d.beta = be;
    gemm_run(d, nullptr, stream());
    sync_if_needed();
}

template <typename T>
void trsm_core(side_t side, uplo_t uploa, trans_t transa, diag_t diaga, dim_t m, dim_t n,
               const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
               T* b, inc_t rs_b, inc_t cs_b) {
    init_once();
    require(side == BLIS_LEFT || side == BLIS_RIGHT, BLIS_INVALID_SIDE, "trsm: side");
    require(uploa == BLIS_LOWER || uploa == BLIS_UPPER, BLIS_INVALID_UPLO, "trsm: uplo");
    require(valid_trans(transa), BLIS_INVALID_TRANS, "trsm: transa");
    require(diaga == BLIS_NONUNIT_DIAG || diaga == BLIS_UNIT_DIAG, BLIS_INVALID_DIAG, "trsm: diag");
    require(m >= 0 && n >= 0, BLIS_NEGATIVE_DIMENSION, "trsm: dimensions");
    if (m <= 0 || n <= 0) return;
    const double al = (double)*alpha;
    const Dt dt = dt_for<T>();
    const dim_t mn = (side == BLIS_LEFT) ? m : n;
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb;
    stage(sb, b, dt, m, n, rs_b, cs_b, true);
    if (al == 0.0) {
        // alpha == 0: B := 0 (bli_trsm_ex passes BLIS_ZERO as beta to the
        // trivial-case check, frame/3/bli_l3_oapi_ex.c:710).
        scalm_run(dt, m, n, 0.0, View{ sb.dev, rs_b, cs_b }, stream());
        sync_if_needed();
        return;
    }
    stage(sa, a, dt, mn, mn, rs_a, cs_a, false);
    trsm_run(dt, side, uploa, transa, diaga, m, n, al, View{ sa.dev, rs_a, cs_a },
             View{ sb.dev, rs_b, cs_b }, stream());
    sync_if_needed();
}

template <typename T>
void gemm_batch_core(const trans_t* transa, const trans_t* transb,
                     const dim_t* m, const dim_t* n, const dim_t* k,
                     const T* alpha, const T* const* a, const inc_t* rs_a, const inc_t* cs_a,
                     const T* const* b, const inc_t* rs_b, const inc_t* cs_b,
                     const T* beta, T* const* c, const inc_t* rs_c, const inc_t* cs_c,
                     dim_t group_count, const dim_t* group_size) {
    init_once();
    require(group_count >= 0, BLIS_NEGATIVE_DIMENSION, "gemm_batch: group_count");
    const Dt dt = dt_for<T>();
    std::vector<BatchProblem> probs;
    std::deque<Staged> stages;
    dim_t total = 0;
    for (dim_t g = 0; g < group_count; ++g) {
        require(group_size[g] >= 0, BLIS_NEGATIVE_DIMENSION, "gemm_batch: group_size");
        require(valid_trans(transa[g]) && valid_trans(transb[g]), BLIS_INVALID_TRANS,
                "gemm_batch: trans");
        require(m[g] >= 0 && n[g] >= 0 && k[g] >= 0, BLIS_NEGATIVE_DIMENSION,
                "gemm_batch: dimensions");
        total += group_size[g] > 0 ? group_size[g] : 0;
    }
    std::lock_guard<std::mutex> lock(api_mutex());
    probs.reserve((size_t)total);
    DeviceRanges ranges;
    dim_t idx = 0;
    for (dim_t g = 0; g < group_count; ++g) {
        const bool ta = does_trans(transa[g]), tb = does_trans(transb[g]);
        for (dim_t s = 0; s < group_size[g]; ++s, ++idx) {
            if (m[g] <= 0 || n[g] <= 0) continue;
            BatchProblem p;
            p.m = m[g]; p.n = n[g]; p.k = k[g];
            p.alpha = (double)alpha[g]; p.beta = (double)beta[g];
            const bool touch_ab = (p.k > 0 && p.alpha != 0.0);
            if (!touch_ab) p.k = 0;   // C := beta*C without reading A or B
            stages.emplace_back();
            Staged& sa = stages.back();
            if (touch_ab) stage(sa, a[idx], dt, ta ? k[g] : m[g], ta ? m[g] : k[g], rs_a[g], cs_a[g], false, &ranges);
            stages.emplace_back();
            Staged& sb = stages.back();
            if (touch_ab) stage(sb, b[idx], dt, tb ? n[g] : k[g], tb ? k[g] : n[g], rs_b[g], cs_b[g], false, &ranges);
            stages.emplace_back();
            Staged& sc = stages.back();
            stage(sc, c[idx], dt, m[g], n[g], rs_c[g], cs_c[g], true, &ranges);
            p.a = { sa.dev, ta ? cs_a[g] : rs_a[g], ta ? rs_a[g] : cs_a[g] };
            p.b = { sb.dev, tb ? cs_b[g] : rs_b[g], tb ? rs_b[g] : cs_b[g] };
            p.c = { sc.dev, rs_c[g], cs_c[g] };
            probs.push_back(p);
        }
    }
    if (!probs.empty()) {
        gemm_batch_run(dt, probs, stream());
        sync_if_needed();
    }
}

// ---- level-1/2 cores ----------------------------------------------------

template <typename T>
void axpyv_core(dim_t n, const T* alpha, const T* x, inc_t incx, T* y, inc_t incy) {
    init_once();
    if (n <= 0) return;
    const double al = (double)*alpha;
    if (al == 0.0) return;
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sx, sy;
    stage_vec(sx, x, dt, n, incx, false);
    stage_vec(sy, y, dt, n, incy, true);
    axpyv_run(dt, n, al, sx.dev, incx, sy.dev, incy, stream());
    sync_if_needed();
}

template <typename T>
void dotv_core(dim_t n, const T* x, inc_t incx, const T* y, inc_t incy, T* rho) {
    init_once();
    const Dt dt = dt_for<T>();
    const bool rho_dev = is_device_pointer(rho);
    std::lock_guard<std::mutex> lock(api_mutex());
    T* out = rho_dev ? rho : (T*)workspace(sizeof(T), 7);
    Staged sx, sy;
    if (n > 0) {
        stage_vec(sx, x, dt, n, incx, false);
        stage_vec(sy, y, dt, n, incy, false);
    }
    dotv_run(dt, n > 0 ? n : 0, sx.dev, incx, sy.dev, incy, out, stream());
    if (!rho_dev) {
        HIP_CHECK(hipMemcpyAsync(rho, out, sizeof(T), hipMemcpyDeviceToHost, stream()));
        HIP_CHECK(hipStreamSynchronize(stream()));
    } else {
        sync_if_needed();
    }
}

template <typename T>
void gemv_core(trans_t transa, dim_t m, dim_t n, const T* alpha,
               const T* a, inc_t rs_a, inc_t cs_a, const T* x, inc_t incx,
               const T* beta, T* y, inc_t incy) {
    init_once();
    require(valid_trans(transa), BLIS_INVALID_TRANS, "gemv: transa");
    require(m >= 0 && n >= 0, BLIS_NEGATIVE_DIMENSION, "gemv: dimensions");
    const bool ta = does_trans(transa);
    const dim_t m_y = ta ? n : m, n_x = ta ? m : n;
    if (m_y <= 0) return;
    const double al = (double)*alpha, be = (double)*beta;
    const Dt dt = dt_for<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sx, sy;
    stage_vec(sy, y, dt, m_y, incy, true);
    if (n_x <= 0 || al == 0.0) {
        scalv_run(dt, m_y, be, sy.dev, incy, stream());
        sync_if_needed();
        return;
    }
    stage(sa, a, dt, m, n, rs_a, cs_a, false);
    stage_vec(sx, x, dt, n_x, incx, false);
    gemv_run(dt, transa, m, n, al, View{ sa.dev, rs_a, cs_a }, sx.dev, incx, be, sy.dev, incy,
             stream());
    sync_if_needed();
}

// ---- obj_t helpers --------------------------------------------------------

// ---- BLAS helpers --------------------------------------------------------

template <typename T>
void blas_gemm(const char* dtc, const f77_char* transa, const f77_char* transb,
               const f77_int* m, const f77_int* n, const f77_int* k,
               const T* alpha, const T* a, const f77_int* lda,
               const T* b, const f77_int* ldb, const T* beta, T* c, const f77_int* ldc) {
    // frame/compat/check/bla_gemm_check.h:39-82
    const bool nota = up(*transa) == 'N', notb = up(*transb) == 'N';
    const bool conja = up(*transa) == 'C', conjb = up(*transb) == 'C';
    const bool ta = up(*transa) == 'T', tb = up(*transb) == 'T';
    const f77_int nrowa = nota ? *m : *k, nrowb = notb ? *k : *n;
    f77_int info = 0;
    if (!nota && !conja && !ta) info = 1;
    else if (!notb && !conjb && !tb) info = 2;
    else if (*m < 0) info = 3;
    else if (*n < 0) info = 4;
    else if (*k < 0) info = 5;
    else if (*lda < std::max<f77_int>(1, nrowa)) info = 8;
    else if (*ldb < std::max<f77_int>(1, nrowb)) info = 10;
    else if (*ldc < std::max<f77_int>(1, *m)) info = 13;
    if (info != 0) { blas_error(dtc, "gemm", info); return; }

    trans_t bta, btb;
    map_trans(*transa, &bta);
    map_trans(*transb, &btb);
    const dim_t m0 = *m, n0 = *n, k0 = *k;
    const inc_t rs_a = 1, cs_a = *lda, rs_b = 1, cs_b = *ldb, rs_c = 1, cs_c = *ldc;
    // m == 1 or n == 1 goes to gemv (frame/compat/bla_gemm.c:181-222).
    if (n0 == 1) {
        const dim_t m0t = does_trans(bta) ? k0 : m0, k0t = does_trans(bta) ? m0 : k0;
        gemv_core<T>(bta, m0t, k0t, alpha, a, rs_a, cs_a,
                     b, does_trans(btb) ? cs_b : rs_b, beta, c, rs_c);
        return;
    }
    if (m0 == 1) {
        const dim_t n0t = does_trans(btb) ? k0 : n0, k0t = does_trans(btb) ? n0 : k0;
        gemv_core<T>(btb, n0t, k0t, alpha, b, cs_b, rs_b,
                     a, does_trans(bta) ? rs_a : cs_a, beta, c, cs_c);
        return;
    }
    gemm_core<T>(bta, btb, m0, n0, k0, alpha, a, rs_a, cs_a, b, rs_b, cs_b, beta, c, rs_c, cs_c);
}

template <typename T>
void blas_trsm(const char* dtc, const f77_char* side, const f77_char* uploa,
               const f77_char* transa, const f77_char* diaga,
               const f77_int* m, const f77_int* n, const T* alpha,
               const T* a, const f77_int* lda, T* b, const f77_int* ldb) {
    // frame/compat/check/bla_trmm_check.h (bla_trsm_check is bla_trmm_check)
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
    if (info != 0) { blas_error(dtc, "trsm", info); return; }
    trans_t bt;
    map_trans(*transa, &bt);
    trsm_core<T>(left ? BLIS_LEFT : BLIS_RIGHT, lower ? BLIS_LOWER : BLIS_UPPER, bt,
                 unita ? BLIS_UNIT_DIAG : BLIS_NONUNIT_DIAG, *m, *n, alpha,
                 a, 1, *lda, b, 1, *ldb);
}

template <typename T>
void blas_gemv(const char* dtc, const f77_char* transa, const f77_int* m, const f77_int* n,
               const T* alpha, const T* a, const f77_int* lda, const T* x, const f77_int* incx,
               const T* beta, T* y, const f77_int* incy) {
    // frame/compat/check/bla_gemv_check.h
    const bool nota = up(*transa) == 'N', ta = up(*transa) == 'T', conja = up(*transa) == 'C';
    f77_int info = 0;
    if (!nota && !ta && !conja) info = 1;
    else if (*m < 0) info = 2;
    else if (*n < 0) info = 3;
    else if (*lda < std::max<f77_int>(1, *m)) info = 6;
    else if (*incx == 0) info = 8;
    else if (*incy == 0) info = 11;
    if (info != 0) { blas_error(dtc, "gemv", info); return; }
    trans_t bt;
    map_trans(*transa, &bt);
    const dim_t m0 = *m, n0 = *n;
    const dim_t m_y = does_trans(bt) ? n0 : m0, n_x = does_trans(bt) ? m0 : n0;
    // BLAS returns without scaling y when x is empty (bla_gemv.c:92-111).
    if (m_y > 0 && n_x == 0) return;
    const T* x0 = (*incx < 0) ? x + (n_x - 1) * (-(inc_t)*incx) : x;
    T* y0 = (*incy < 0) ? y + (m_y - 1) * (-(inc_t)*incy) : y;
    gemv_core<T>(bt, m0, n0, alpha, a, 1, *lda, x0, *incx, beta, y0, *incy);
}

template <typename T>
void blas_axpy(const f77_int* n, const T* alpha, const T* x, const f77_int* incx,
               T* y, const f77_int* incy) {
    const dim_t n0 = *n < 0 ? 0 : *n;
    const T* x0 = (*incx < 0) ? x + (n0 - 1) * (-(inc_t)*incx) : x;
    T* y0 = (*incy < 0) ? y + (n0 - 1) * (-(inc_t)*incy) : y;
    axpyv_core<T>(n0, alpha, x0, *incx, y0, *incy);
}

template <typename T>
T blas_dot(const f77_int* n, const T* x, const f77_int* incx, const T* y, const f77_int* incy) {
    const dim_t n0 = *n < 0 ? 0 : *n;
    const T* x0 = (*incx < 0) ? x + (n0 - 1) * (-(inc_t)*incx) : x;
    const T* y0 = (*incy < 0) ? y + (n0 - 1) * (-(inc_t)*incy) : y;
    T rho = T(0);
    dotv_core<T>(n0, x0, *incx, y0, *incy, &rho);
    return rho;
}

template <typename T>
void blas_gemm_batch(const char* dtc, const f77_char* transa_array, const f77_char* transb_array,
                     const f77_int* m_array, const f77_int* n_array, const f77_int* k_array,
                     const T* alpha_array, const T** a_array, const f77_int* lda_array,
                     const T** b_array, const f77_int* ldb_array, const T* beta_array,
                     T** c_array, const f77_int* ldc_array,
                     const f77_int* group_count, const f77_int* group_size) {
    // Each group is checked as a gemm (bla_gemm_batch.c:152-168).
    for (f77_int g = 0; g < *group_count; ++g) {
        const char ta = up(transa_array[g]), tb = up(transb_array[g]);
        const bool nota = ta == 'N', notb = tb == 'N';
        const f77_int nrowa = nota ? m_array[g] : k_array[g];
        const f77_int nrowb = notb ? k_array[g] : n_array[g];
        f77_int info = 0;
        if (!nota && ta != 'C' && ta != 'T') info = 1;
        else if (!notb && tb != 'C' && tb != 'T') info = 2;
        else if (m_array[g] < 0) info = 3;
        else if (n_array[g] < 0) info = 4;
        else if (k_array[g] < 0) info = 5;
        else if (lda_array[g] < std::max<f77_int>(1, nrowa)) info = 8;
        else if (ldb_array[g] < std::max<f77_int>(1, nrowb)) info = 10;
        else if (ldc_array[g] < std::max<f77_int>(1, m_array[g])) info = 13;
        if (info != 0) { blas_error(dtc, "gemm", info); return; }
    }
    const dim_t G = *group_count;
    std::vector<trans_t> ta(G), tb(G);
    std::vector<dim_t> m(G), n(G), k(G), gs(G);
    std::vector<inc_t> ones(G, 1), lda(G), ldb(G), ldc(G);
    for (dim_t g = 0; g < G; ++g) {
        map_trans(transa_array[g], &ta[g]);
        map_trans(transb_array[g], &tb[g]);
        m[g] = m_array[g]; n[g] = n_array[g]; k[g] = k_array[g];
        lda[g] = lda_array[g]; ldb[g] = ldb_array[g]; ldc[g] = ldc_array[g];
        gs[g] = group_size[g];
    }
    gemm_batch_core<T>(ta.data(), tb.data(), m.data(), n.data(), k.data(), alpha_array,
                       a_array, ones.data(), lda.data(), b_array, ones.data(), ldb.data(),
                       beta_array, c_array, ones.data(), ldc.data(), G, gs.data());
}

} // namespace

// ===========================================================================
// Exported symbols
// ===========================================================================

BLIS_EXPORT obj_t BLIS_TWO       = make_const(&k_two);
BLIS_EXPORT obj_t BLIS_ONE       = make_const(&k_one);
BLIS_EXPORT obj_t BLIS_ZERO      = make_const(&k_zero);
BLIS_EXPORT obj_t BLIS_MINUS_ONE = make_const(&k_mone);
BLIS_EXPORT obj_t BLIS_MINUS_TWO = make_const(&k_mtwo);

BLIS_EXPORT void bli_init(void) { init_once(); }
BLIS_EXPORT void bli_finalize(void) {}
BLIS_EXPORT const char* bli_info_get_version_str(void) { return BLIS_VERSION_STRING; }
BLIS_EXPORT void bli_error_checking_level_set(errlev_t l) { errlev() = l; }
BLIS_EXPORT errlev_t bli_error_checking_level(void) { return errlev(); }
BLIS_EXPORT bool bli_error_checking_is_enabled(void) { return check_on(); }

BLIS_EXPORT const char* bli_error_string_for_code(gint_t code) {
    switch ((err_t)code) {
    case BLIS_SUCCESS: return "Success.";
    case BLIS_FAILURE: return "Failure.";
    case BLIS_NULL_POINTER: return "Encountered unexpected null pointer.";
    case BLIS_NOT_YET_IMPLEMENTED: return "Requested functionality not yet implemented.";
    case BLIS_INVALID_SIDE: return "Invalid side parameter value.";
    case BLIS_INVALID_UPLO: return "Invalid uplo_t parameter value.";
    case BLIS_INVALID_TRANS: return "Invalid trans_t parameter value.";
    case BLIS_INVALID_CONJ: return "Invalid conj_t parameter value.";
    case BLIS_INVALID_DIAG: return "Invalid diag_t parameter value.";
    case BLIS_INVALID_DATATYPE: return "Invalid datatype value.";
    case BLIS_EXPECTED_REAL_DATATYPE: return "Expected real datatype value.";
    case BLIS_INCONSISTENT_DATATYPES: return "Expected consistent datatypes (equal, or one being constant).";
    case BLIS_NONCONFORMAL_DIMENSIONS: return "Encountered non-conformal dimensions between objects.";
    case BLIS_EXPECTED_SCALAR_OBJECT: return "Expected scalar object.";
    case BLIS_EXPECTED_VECTOR_OBJECT: return "Expected vector object.";
    case BLIS_UNEQUAL_VECTOR_LENGTHS: return "Encountered unequal vector lengths.";
    case BLIS_EXPECTED_SQUARE_OBJECT: return "Expected square object.";
    case BLIS_NEGATIVE_DIMENSION: return "Encountered negative dimension.";
    case BLIS_EXPECTED_GENERAL_OBJECT: return "Expected general object.";
    case BLIS_EXPECTED_TRIANGULAR_OBJECT: return "Expected triangular object.";
    case BLIS_EXPECTED_UPPER_OR_LOWER_OBJECT: return "Expected upper or lower triangular object.";
    case BLIS_EXPECTED_NONNULL_OBJECT_BUFFER: return "Encountered object with non-NULL buffer.";
    case BLIS_MALLOC_RETURNED_NULL: return "malloc() returned NULL; heap memory is likely exhausted.";
    case BLIS_INVALID_UKR_ID: return "Invalid ukr_t value.";
    default: return "Undefined error code.";
    }
}

// ---- objects ----------------------------------------------------------------

BLIS_EXPORT void bli_obj_create_without_buffer(num_t dt, dim_t m, dim_t n, obj_t* obj) {
    require(obj != nullptr, BLIS_NULL_POINTER, "bli_obj_create_without_buffer");
    require(m >= 0 && n >= 0, BLIS_NEGATIVE_DIMENSION, "bli_obj_create_without_buffer");
    obj_t o = BLIS_OBJECT_INITIALIZER;
    o.root = obj;
    o.dim[0] = m; o.dim[1] = n;
    o.info = ( o.info & ~BLIS_DATATYPE_BITS ) | (objbits_t)dt;
    // The computation precision defaults to the storage precision
    // (frame/base/bli_obj.c:111).
    if (dt == BLIS_FLOAT || dt == BLIS_DOUBLE || dt == BLIS_SCOMPLEX || dt == BLIS_DCOMPLEX)
        o.info |= ( (objbits_t)dt & BLIS_PRECISION_BIT ) << BLIS_COMP_PREC_SHIFT;
    o.info2 = ( (objbits_t)dt << BLIS_SCALAR_DT_SHIFT ) & BLIS_SCALAR_DT_BITS;
    o.elem_size = (dt == BLIS_DOUBLE) ? 8 : (dt == BLIS_FLOAT) ? 4 :
                  (dt == BLIS_SCOMPLEX) ? 8 : (dt == BLIS_DCOMPLEX) ? 16 : 8;
    o.scalar.real = 1.0; o.scalar.imag = 0.0;
    *obj = o;
}

BLIS_EXPORT void bli_obj_attach_buffer(void* p, inc_t rs, inc_t cs, inc_t is, obj_t* obj) {
    if (rs == 0 && cs == 0) {
        // Default to column storage (frame/base/bli_obj.c, bli_adjust_strides).
        rs = 1;
        cs = obj->dim[0] > 0 ? obj->dim[0] : 1;
    }
    obj->buffer = p;
    obj->rs = rs; obj->cs = cs; obj->is = is;
}

BLIS_EXPORT void bli_obj_create_with_attached_buffer(num_t dt, dim_t m, dim_t n, void* p,
                                                     inc_t rs, inc_t cs, obj_t* obj) {
    bli_obj_create_without_buffer(dt, m, n, obj);
    bli_obj_attach_buffer(p, rs, cs, 1, obj);
}

// Buffers created here are device memory (the library's matrices live in
// HBM); bli_obj_free releases only buffers created by bli_obj_create.
BLIS_EXPORT void bli_obj_create(num_t dt, dim_t m, dim_t n, inc_t rs, inc_t cs, obj_t* obj) {
    init_once();
    bli_obj_create_without_buffer(dt, m, n, obj);
    if (rs == 0 && cs == 0) { rs = 1; cs = m > 0 ? m : 1; }
    const inc_t span_r = (m > 0 ? (m - 1) * (rs < 0 ? -rs : rs) : 0);
    const inc_t span_c = (n > 0 ? (n - 1) * (cs < 0 ? -cs : cs) : 0);
    size_t bytes = (size_t)(span_r + span_c + 1) * obj->elem_size;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        abort_with(BLIS_MALLOC_RETURNED_NULL, "bli_obj_create");
    }
    { std::lock_guard<std::mutex> g(g_alloc_mutex); g_owned.insert(p); }
    bli_obj_attach_buffer(p, rs, cs, 1, obj);
}

BLIS_EXPORT void bli_obj_create_1x1(num_t dt, obj_t* obj) {
    bli_obj_create(dt, 1, 1, 1, 1, obj);
}

BLIS_EXPORT void bli_obj_create_1x1_with_attached_buffer(num_t dt, void* p, obj_t* obj) {
    bli_obj_create_with_attached_buffer(dt, 1, 1, p, 1, 1, obj);
}

BLIS_EXPORT void bli_obj_scalar_init_detached(num_t dt, obj_t* beta) {
    bli_obj_create_without_buffer(dt, 1, 1, beta);
    beta->scalar.real = 0.0; beta->scalar.imag = 0.0;
    bli_obj_attach_buffer(&beta->scalar, 1, 1, 1, beta);
}

BLIS_EXPORT void bli_obj_free(obj_t* obj) {
    if (!obj || !obj->buffer) return;
    bool owned = false;
    {
        std::lock_guard<std::mutex> g(g_alloc_mutex);
        auto it = g_owned.find(obj->buffer);
        if (it != g_owned.end()) { owned = true; g_owned.erase(it); }
    }
    if (owned) HIP_CHECK(hipFree(obj->buffer));
    obj->buffer = nullptr;
}

BLIS_EXPORT void bli_setsc(double zeta_r, double zeta_i, obj_t* chi) {
    (void)zeta_i;
    const num_t dt = bli_obj_dt(chi);
    void* p = obj_origin(chi);
    if (dt == BLIS_DOUBLE) {
        if (is_device_pointer(p)) HIP_CHECK(hipMemcpy(p, &zeta_r, 8, hipMemcpyHostToDevice));
        else *(double*)p = zeta_r;
    } else if (dt == BLIS_FLOAT) {
        float v = (float)zeta_r;
        if (is_device_pointer(p)) HIP_CHECK(hipMemcpy(p, &v, 4, hipMemcpyHostToDevice));
        else *(float*)p = v;
    } else {
        abort_with(BLIS_EXPECTED_REAL_DATATYPE, "bli_setsc");
    }
}

BLIS_EXPORT void bli_getsc(const obj_t* chi, double* zeta_r, double* zeta_i) {
    *zeta_r = obj_scalar(chi);
    if (zeta_i) *zeta_i = 0.0;
}

// ---- level-3 typed ----------------------------------------------------------

#define GEMM_TYPED(ch, T)                                                              \
BLIS_EXPORT void bli_##ch##gemm_ex(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k, \
        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, const T* b, inc_t rs_b,    \
        inc_t cs_b, const T* beta, T* c, inc_t rs_c, inc_t cs_c,                       \
        const cntx_t* cntx, const rntm_t* rntm) {                                      \
    (void)cntx; (void)rntm;                                                            \
    gemm_core<T>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,         \
                 beta, c, rs_c, cs_c);                                                 \
}                                                                                      \
BLIS_EXPORT void bli_##ch##gemm(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k, \
        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, const T* b, inc_t rs_b,    \
        inc_t cs_b, const T* beta, T* c, inc_t rs_c, inc_t cs_c) {                     \
    gemm_core<T>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,         \
                 beta, c, rs_c, cs_c);                                                 \
}
GEMM_TYPED(s, float)
GEMM_TYPED(d, double)

#define TRSM_TYPED(ch, T)                                                              \
BLIS_EXPORT void bli_##ch##trsm_ex(side_t side, uplo_t uploa, trans_t transa, diag_t diaga, \
        dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,          \
        T* b, inc_t rs_b, inc_t cs_b, const cntx_t* cntx, const rntm_t* rntm) {        \
    (void)cntx; (void)rntm;                                                            \
    trsm_core<T>(side, uploa, transa, diaga, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b); \
}                                                                                      \
BLIS_EXPORT void bli_##ch##trsm(side_t side, uplo_t uploa, trans_t transa, diag_t diaga, \
        dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,          \
        T* b, inc_t rs_b, inc_t cs_b) {                                                \
    trsm_core<T>(side, uploa, transa, diaga, m, n, alpha, a, rs_a, cs_a, b, rs_b, cs_b); \
}
TRSM_TYPED(s, float)
TRSM_TYPED(d, double)

#define GEMM_BATCH_TYPED(ch, T)                                                        \
BLIS_EXPORT void bli_##ch##gemm_batch(const trans_t* transa, const trans_t* transb,    \
        const dim_t* m, const dim_t* n, const dim_t* k, const T* alpha,                \
        const T* const* a, const inc_t* rs_a, const inc_t* cs_a,                       \
        const T* const* b, const inc_t* rs_b, const inc_t* cs_b, const T* beta,        \
        T* const* c, const inc_t* rs_c, const inc_t* cs_c,                             \
        dim_t group_count, const dim_t* group_size) {                                  \
    gemm_batch_core<T>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,   \
                       beta, c, rs_c, cs_c, group_count, group_size);                  \
}
GEMM_BATCH_TYPED(s, float)
GEMM_BATCH_TYPED(d, double)

// ---- level-3 object API -----------------------------------------------------

BLIS_EXPORT void bli_gemm_ex(const obj_t* alpha, const obj_t* a, const obj_t* b,
                             const obj_t* beta, const obj_t* c,
                             const cntx_t* cntx, const rntm_t* rntm) {
    (void)cntx; (void)rntm;
    init_once();
    require(a != nullptr && b != nullptr && c != nullptr, BLIS_NULL_POINTER, "bli_gemm");
    {
        // Complex operands, mixed datatypes, or a computation precision other
        // than C's storage precision take the mixed-datatype path.
        const num_t dc = bli_obj_dt(c);
        const unsigned cp = (c->info & BLIS_COMP_PREC_BIT) >> BLIS_COMP_PREC_SHIFT;
        const bool plain = (dc == BLIS_FLOAT || dc == BLIS_DOUBLE) &&
                           bli_obj_dt(a) == dc && bli_obj_dt(b) == dc &&
                           cp == ((unsigned)dc & BLIS_PRECISION_BIT);
        if (!plain) { gemm_obj_md(alpha, a, b, beta, c); return; }
    }
    check_matrix_obj(a, "bli_gemm: a");
    check_matrix_obj(b, "bli_gemm: b");
    check_matrix_obj(c, "bli_gemm: c");
    const dim_t m = c->dim[0], n = c->dim[1], k = bli_obj_width_after_trans(a);
    require(bli_obj_length_after_trans(a) == m && bli_obj_width_after_trans(b) == n &&
            bli_obj_length_after_trans(b) == k, BLIS_NONCONFORMAL_DIMENSIONS, "bli_gemm");
    require(bli_obj_struc(c) == BLIS_GENERAL, BLIS_EXPECTED_GENERAL_OBJECT, "bli_gemm: c");
    const double al = obj_scalar(alpha), be = obj_scalar(beta);
    const trans_t ta = bli_obj_onlytrans_status(a), tb = bli_obj_onlytrans_status(b);
    if (bli_obj_dt(c) == BLIS_DOUBLE) {
        gemm_core<double>(ta, tb, m, n, k, &al, (const double*)obj_origin(a), a->rs, a->cs,
                          (const double*)obj_origin(b), b->rs, b->cs, &be,
                          (double*)obj_origin(c), c->rs, c->cs);
    } else {
        const float fa = (float)al, fb = (float)be;
        gemm_core<float>(ta, tb, m, n, k, &fa, (const float*)obj_origin(a), a->rs, a->cs,
                         (const float*)obj_origin(b), b->rs, b->cs, &fb,
                         (float*)obj_origin(c), c->rs, c->cs);
    }
}

BLIS_EXPORT void bli_gemm(const obj_t* alpha, const obj_t* a, const obj_t* b,
                          const obj_t* beta, const obj_t* c) {
    bli_gemm_ex(alpha, a, b, beta, c, nullptr, nullptr);
}

BLIS_EXPORT void bli_trsm_ex(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                             const cntx_t* cntx, const rntm_t* rntm) {
    (void)cntx; (void)rntm;
    init_once();
    check_matrix_obj(a, "bli_trsm: a");
    check_matrix_obj(b, "bli_trsm: b");
    require(bli_obj_dt(a) == bli_obj_dt(b), BLIS_INCONSISTENT_DATATYPES, "bli_trsm");
    require(bli_obj_struc(a) == BLIS_TRIANGULAR, BLIS_EXPECTED_TRIANGULAR_OBJECT, "bli_trsm: a");
    require(a->dim[0] == a->dim[1], BLIS_EXPECTED_SQUARE_OBJECT, "bli_trsm: a");
    const dim_t m = b->dim[0], n = b->dim[1];
    require(a->dim[0] == (side == BLIS_LEFT ? m : n), BLIS_NONCONFORMAL_DIMENSIONS, "bli_trsm");
    const uplo_t uplo = bli_obj_uplo(a);
    require(uplo == BLIS_LOWER || uplo == BLIS_UPPER, BLIS_EXPECTED_UPPER_OR_LOWER_OBJECT,
            "bli_trsm: a");
    const double al = obj_scalar(alpha);
    const trans_t ta = bli_obj_onlytrans_status(a);
    const diag_t dg = bli_obj_diag(a);
    if (bli_obj_dt(b) == BLIS_DOUBLE) {
        trsm_core<double>(side, uplo, ta, dg, m, n, &al, (const double*)obj_origin(a),
                          a->rs, a->cs, (double*)obj_origin(b), b->rs, b->cs);
    } else {
        const float fa = (float)al;
        trsm_core<float>(side, uplo, ta, dg, m, n, &fa, (const float*)obj_origin(a),
                         a->rs, a->cs, (float*)obj_origin(b), b->rs, b->cs);
    }
}

BLIS_EXPORT void bli_trsm(side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b) {
    bli_trsm_ex(side, alpha, a, b, nullptr, nullptr);
}

// ---- level-1/2 typed ---------------------------------------------------------

#define L12_TYPED(ch, T)                                                               \
BLIS_EXPORT void bli_##ch##axpyv_ex(conj_t conjx, dim_t n, const T* alpha, const T* x, \
        inc_t incx, T* y, inc_t incy, const cntx_t* cntx, const rntm_t* rntm) {        \
    (void)conjx; (void)cntx; (void)rntm;                                               \
    axpyv_core<T>(n, alpha, x, incx, y, incy);                                         \
}                                                                                      \
BLIS_EXPORT void bli_##ch##axpyv(conj_t conjx, dim_t n, const T* alpha, const T* x,    \
        inc_t incx, T* y, inc_t incy) {                                                \
    (void)conjx;                                                                       \
    axpyv_core<T>(n, alpha, x, incx, y, incy);                                         \
}                                                                                      \
BLIS_EXPORT void bli_##ch##dotv_ex(conj_t conjx, conj_t conjy, dim_t n, const T* x,    \
        inc_t incx, const T* y, inc_t incy, T* rho, const cntx_t* cntx,                \
        const rntm_t* rntm) {                                                          \
    (void)conjx; (void)conjy; (void)cntx; (void)rntm;                                  \
    dotv_core<T>(n, x, incx, y, incy, rho);                                            \
}                                                                                      \
BLIS_EXPORT void bli_##ch##dotv(conj_t conjx, conj_t conjy, dim_t n, const T* x,       \
        inc_t incx, const T* y, inc_t incy, T* rho) {                                  \
    (void)conjx; (void)conjy;                                                          \
    dotv_core<T>(n, x, incx, y, incy, rho);                                            \
}                                                                                      \
BLIS_EXPORT void bli_##ch##gemv_ex(trans_t transa, conj_t conjx, dim_t m, dim_t n,     \
        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, const T* x, inc_t incx,    \
        const T* beta, T* y, inc_t incy, const cntx_t* cntx, const rntm_t* rntm) {     \
    (void)conjx; (void)cntx; (void)rntm;                                               \
    gemv_core<T>(transa, m, n, alpha, a, rs_a, cs_a, x, incx, beta, y, incy);          \
}                                                                                      \
BLIS_EXPORT void bli_##ch##gemv(trans_t transa, conj_t conjx, dim_t m, dim_t n,        \
        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, const T* x, inc_t incx,    \
        const T* beta, T* y, inc_t incy) {                                             \
    (void)conjx;                                                                       \
    gemv_core<T>(transa, m, n, alpha, a, rs_a, cs_a, x, incx, beta, y, incy);          \
}
L12_TYPED(s, float)
L12_TYPED(d, double)

// ---- level-1/2 object API ------------------------------------------------------

BLIS_EXPORT void bli_axpyv(const obj_t* alpha, const obj_t* x, const obj_t* y) {
    init_once();
    require(x && y, BLIS_NULL_POINTER, "bli_axpyv");
    require(obj_is_real_fp(x) && bli_obj_dt(x) == bli_obj_dt(y), BLIS_INCONSISTENT_DATATYPES,
            "bli_axpyv");
    require(obj_vec_dim(x) == obj_vec_dim(y), BLIS_UNEQUAL_VECTOR_LENGTHS, "bli_axpyv");
    const double al = obj_scalar(alpha);
    if (bli_obj_dt(y) == BLIS_DOUBLE)
        axpyv_core<double>(obj_vec_dim(x), &al, (const double*)obj_origin(x), obj_vec_inc(x),
                           (double*)obj_origin(y), obj_vec_inc(y));
    else {
        const float fa = (float)al;
        axpyv_core<float>(obj_vec_dim(x), &fa, (const float*)obj_origin(x), obj_vec_inc(x),
                          (float*)obj_origin(y), obj_vec_inc(y));
    }
}

BLIS_EXPORT void bli_dotv(const obj_t* x, const obj_t* y, const obj_t* rho) {
    init_once();
    require(x && y && rho, BLIS_NULL_POINTER, "bli_dotv");
    require(obj_is_real_fp(x) && bli_obj_dt(x) == bli_obj_dt(y), BLIS_INCONSISTENT_DATATYPES,
            "bli_dotv");
    require(obj_vec_dim(x) == obj_vec_dim(y), BLIS_UNEQUAL_VECTOR_LENGTHS, "bli_dotv");
    if (bli_obj_dt(x) == BLIS_DOUBLE) {
        double r = 0.0;
        dotv_core<double>(obj_vec_dim(x), (const double*)obj_origin(x), obj_vec_inc(x),
                          (const double*)obj_origin(y), obj_vec_inc(y), &r);
        bli_setsc(r, 0.0, const_cast<obj_t*>(rho));
    } else {
        float r = 0.0f;
        dotv_core<float>(obj_vec_dim(x), (const float*)obj_origin(x), obj_vec_inc(x),
                         (const float*)obj_origin(y), obj_vec_inc(y), &r);
        bli_setsc((double)r, 0.0, const_cast<obj_t*>(rho));
    }
}

BLIS_EXPORT void bli_gemv(const obj_t* alpha, const obj_t* a, const obj_t* x,
                          const obj_t* beta, const obj_t* y) {
    init_once();
    check_matrix_obj(a, "bli_gemv: a");
    require(x && y, BLIS_NULL_POINTER, "bli_gemv");
    require(bli_obj_dt(a) == bli_obj_dt(x) && bli_obj_dt(a) == bli_obj_dt(y),
            BLIS_INCONSISTENT_DATATYPES, "bli_gemv");
    require(bli_obj_length_after_trans(a) == obj_vec_dim(y) &&
            bli_obj_width_after_trans(a) == obj_vec_dim(x), BLIS_NONCONFORMAL_DIMENSIONS,
            "bli_gemv");
    const double al = obj_scalar(alpha), be = obj_scalar(beta);
    const trans_t ta = bli_obj_onlytrans_status(a);
    if (bli_obj_dt(a) == BLIS_DOUBLE)
        gemv_core<double>(ta, a->dim[0], a->dim[1], &al, (const double*)obj_origin(a), a->rs,
                          a->cs, (const double*)obj_origin(x), obj_vec_inc(x), &be,
                          (double*)obj_origin(y), obj_vec_inc(y));
    else {
        const float fa = (float)al, fb = (float)be;
        gemv_core<float>(ta, a->dim[0], a->dim[1], &fa, (const float*)obj_origin(a), a->rs,
                         a->cs, (const float*)obj_origin(x), obj_vec_inc(x), &fb,
                         (float*)obj_origin(y), obj_vec_inc(y));
    }
}

// ---- BLAS compatibility -------------------------------------------------------

BLIS_EXPORT __attribute__((weak)) int xerbla_(const char* srname, const f77_int* info,
                                              ftnlen srname_len) {
    int len = 0;
    while (len < (int)srname_len && len < 6 && srname[len] != '\0') ++len;
    std::printf("** On entry to %6.*s, parameter number %2i had an illegal value\n", len,
                srname, (int)*info);
    return 0;
}

BLIS_EXPORT f77_int lsame_(const char* ca, const char* cb, ftnlen, ftnlen) {
    return up(*ca) == up(*cb) ? 1 : 0;
}

BLIS_EXPORT void sgemm_(const f77_char* transa, const f77_char* transb, const f77_int* m,
                        const f77_int* n, const f77_int* k, const float* alpha, const float* a,
                        const f77_int* lda, const float* b, const f77_int* ldb,
                        const float* beta, float* c, const f77_int* ldc) {
    blas_gemm<float>("s", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
BLIS_EXPORT void dgemm_(const f77_char* transa, const f77_char* transb, const f77_int* m,
                        const f77_int* n, const f77_int* k, const double* alpha,
                        const double* a, const f77_int* lda, const double* b,
                        const f77_int* ldb, const double* beta, double* c, const f77_int* ldc) {
    blas_gemm<double>("d", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
BLIS_EXPORT void strsm_(const f77_char* side, const f77_char* uploa, const f77_char* transa,
                        const f77_char* diaga, const f77_int* m, const f77_int* n,
                        const float* alpha, const float* a, const f77_int* lda, float* b,
                        const f77_int* ldb) {
    blas_trsm<float>("s", side, uploa, transa, diaga, m, n, alpha, a, lda, b, ldb);
}
BLIS_EXPORT void dtrsm_(const f77_char* side, const f77_char* uploa, const f77_char* transa,
                        const f77_char* diaga, const f77_int* m, const f77_int* n,
                        const double* alpha, const double* a, const f77_int* lda, double* b,
                        const f77_int* ldb) {
    blas_trsm<double>("d", side, uploa, transa, diaga, m, n, alpha, a, lda, b, ldb);
}
BLIS_EXPORT void sgemv_(const f77_char* transa, const f77_int* m, const f77_int* n,
                        const float* alpha, const float* a, const f77_int* lda, const float* x,
                        const f77_int* incx, const float* beta, float* y, const f77_int* incy) {
    blas_gemv<float>("s", transa, m, n, alpha, a, lda, x, incx, beta, y, incy);
}
BLIS_EXPORT void dgemv_(const f77_char* transa, const f77_int* m, const f77_int* n,
                        const double* alpha, const double* a, const f77_int* lda,
                        const double* x, const f77_int* incx, const double* beta, double* y,
                        const f77_int* incy) {
    blas_gemv<double>("d", transa, m, n, alpha, a, lda, x, incx, beta, y, incy);
}
BLIS_EXPORT void saxpy_(const f77_int* n, const float* alpha, const float* x,
                        const f77_int* incx, float* y, const f77_int* incy) {
    blas_axpy<float>(n, alpha, x, incx, y, incy);
}
BLIS_EXPORT void daxpy_(const f77_int* n, const double* alpha, const double* x,
                        const f77_int* incx, double* y, const f77_int* incy) {
    blas_axpy<double>(n, alpha, x, incx, y, incy);
}
BLIS_EXPORT float sdot_(const f77_int* n, const float* x, const f77_int* incx, const float* y,
                        const f77_int* incy) {
    return blas_dot<float>(n, x, incx, y, incy);
}
BLIS_EXPORT double ddot_(const f77_int* n, const double* x, const f77_int* incx,
                         const double* y, const f77_int* incy) {
    return blas_dot<double>(n, x, incx, y, incy);
}
BLIS_EXPORT void sgemm_batch_(const f77_char* transa_array, const f77_char* transb_array,
                              const f77_int* m_array, const f77_int* n_array,
                              const f77_int* k_array, const float* alpha_array,
                              const float** a_array, const f77_int* lda_array,
                              const float** b_array, const f77_int* ldb_array,
                              const float* beta_array, float** c_array,
                              const f77_int* ldc_array, const f77_int* group_count,
                              const f77_int* group_size) {
    blas_gemm_batch<float>("s", transa_array, transb_array, m_array, n_array, k_array,
                           alpha_array, a_array, lda_array, b_array, ldb_array, beta_array,
                           c_array, ldc_array, group_count, group_size);
}
BLIS_EXPORT void dgemm_batch_(const f77_char* transa_array, const f77_char* transb_array,
                              const f77_int* m_array, const f77_int* n_array,
                              const f77_int* k_array, const double* alpha_array,
                              const double** a_array, const f77_int* lda_array,
                              const double** b_array, const f77_int* ldb_array,
                              const double* beta_array, double** c_array,
                              const f77_int* ldc_array, const f77_int* group_count,
                              const f77_int* group_size) {
    blas_gemm_batch<double>("d", transa_array, transb_array, m_array, n_array, k_array,
                            alpha_array, a_array, lda_array, b_array, ldb_array, beta_array,
                            c_array, ldc_array, group_count, group_size);
}
