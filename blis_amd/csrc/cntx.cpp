// cntx.cpp — the kernel/blocksize registry (BLIS's gks and cntx_t) for
// gfx950, and the configuration queries of bli_info.
//
// Reference: frame/base/bli_gks.c:58-609 (arch -> context, kernel lookup),
// frame/base/bli_cntx.c:109-454 and frame/base/bli_cntx.h:73-204 (blocksize
// def/max values, blocksize multiples, kernel row/column preferences, the
// small-problem thresholds), frame/base/bli_info.c (configuration strings and
// values), frame/3/bli_l3_ukr_tapi.c (bli_?gemm_ukernel).
//
// On the GPU, BLIS's blocking parameters are the tile table's shapes: the
// registered context reports, per datatype, MR x NR = the wave tile of the
// default large workgroup tile (the register block of MFMA accumulators one
// wave owns), MC x NC = that workgroup tile, KC = its k slab, KR = the MFMA
// k step. MT/NT/KT are the thresholds below which the library takes its
// small-problem path. The registered gemm microkernel takes BLIS's packed
// micro-panels (A: MR x k, column-major with ld MR; B: k x NR, row-major
// with ld NR) and runs them through the GEMM tiles.
#include "api_common.hpp"
#include "gemm_tiles.hpp"

using namespace blis_amd;
using namespace blis_amd::api;

#define BLIS_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

inline int dt_index(num_t dt) {
    switch (dt) {
    case BLIS_FLOAT: return 0;
    case BLIS_SCOMPLEX: return 1;
    case BLIS_DOUBLE: return 2;
    case BLIS_DCOMPLEX: return 3;
    default: return -1;
    }
}

// The registry's microkernels: one MR x NR block from packed micro-panels.
template <typename T>
void gemm_ukernel_t(void (*gemm)(trans_t, trans_t, dim_t, dim_t, dim_t, const T*, const T*,
                                 inc_t, inc_t, const T*, inc_t, inc_t, const T*, T*, inc_t,
                                 inc_t),
                    dim_t mr, dim_t nr, dim_t m, dim_t n, dim_t k, const T* alpha, const T* a,
                    const T* b, const T* beta, T* c, inc_t rs_c, inc_t cs_c) {
    require(m <= mr && n <= nr, BLIS_NONCONFORMAL_DIMENSIONS, "gemm ukernel: m > MR or n > NR");
    gemm(BLIS_NO_TRANSPOSE, BLIS_NO_TRANSPOSE, m, n, k, alpha, a, 1, mr, b, nr, 1, beta, c,
         rs_c, cs_c);
}

struct Registry {
    cntx_t nat;
    Registry() {
        std::memset(&nat, 0, sizeof(nat));
        nat.method = BLIS_NAT;
        // Per datatype: the default large tile of each real type (FP32 tile
        // 12, FP64 tile 6). The complex types run on the real tiles in 1m
        // form (2m x n x 2k real products), so their register blocks cover
        // half the rows.
        const TileShape& ts = tile_shape(Dt::F32, 12);
        const TileShape& td = tile_shape(Dt::F64, 6);
        auto set = [&](bszid_t id, dim_t s, dim_t c, dim_t d, dim_t z) {
            nat.blkszs[id].v[0] = s; nat.blkszs[id].v[1] = c;
            nat.blkszs[id].v[2] = d; nat.blkszs[id].v[3] = z;
            for (int i = 0; i < 4; ++i) nat.blkszs[id].e[i] = nat.blkszs[id].v[i];
        };
        set(BLIS_KR, 4, 4, 4, 4);
        set(BLIS_MR, ts.wm, ts.wm / 2, td.wm, td.wm / 2);
        set(BLIS_NR, ts.wn, ts.wn, td.wn, td.wn);
        set(BLIS_MC, ts.bm, ts.bm / 2, td.bm, td.bm / 2);
        set(BLIS_KC, ts.bk, ts.bk / 2, td.bk, td.bk / 2);
        set(BLIS_NC, ts.bn, ts.bn, td.bn, td.bn);
        set(BLIS_BBM, 1, 1, 1, 1);
        set(BLIS_BBN, 1, 1, 1, 1);
        // Level-2 and level-1f: the gemv kernels' row block and x staging.
        set(BLIS_M2, 1024, 1024, 512, 512);
        set(BLIS_N2, 256, 256, 256, 256);
        set(BLIS_AF, 4, 4, 2, 2);
        set(BLIS_DF, 4, 4, 2, 2);
        set(BLIS_XF, 4, 4, 2, 2);
        // Small-problem thresholds (the single-call small path).
        set(BLIS_MT, SMALL_GEMM_MAX_DIM + 1, SMALL_GEMM_MAX_DIM + 1, SMALL_GEMM_MAX_DIM + 1,
            SMALL_GEMM_MAX_DIM + 1);
        set(BLIS_NT, SMALL_GEMM_MAX_DIM + 1, SMALL_GEMM_MAX_DIM + 1, SMALL_GEMM_MAX_DIM + 1,
            SMALL_GEMM_MAX_DIM + 1);
        set(BLIS_KT, SMALL_GEMM_MAX_DIM + 1, SMALL_GEMM_MAX_DIM + 1, SMALL_GEMM_MAX_DIM + 1,
            SMALL_GEMM_MAX_DIM + 1);
        for (int i = 0; i < BLIS_NUM_BLKSZS; ++i) nat.bmults[i] = (bszid_t)i;
        nat.bmults[BLIS_MC] = BLIS_MR;
        nat.bmults[BLIS_NC] = BLIS_NR;
        nat.bmults[BLIS_KC] = BLIS_KR;
        // Kernels: the gemm microkernel for every datatype.
        void_fp* g = nat.ukr2s[bli_ker_idx(BLIS_GEMM_UKR)];
        g[0] = (void_fp)bli_sgemm_ukernel;
        g[1] = (void_fp)bli_cgemm_ukernel;
        g[2] = (void_fp)bli_dgemm_ukernel;
        g[3] = (void_fp)bli_zgemm_ukernel;
        // Column preference: every tile writes column-stored C (row-stored C
        // is handled by the operation transpose).
        for (int p = 0; p < BLIS_NUM_UKR_PREFS; ++p)
            for (int i = 0; i < 4; ++i) nat.ukr_prefs[p].v[i] = false;
    }
};

Registry& registry() {
    static Registry r;
    return r;
}

} // namespace

// ---- microkernels ------------------------------------------------------------

BLIS_EXPORT void bli_sgemm_ukernel(dim_t m, dim_t n, dim_t k, const float* alpha, const float* a,
                                   const float* b, const float* beta, float* c, inc_t rs_c,
                                   inc_t cs_c, auxinfo_t* data, const cntx_t* cntx) {
    (void)data;
    if (cntx == nullptr) cntx = &registry().nat;
    gemm_ukernel_t<float>(bli_sgemm, cntx->blkszs[BLIS_MR].v[0], cntx->blkszs[BLIS_NR].v[0], m,
                          n, k, alpha, a, b, beta, c, rs_c, cs_c);
}

BLIS_EXPORT void bli_dgemm_ukernel(dim_t m, dim_t n, dim_t k, const double* alpha,
                                   const double* a, const double* b, const double* beta,
                                   double* c, inc_t rs_c, inc_t cs_c, auxinfo_t* data,
                                   const cntx_t* cntx) {
    (void)data;
    if (cntx == nullptr) cntx = &registry().nat;
    gemm_ukernel_t<double>(bli_dgemm, cntx->blkszs[BLIS_MR].v[2], cntx->blkszs[BLIS_NR].v[2],
                           m, n, k, alpha, a, b, beta, c, rs_c, cs_c);
}

BLIS_EXPORT void bli_cgemm_ukernel(dim_t m, dim_t n, dim_t k, const scomplex* alpha,
                                   const scomplex* a, const scomplex* b, const scomplex* beta,
                                   scomplex* c, inc_t rs_c, inc_t cs_c, auxinfo_t* data,
                                   const cntx_t* cntx) {
    (void)data;
    if (cntx == nullptr) cntx = &registry().nat;
    gemm_ukernel_t<scomplex>(bli_cgemm, cntx->blkszs[BLIS_MR].v[1], cntx->blkszs[BLIS_NR].v[1],
                             m, n, k, alpha, a, b, beta, c, rs_c, cs_c);
}

BLIS_EXPORT void bli_zgemm_ukernel(dim_t m, dim_t n, dim_t k, const dcomplex* alpha,
                                   const dcomplex* a, const dcomplex* b, const dcomplex* beta,
                                   dcomplex* c, inc_t rs_c, inc_t cs_c, auxinfo_t* data,
                                   const cntx_t* cntx) {
    (void)data;
    if (cntx == nullptr) cntx = &registry().nat;
    gemm_ukernel_t<dcomplex>(bli_zgemm, cntx->blkszs[BLIS_MR].v[3], cntx->blkszs[BLIS_NR].v[3],
                             m, n, k, alpha, a, b, beta, c, rs_c, cs_c);
}

// ---- gks ----------------------------------------------------------------------

BLIS_EXPORT const cntx_t* bli_gks_query_cntx(void) { return &registry().nat; }
BLIS_EXPORT const cntx_t* bli_gks_query_nat_cntx(void) { return &registry().nat; }

// Complex operations use the 1m method on the real tiles; the context for
// it is the native one (its complex entries already describe 1m).
BLIS_EXPORT const cntx_t* bli_gks_query_ind_cntx(ind_t ind) {
    require(ind == BLIS_NAT || ind == BLIS_1M, BLIS_NOT_YET_IMPLEMENTED, "bli_gks_query_ind_cntx");
    return &registry().nat;
}

BLIS_EXPORT kimpl_t bli_gks_l3_ukr_impl_type(ukr_t ukr, ind_t method, num_t dt) {
    if (dt_index(dt) < 0) return BLIS_NOTAPPLIC_UKERNEL;
    if (ukr == BLIS_GEMM_UKR)
        return (method == BLIS_1M || bli_dt_is_complex(dt)) ? BLIS_VIRTUAL_UKERNEL
                                                             : BLIS_OPTIMIZED_UKERNEL;
    if (ukr == BLIS_GEMMTRSM_L_UKR || ukr == BLIS_GEMMTRSM_U_UKR || ukr == BLIS_TRSM_L_UKR ||
        ukr == BLIS_TRSM_U_UKR)
        return bli_dt_is_complex(dt) ? BLIS_NOTAPPLIC_UKERNEL : BLIS_OPTIMIZED_UKERNEL;
    return BLIS_NOTAPPLIC_UKERNEL;
}

BLIS_EXPORT const char* bli_gks_l3_ukr_impl_string(ukr_t ukr, ind_t method, num_t dt) {
    switch (bli_gks_l3_ukr_impl_type(ukr, method, dt)) {
    case BLIS_REFERENCE_UKERNEL: return "refrnce";
    case BLIS_VIRTUAL_UKERNEL: return "virtual";
    case BLIS_OPTIMIZED_UKERNEL: return "optimzd";
    default: return "notappl";
    }
}

BLIS_EXPORT arch_t bli_arch_query_id(void) { return BLIS_ARCH_GFX950; }
BLIS_EXPORT const char* bli_arch_string(arch_t id) {
    return id == BLIS_ARCH_GFX950 ? "gfx950" : "unknown";
}

// ---- cntx queries ---------------------------------------------------------------

BLIS_EXPORT const blksz_t* bli_cntx_get_blksz(bszid_t bs_id, const cntx_t* cntx) {
    require(cntx != nullptr, BLIS_NULL_POINTER, "bli_cntx_get_blksz");
    require((int)bs_id >= 0 && (int)bs_id < BLIS_NUM_BLKSZS, BLIS_OUT_OF_BOUNDS,
            "bli_cntx_get_blksz");
    return &cntx->blkszs[bs_id];
}

BLIS_EXPORT dim_t bli_cntx_get_blksz_def_dt(num_t dt, bszid_t bs_id, const cntx_t* cntx) {
    const int i = dt_index(dt);
    require(i >= 0, BLIS_INVALID_DATATYPE, "bli_cntx_get_blksz_def_dt");
    return bli_cntx_get_blksz(bs_id, cntx)->v[i];
}

BLIS_EXPORT dim_t bli_cntx_get_blksz_max_dt(num_t dt, bszid_t bs_id, const cntx_t* cntx) {
    const int i = dt_index(dt);
    require(i >= 0, BLIS_INVALID_DATATYPE, "bli_cntx_get_blksz_max_dt");
    return bli_cntx_get_blksz(bs_id, cntx)->e[i];
}

BLIS_EXPORT dim_t bli_cntx_get_bmult_dt(num_t dt, bszid_t bs_id, const cntx_t* cntx) {
    require(cntx != nullptr, BLIS_NULL_POINTER, "bli_cntx_get_bmult_dt");
    require((int)bs_id >= 0 && (int)bs_id < BLIS_NUM_BLKSZS, BLIS_OUT_OF_BOUNDS,
            "bli_cntx_get_bmult_dt");
    return bli_cntx_get_blksz_def_dt(dt, cntx->bmults[bs_id], cntx);
}

BLIS_EXPORT void_fp bli_cntx_get_ukr_dt(num_t dt, ukr_t ukr_id, const cntx_t* cntx) {
    require(cntx != nullptr, BLIS_NULL_POINTER, "bli_cntx_get_ukr_dt");
    const int i = dt_index(dt);
    require(i >= 0, BLIS_INVALID_DATATYPE, "bli_cntx_get_ukr_dt");
    const kerid_t idx = bli_ker_idx(ukr_id);
    if (bli_ker_ntype(ukr_id) == 2)
        return idx < BLIS_NUM_UKR2S ? cntx->ukr2s[idx][i] : nullptr;
    return idx < BLIS_NUM_UKRS ? cntx->ukrs[idx][i] : nullptr;
}

BLIS_EXPORT bool bli_cntx_get_ukr_prefs_dt(num_t dt, ukr_pref_t pref_id, const cntx_t* cntx) {
    require(cntx != nullptr, BLIS_NULL_POINTER, "bli_cntx_get_ukr_prefs_dt");
    const int i = dt_index(dt);
    require(i >= 0 && (int)pref_id >= 0 && (int)pref_id < BLIS_NUM_UKR_PREFS, BLIS_OUT_OF_BOUNDS,
            "bli_cntx_get_ukr_prefs_dt");
    return cntx->ukr_prefs[pref_id].v[i];
}

BLIS_EXPORT bool bli_cntx_ukr_prefers_rows_dt(num_t dt, ukr_t ukr_id, const cntx_t* cntx) {
    ukr_pref_t p;
    switch (ukr_id) {
    case BLIS_GEMM_UKR: p = BLIS_GEMM_UKR_ROW_PREF; break;
    case BLIS_GEMMTRSM_L_UKR: p = BLIS_GEMMTRSM_L_UKR_ROW_PREF; break;
    case BLIS_GEMMTRSM_U_UKR: p = BLIS_GEMMTRSM_U_UKR_ROW_PREF; break;
    case BLIS_TRSM_L_UKR: p = BLIS_TRSM_L_UKR_ROW_PREF; break;
    case BLIS_TRSM_U_UKR: p = BLIS_TRSM_U_UKR_ROW_PREF; break;
    default: return false;
    }
    return bli_cntx_get_ukr_prefs_dt(dt, p, cntx);
}

BLIS_EXPORT bool bli_cntx_ukr_prefers_cols_dt(num_t dt, ukr_t ukr_id, const cntx_t* cntx) {
    return !bli_cntx_ukr_prefers_rows_dt(dt, ukr_id, cntx);
}

BLIS_EXPORT bool bli_cntx_l3_sup_thresh_is_met(num_t dt, dim_t m, dim_t n, dim_t k,
                                               const cntx_t* cntx) {
    if (m < bli_cntx_get_blksz_def_dt(dt, BLIS_MT, cntx)) return true;
    if (n < bli_cntx_get_blksz_def_dt(dt, BLIS_NT, cntx)) return true;
    if (k < bli_cntx_get_blksz_def_dt(dt, BLIS_KT, cntx)) return true;
    return false;
}

// ---- bli_info ---------------------------------------------------------------------

BLIS_EXPORT gint_t bli_info_get_int_type_size(void) { return BLIS_INT_TYPE_SIZE; }
BLIS_EXPORT const char* bli_info_get_int_type_size_str(void) { return "64"; }
BLIS_EXPORT gint_t bli_info_get_num_fp_types(void) { return BLIS_NUM_FP_TYPES; }
BLIS_EXPORT gint_t bli_info_get_max_type_size(void) { return (gint_t)sizeof(dcomplex); }
BLIS_EXPORT gint_t bli_info_get_page_size(void) { return 4096; }
// A wavefront is 64 lanes; an MFMA register block is 16 x 16 elements.
BLIS_EXPORT gint_t bli_info_get_simd_size(void) { return 64 * 8; }
BLIS_EXPORT gint_t bli_info_get_simd_num_registers(void) { return 512; }
BLIS_EXPORT gint_t bli_info_get_simd_align_size(void) { return 16; }
BLIS_EXPORT gint_t bli_info_get_stack_buf_max_size(void) { return 0; }
BLIS_EXPORT gint_t bli_info_get_stack_buf_align_size(void) { return 256; }
BLIS_EXPORT gint_t bli_info_get_heap_addr_align_size(void) { return 256; }
BLIS_EXPORT gint_t bli_info_get_enable_blas(void) { return 1; }
BLIS_EXPORT gint_t bli_info_get_enable_cblas(void) { return 1; }
BLIS_EXPORT gint_t bli_info_get_blas_int_type_size(void) { return BLIS_BLAS_INT_TYPE_SIZE; }
BLIS_EXPORT gint_t bli_info_get_enable_threading(void) { return 0; }
BLIS_EXPORT gint_t bli_info_get_enable_openmp(void) { return 0; }
BLIS_EXPORT gint_t bli_info_get_enable_pthreads(void) { return 0; }
BLIS_EXPORT gint_t bli_info_get_enable_sandbox(void) { return 0; }

BLIS_EXPORT const char* bli_info_get_gemm_ukr_impl_string(ind_t method, num_t dt) {
    return bli_gks_l3_ukr_impl_string(BLIS_GEMM_UKR, method, dt);
}
BLIS_EXPORT const char* bli_info_get_gemmtrsm_l_ukr_impl_string(ind_t method, num_t dt) {
    return bli_gks_l3_ukr_impl_string(BLIS_GEMMTRSM_L_UKR, method, dt);
}
BLIS_EXPORT const char* bli_info_get_gemmtrsm_u_ukr_impl_string(ind_t method, num_t dt) {
    return bli_gks_l3_ukr_impl_string(BLIS_GEMMTRSM_U_UKR, method, dt);
}
BLIS_EXPORT const char* bli_info_get_trsm_l_ukr_impl_string(ind_t method, num_t dt) {
    return bli_gks_l3_ukr_impl_string(BLIS_TRSM_L_UKR, method, dt);
}
BLIS_EXPORT const char* bli_info_get_trsm_u_ukr_impl_string(ind_t method, num_t dt) {
    return bli_gks_l3_ukr_impl_string(BLIS_TRSM_U_UKR, method, dt);
}

// Level-3 implementation strings (frame/base/bli_info.c): complex GEMM and
// the operations built on it run the 1m method; TRSM is native real only.
namespace {
const char* l3_impl(num_t dt, bool complex_supported) {
    if (dt_index(dt) < 0) return "notappl";
    if (!bli_dt_is_complex(dt)) return "native";
    return complex_supported ? "1m" : "notappl";
}
} // namespace
BLIS_EXPORT const char* bli_info_get_gemm_impl_string(num_t dt) { return l3_impl(dt, true); }
BLIS_EXPORT const char* bli_info_get_gemmt_impl_string(num_t dt) { return l3_impl(dt, false); }
BLIS_EXPORT const char* bli_info_get_symm_impl_string(num_t dt) { return l3_impl(dt, false); }
BLIS_EXPORT const char* bli_info_get_syrk_impl_string(num_t dt) { return l3_impl(dt, false); }
BLIS_EXPORT const char* bli_info_get_syr2k_impl_string(num_t dt) { return l3_impl(dt, false); }
BLIS_EXPORT const char* bli_info_get_trmm_impl_string(num_t dt) { return l3_impl(dt, false); }
BLIS_EXPORT const char* bli_info_get_trmm3_impl_string(num_t dt) { return l3_impl(dt, false); }
BLIS_EXPORT const char* bli_info_get_trsm_impl_string(num_t dt) { return l3_impl(dt, false); }
