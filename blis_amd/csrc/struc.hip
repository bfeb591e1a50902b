// struc.hip — structured level-3 operations on the GEMM tiles: gemmt (and
// through it syrk and syr2k), symm, trmm and trmm3.
//
// BLIS runs all of these through its GEMM control tree with structure-aware
// packing (frame/3/bli_l3_oapi_ex.c:151-690; frame/1m/packm/
// bli_packm_struc_cxk.c:38-306): packm densifies a symmetric operand by
// mirroring its stored triangle, zero-fills the unstored triangle of a
// triangular one (writing ones on a unit diagonal), and the macro-kernel
// skips micro-tiles of C outside gemmt's stored triangle or of a
// triangular product's zero region. The GPU form:
//   structured packm   -> struc_pack_kernel: one pass over the stored
//                         triangle into a dense column-major copy in
//                         workspace (64 x 64 LDS tiles, coalesced both ways);
//   masked C           -> gemm_struc_kernel with MASK: workgroups whose tile
//                         lies wholly outside C's stored triangle return at
//                         once, tiles on the diagonal store element-masked;
//   triangular skip    -> gemm_struc_kernel with a per-tile k range: a tile
//                         runs only over the k slabs where the triangular
//                         operand can be nonzero, which halves the flops.
// syrk = gemmt(A, A^T) and syr2k = gemmt(A, B^T) then gemmt(B, A^T) with
// beta = 1, as bli_syrk_ex / bli_syr2k_ex do (bli_l3_oapi_ex.c:260-330).
#include "gemm_plan.hpp"

#include <algorithm>

namespace blis_amd {

namespace {

// Per-tile k range of a product with one triangular operand. With A_eff
// (m x k) and B_eff (k x n):
//   ROW_END:   A_eff(i, p) may be nonzero only for p <= i + koff;
//   ROW_BEGIN: A_eff(i, p) may be nonzero only for p >= i + koff;
//   COL_BEGIN: B_eff(p, j) may be nonzero only for p >= j + koff;
//   COL_END:   B_eff(p, j) may be nonzero only for p <= j + koff.
enum KMode : int { K_FULL = 0, K_ROW_END = 1, K_ROW_BEGIN = 2, K_COL_BEGIN = 3, K_COL_END = 4 };

struct StrucSpec {
    int     cmask = 0;     // C's stored triangle: 0 all, 1 lower, 2 upper
    int64_t cdoff = 0;     // its diagonal offset
    int     kmode = K_FULL;
    int64_t koff = 0;
    int     reverse = 0;   // walk the tiles last to first
};

// The same problem after the operation transpose C^T = B^T A^T.
StrucSpec transposed(const StrucSpec& s) {
    StrucSpec t = s;
    if (s.cmask) { t.cmask = 3 - s.cmask; t.cdoff = -s.cdoff; }
    switch (s.kmode) {
    case K_ROW_END:   t.kmode = K_COL_END; break;
    case K_ROW_BEGIN: t.kmode = K_COL_BEGIN; break;
    case K_COL_BEGIN: t.kmode = K_ROW_BEGIN; break;
    case K_COL_END:   t.kmode = K_ROW_END; break;
    default: break;
    }
    return t;
}

} // namespace

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// One tile of a structured product. Workgroups are uniform in (tm, tn), so
// the early return is taken by the whole workgroup before any barrier.
template <typename T, class TL, bool A_K, bool B_K, bool MASK>
__global__ __launch_bounds__(TL::NT) void gemm_struc_kernel(TileProblem<T> p, GridInfo g,
                                                            StrucSpec s) {
    __shared__ T smem[2 * TL::STAGE];
    int tm, tn;
    const int bid = s.reverse ? (int)gridDim.x - 1 - (int)blockIdx.x : (int)blockIdx.x;
    tile_coords(bid, gridDim.x, g, tm, tn);
    const int64_t m0 = (int64_t)tm * TL::BM, n0 = (int64_t)tn * TL::BN;
    if constexpr (MASK) {
        // lower: stored rows of column j are i >= j - doff; upper: i <= j - doff.
        if (s.cmask == 1 && m0 + TL::BM - 1 < n0 - s.cdoff) return;
        if (s.cmask == 2 && m0 > n0 + TL::BN - 1 - s.cdoff) return;
    }
    int64_t kb = 0, ke = p.k;
    switch (s.kmode) {
    case K_ROW_END:   ke = m0 + TL::BM + s.koff; break;
    case K_ROW_BEGIN: kb = m0 + s.koff; break;
    case K_COL_BEGIN: kb = n0 + s.koff; break;
    case K_COL_END:   ke = n0 + TL::BN + s.koff; break;
    default: break;
    }
    // Start on a multiple of 4 elements so that 16-byte operand vectors stay
    // aligned; the k range may only grow, which adds zero terms.
    kb = kb < 0 ? 0 : (kb > p.k ? p.k : kb);
    kb &= ~(int64_t)3;
    ke = ke > p.k ? p.k : ke;
    if (ke < kb) ke = kb;
    p.a += A_K ? kb : kb * p.lda;
    p.b += B_K ? kb : kb * p.ldb;
    p.k = ke - kb;
    gemm_tile<T, TL, A_K, B_K, 1, false, false, false, MASK>(p, tm, tn, smem, nullptr,
                                                             CMask{s.cmask, s.cdoff});
}

// Dense column-major copy (ld = ldd) of an n x n structured matrix stored
// in one triangle of src (element (i, j) at src[i*rs + j*cs]); (i, j) is
// stored when j - i <= doff (lower) or j - i >= doff (upper).
//   SYM:  dst(i, j) = src(i, j) if stored, else src(j, i)  (symmetric);
//   TRI:  dst(i, j) = src(i, j) if stored, else 0; a unit diagonal
//         (j - i == doff) is written as 1 and not read  (triangular).
// Only stored elements are read. A workgroup handles 64 x 64 output tiles:
// it loads the tile and its mirror through LDS with lanes along src's unit
// stride, and writes dst down its columns.
template <typename T, bool SYM>
__global__ __launch_bounds__(256) void struc_pack_kernel(const T* __restrict__ src, int64_t n,
                                                         int64_t rs, int64_t cs, int lower,
                                                         int64_t doff, int unit,
                                                         T* __restrict__ dst, int64_t ldd) {
    __shared__ T t1[64][65];   // t1[r][c] = src(I0 + r, J0 + c)
    __shared__ T t2[64][65];   // t2[r][c] = src(J0 + r, I0 + c)
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const bool row_major = (rs != 1) && (cs == 1);
    const int64_t nt = (n + 63) / 64;
    auto stored = [&](int64_t i, int64_t j) {
        return lower ? (j - i <= doff) : (j - i >= doff);
    };
    for (int64_t t = blockIdx.x; t < nt * nt; t += gridDim.x) {
        const int64_t I0 = (t % nt) * 64, J0 = (t / nt) * 64;
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            // lanes along the unit stride of src
            const int r = row_major ? ty + 4 * q : tx;
            const int c = row_major ? tx : ty + 4 * q;
            const int64_t i1 = I0 + r, j1 = J0 + c;
            if (i1 < n && j1 < n && stored(i1, j1) && !(unit && j1 - i1 == doff))
                t1[r][c] = src[i1 * rs + j1 * cs];
            if (SYM) {
                const int64_t i2 = J0 + r, j2 = I0 + c;
                if (i2 < n && j2 < n && stored(i2, j2)) t2[r][c] = src[i2 * rs + j2 * cs];
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const int r = tx, c = ty + 4 * q;
            const int64_t i = I0 + r, j = J0 + c;
            if (i < n && j < n) {
                T v;
                if (stored(i, j)) v = (unit && j - i == doff) ? T(1) : t1[r][c];
                else v = SYM ? t2[c][r] : T(0);
                dst[i + j * ldd] = v;
            }
        }
        __syncthreads();
    }
}

// C := beta*C on the stored triangle only (gemmt with alpha == 0 or k == 0).
// beta == 0 stores zeros without reading C.
template <typename T>
__global__ void scalm_struc_kernel(T* c, int64_t m, int64_t n, int64_t rs, int64_t cs, T beta,
                                   int lower, int64_t doff) {
    const int64_t total = m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        if (lower ? (j - i > doff) : (j - i < doff)) continue;
        T* dst = c + i * rs + j * cs;
        *dst = (beta == T(0)) ? T(0) : beta * *dst;
    }
}

// ---------------------------------------------------------------------------
// Launch plumbing
// ---------------------------------------------------------------------------

namespace {

int grid_for(int64_t total) {
    int64_t g = (total + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

template <typename T, class TL>
void launch_struc_tl(const TileProblem<T>& p, bool a_k, bool b_k, const GridInfo& g,
                     const StrucSpec& s, hipStream_t st) {
    const dim3 grid(g.grid_m * g.grid_n), block(TL::NT);
#define BLIS_AMD_STRUC_LAUNCH(AK, BK)                                                        \
    if (s.cmask)                                                                             \
        hipLaunchKernelGGL((gemm_struc_kernel<T, TL, AK, BK, true>), grid, block, 0, st, p,  \
                           g, s);                                                            \
    else                                                                                     \
        hipLaunchKernelGGL((gemm_struc_kernel<T, TL, AK, BK, false>), grid, block, 0, st, p, \
                           g, s);
    if (!a_k && !b_k) { BLIS_AMD_STRUC_LAUNCH(false, false) }
    else if (!a_k && b_k) { BLIS_AMD_STRUC_LAUNCH(false, true) }
    else if (a_k && !b_k) { BLIS_AMD_STRUC_LAUNCH(true, false) }
    else { BLIS_AMD_STRUC_LAUNCH(true, true) }
#undef BLIS_AMD_STRUC_LAUNCH
    HIP_CHECK(hipGetLastError());
}

// The structured kernels are built for the tiles the size heuristic picks
// (gemm_default_tile): FP64 6, 10, 3, 4 and FP32 1, 12, 9, 10, 3, 4.
template <typename T>
void launch_struc(int tile, const TileProblem<T>& p, bool a_k, bool b_k, GridInfo g,
                  const StrucSpec& s, hipStream_t st) {
    const TileShape& ts = tile_shape(dt_of<T>(), tile);
    g.grid_m = (int)((p.m + ts.bm - 1) / ts.bm);
    g.grid_n = (int)((p.n + ts.bn - 1) / ts.bn);
#define BLIS_AMD_STRUC_CASE(idx, BM, BN, BK, WM, WN) \
    case idx: launch_struc_tl<T, TileCfg<T, BM, BN, BK, WM, WN>>(p, a_k, b_k, g, s, st); return;
    if constexpr (std::is_same<T, double>::value) {
        switch (tile) {
        BLIS_AMD_STRUC_CASE(6, 128, 128, 16, 64, 32)
        BLIS_AMD_STRUC_CASE(10, 64, 128, 16, 32, 32)
        BLIS_AMD_STRUC_CASE(3, 64, 64, 16, 32, 32)
        BLIS_AMD_STRUC_CASE(4, 32, 32, 16, 16, 16)
        default: break;
        }
    } else {
        switch (tile) {
        BLIS_AMD_STRUC_CASE(1, 128, 64, 32, 64, 32)
        BLIS_AMD_STRUC_CASE(12, 256, 128, 16, 64, 64)
        BLIS_AMD_STRUC_CASE(9, 128, 64, 32, 32, 32)
        BLIS_AMD_STRUC_CASE(10, 64, 128, 32, 32, 32)
        BLIS_AMD_STRUC_CASE(3, 64, 64, 32, 32, 32)
        BLIS_AMD_STRUC_CASE(4, 32, 32, 32, 16, 16)
        default: break;
        }
    }
#undef BLIS_AMD_STRUC_CASE
    abort_with(BLIS_INVALID_UKR_ID, "structured gemm: tile without a structured kernel");
}

void scalm_struc_run(Dt dt, int64_t m, int64_t n, double beta, View c, int lower, int64_t doff,
                     hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const int g = grid_for(m * n);
    if (dt == Dt::F64)
        hipLaunchKernelGGL(scalm_struc_kernel<double>, dim3(g), dim3(256), 0, st,
                           (double*)c.ptr, m, n, c.rs, c.cs, beta, lower, doff);
    else
        hipLaunchKernelGGL(scalm_struc_kernel<float>, dim3(g), dim3(256), 0, st,
                           (float*)c.ptr, m, n, c.rs, c.cs, (float)beta, lower, doff);
    HIP_CHECK(hipGetLastError());
}

// C := beta*C + alpha*A_eff*B_eff under structure s (stated for the problem
// as given; the operation transpose is applied here).
void gemm_struc_run(const GemmDesc& d, StrucSpec s, hipStream_t st) {
    if (d.m <= 0 || d.n <= 0) return;
    if (d.k <= 0 || d.alpha == 0.0) {
        if (s.cmask) scalm_struc_run(d.dt, d.m, d.n, d.beta, d.c, s.cmask == 1, s.cdoff, st);
        else scalm_run(d.dt, d.m, d.n, d.beta, d.c, st);
        return;
    }
    if (d.c.rs != 1 && d.c.cs == 1) s = transposed(s);   // layout() transposes the same way
    Plan pl = layout(d);
    Scratch sc;
    pack_operands(pl, d.dt, sc, st);
    // The work per tile varies across the grid (tiles outside C's triangle
    // do nothing; a triangular operand's k range grows along the rows or the
    // columns), so the tiles are dealt round-robin over the XCDs instead of
    // in contiguous ranges, which would give some XCDs most of the work.
    // Tiles are numbered by groups of 8 tile rows, columns in order within a
    // group; when the k range grows with the row or the column index the
    // grid runs last tile first, so the longest tiles start first.
    GridInfo g{};
    g.group_m = 8;
    g.xcd_remap = 0;
    s.reverse = (s.kmode == K_ROW_END || s.kmode == K_COL_END) ? 1 : 0;
    const int tile = gemm_default_tile(d.dt, pl.m, pl.n, pl.k);
    if (d.dt == Dt::F64)
        launch_struc<double>(tile, make_problem<double>(pl, d.alpha, d.beta), pl.a_k, pl.b_k, g,
                             s, st);
    else
        launch_struc<float>(tile, make_problem<float>(pl, d.alpha, d.beta), pl.a_k, pl.b_k, g,
                            s, st);
}

// Dense copy of an n x n structured matrix into workspace slot `slot`
// (column-major, ld rounded up to 4 elements so rows start 16-byte aligned).
View struc_pack(Dt dt, bool sym, View a, int64_t n, bool lower, int64_t doff, bool unit,
                int slot, hipStream_t st) {
    const int64_t ldd = (n + 3) / 4 * 4;
    void* w = workspace((size_t)ldd * n * dt_size(dt), slot);
    const int64_t tiles = ((n + 63) / 64) * ((n + 63) / 64);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 16384));
    if (dt == Dt::F64) {
        if (sym)
            hipLaunchKernelGGL((struc_pack_kernel<double, true>), dim3(g), dim3(256), 0, st,
                               (const double*)a.ptr, n, a.rs, a.cs, (int)lower, doff, 0,
                               (double*)w, ldd);
        else
            hipLaunchKernelGGL((struc_pack_kernel<double, false>), dim3(g), dim3(256), 0, st,
                               (const double*)a.ptr, n, a.rs, a.cs, (int)lower, doff, (int)unit,
                               (double*)w, ldd);
    } else {
        if (sym)
            hipLaunchKernelGGL((struc_pack_kernel<float, true>), dim3(g), dim3(256), 0, st,
                               (const float*)a.ptr, n, a.rs, a.cs, (int)lower, doff, 0,
                               (float*)w, ldd);
        else
            hipLaunchKernelGGL((struc_pack_kernel<float, false>), dim3(g), dim3(256), 0, st,
                               (const float*)a.ptr, n, a.rs, a.cs, (int)lower, doff, (int)unit,
                               (float*)w, ldd);
    }
    HIP_CHECK(hipGetLastError());
    return View{ w, 1, ldd };
}

} // namespace

void gemmt_run(const GemmDesc& d, uplo_t uploc, int64_t doff, hipStream_t st) {
    StrucSpec s;
    s.cmask = (uploc == BLIS_LOWER) ? 1 : 2;
    s.cdoff = doff;
    gemm_struc_run(d, s, st);
}

void symm_run(Dt dt, side_t side, uplo_t uploa, int64_t m, int64_t n, double alpha, View a,
              View b, double beta, View c, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    GemmDesc d;
    d.dt = dt; d.m = m; d.n = n;
    d.alpha = alpha; d.beta = beta;
    d.c = c;
    d.k = (side == BLIS_LEFT) ? m : n;
    if (d.k <= 0 || alpha == 0.0) { scalm_run(dt, m, n, beta, c, st); return; }
    const View as = struc_pack(dt, true, a, d.k, uploa == BLIS_LOWER, 0, false, 20, st);
    if (side == BLIS_LEFT) { d.a = as; d.b = b; }
    else { d.a = b; d.b = as; }
    gemm_run(d, nullptr, st);
}

void trmm3_run(Dt dt, side_t side, uplo_t uploa, trans_t transa, diag_t diaga, int64_t doff,
               int64_t m, int64_t n, double alpha, View a, View b, double beta, View c,
               hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    GemmDesc d;
    d.dt = dt; d.m = m; d.n = n;
    d.alpha = alpha; d.beta = beta;
    d.c = c;
    d.k = (side == BLIS_LEFT) ? m : n;
    if (d.k <= 0 || alpha == 0.0) { scalm_run(dt, m, n, beta, c, st); return; }
    const bool lower = uploa == BLIS_LOWER;
    View t = struc_pack(dt, false, a, d.k, lower, doff, diaga == BLIS_UNIT_DIAG, 20, st);
    // Structure of op(A): a transpose swaps the triangles and negates doff.
    bool olower = lower;
    int64_t odoff = doff;
    if ((unsigned)transa & BLIS_TRANS_BIT) {
        std::swap(t.rs, t.cs);
        olower = !lower;
        odoff = -doff;
    }
    StrucSpec s;
    if (side == BLIS_LEFT) {
        // A_eff = op(A): (i, p) stored when p - i <= odoff (lower) / >= odoff.
        d.a = t; d.b = b;
        s.kmode = olower ? K_ROW_END : K_ROW_BEGIN;
        s.koff = odoff;
    } else {
        // B_eff = op(A): (p, j) stored when j - p <= odoff (lower) / >= odoff.
        d.a = b; d.b = t;
        s.kmode = olower ? K_COL_BEGIN : K_COL_END;
        s.koff = -odoff;
    }
    gemm_struc_run(d, s, st);
}

void trmm_run(Dt dt, side_t side, uplo_t uploa, trans_t transa, diag_t diaga, int64_t doff,
              int64_t m, int64_t n, double alpha, View a, View b, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    if (alpha == 0.0) { scalm_run(dt, m, n, 0.0, b, st); return; }
    // B is both an input and the output: the product reads a column-major
    // copy of it (slot 21) and overwrites B.
    const int64_t ldw = (m + 3) / 4 * 4;
    void* w = workspace((size_t)ldw * n * dt_size(dt), 21);
    if (ldw == m) {
        pack_run(dt, b.ptr, m, n, b.rs, b.cs, w, st);
    } else {
        // pack_run writes ld = m; copy column by column into the padded copy.
        for (int64_t j = 0; j < n; ++j)
            pack_run(dt, (const char*)b.ptr + j * b.cs * (int64_t)dt_size(dt), m, 1, b.rs, 1,
                     (char*)w + j * ldw * (int64_t)dt_size(dt), st);
    }
    trmm3_run(dt, side, uploa, transa, diaga, doff, m, n, alpha, a, View{ w, 1, ldw }, 0.0, b,
              st);
}

} // namespace blis_amd
