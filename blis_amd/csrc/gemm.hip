// gemm.hip — gfx950 GEMM launch planner and kernels.
//
// Host side of the BLIS five-loop GEMM (frame/3/gemm/bli_gemm_cntl.c:75-565,
// sandbox/gemmlike/bls_gemm_bp_var1.c:37-318): it normalises operand
// storage, picks a tile configuration (the analogue of BLIS's MC/NC/KC/MR/NR
// blocksizes, config/zen3/bli_cntx_init_zen3.c:204-220) and launches one
// workgroup per output tile. The device side lives in gemm_tile.hpp.
//
// Storage normalisation follows bli_gemm_cntl_init's operation transpose
// (bli_gemm_cntl.c:104-161): when C is row-stored the problem is recast as
// C^T := B^T A^T so that the kernel always writes column-stored C. Operands
// with neither unit row nor unit column stride are first packed into a
// contiguous column-major copy, as packm does for general-stride inputs
// (ref_kernels/1m/bli_packm_cxk_ref.c:78-147).
#include "gemm_plan.hpp"

#include <algorithm>

namespace blis_amd {

using namespace dev;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

template <typename T, class TL, bool A_K, bool B_K>
__global__ __launch_bounds__(TL::NT) void gemm_kernel(TileProblem<T> p, GridInfo g) {
    __shared__ T smem[2 * TL::STAGE];
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, g, tm, tn);
    gemm_tile<T, TL, A_K, B_K>(p, tm, tn, smem);
}

// One k partition of a split-k product per blockIdx.y: partition z covers
// k in [z*kchunk, min(k, (z+1)*kchunk)) and stores alpha * A_z * B_z, without
// reading anything there first, into its own m x n column-major slice of
// `parts` (ld = m). splitk_reduce_kernel then forms C.
template <typename T, class TL, bool A_K, bool B_K>
__global__ __launch_bounds__(TL::NT) void gemm_splitk_kernel(TileProblem<T> p, GridInfo g,
                                                             int64_t kchunk, T* parts) {
    __shared__ T smem[2 * TL::STAGE];
    const int64_t z = blockIdx.y, k0 = z * kchunk;
    p.a += A_K ? k0 : k0 * p.lda;
    p.b += B_K ? k0 : k0 * p.ldb;
    p.k = min(kchunk, p.k - k0);
    p.c = parts + z * p.m * p.n;
    p.c_si = 1;
    p.c_sj = p.m;
    p.beta = T(0);
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, g, tm, tn);
    gemm_tile<T, TL, A_K, B_K>(p, tm, tn, smem);
}

// count products at fixed element strides, one per blockIdx.y:
// C_z := beta*C_z + alpha*A_z*B_z with A_z = A + z*sa, B_z = B + z*sb,
// C_z = C + z*sc.
template <typename T, class TL, bool A_K, bool B_K>
__global__ __launch_bounds__(TL::NT) void gemm_strided_kernel(TileProblem<T> p, GridInfo g,
                                                              int64_t sa, int64_t sb, int64_t sc) {
    __shared__ T smem[2 * TL::STAGE];
    const int64_t z = blockIdx.y;
    p.a += z * sa;
    p.b += z * sb;
    p.c += z * sc;
    int tm, tn;
    tile_coords(blockIdx.x, gridDim.x, g, tm, tn);
    gemm_tile<T, TL, A_K, B_K>(p, tm, tn, smem);
}

// C := beta*C + sum_z parts_z, partitions added in ascending z (a fixed
// order, so the result does not depend on scheduling). beta == 0 does not
// read C.
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const T* __restrict__ parts, int s,
                                                            int64_t m, int64_t n, T* c,
                                                            int64_t c_si, int64_t c_sj, T beta) {
    const int64_t total = m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        T acc = parts[t];
        for (int z = 1; z < s; ++z) acc += parts[z * total + t];
        const int64_t i = t % m, j = t / m;
        T* dst = c + i * c_si + j * c_sj;
        *dst = (beta == T(0)) ? acc : beta * *dst + acc;
    }
}

// C := beta*C for alpha == 0 or k == 0 (bli_l3_return_early_if_trivial,
// frame/3/bli_l3_util.c:41-65). beta == 0 stores zeros without reading C.
template <typename T>
__global__ void scalm_kernel(T* c, int64_t m, int64_t n, int64_t rs, int64_t cs, T beta) {
    const int64_t total = m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        T* dst = c + i * rs + j * cs;
        *dst = (beta == T(0)) ? T(0) : beta * *dst;
    }
}

// dst (column-major m x n, ld = m) := src with arbitrary strides.
template <typename T>
__global__ void pack_kernel(const T* src, int64_t m, int64_t n, int64_t rs, int64_t cs, T* dst) {
    const int64_t total = m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        dst[t] = src[i * rs + j * cs];
    }
}

// dst(r, c) = src(r, c) for an operand stored rows-contiguous along c,
// src(r, c) = src[c + r*lds], written column-wise: dst[r + c*ldd]. 64 x 64
// tiles through LDS, coalesced on both sides.
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const T* __restrict__ src, int64_t rows,
                                                        int64_t cols, int64_t lds,
                                                        T* __restrict__ dst, int64_t ldd) {
    __shared__ T tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t trows = (rows + 63) / 64, tcols = (cols + 63) / 64;
    for (int64_t t = blockIdx.x; t < trows * tcols; t += gridDim.x) {
        const int64_t r0 = (t % trows) * 64, c0 = (t / trows) * 64;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t r = r0 + ty + 4 * q, c = c0 + tx;
            if (r < rows && c < cols) tile[ty + 4 * q][tx] = src[c + r * lds];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t r = r0 + tx, c = c0 + ty + 4 * q;
            if (r < rows && c < cols) dst[r + c * ldd] = tile[tx][ty + 4 * q];
        }
        __syncthreads();
    }
}

static int grid_for(int64_t total) {
    int64_t g = (total + 255) / 256;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (int)g;
}

void scalm_run(Dt dt, int64_t m, int64_t n, double beta, View c, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const int g = grid_for(m * n);
    if (dt == Dt::F64)
        hipLaunchKernelGGL(scalm_kernel<double>, dim3(g), dim3(256), 0, st,
                           (double*)c.ptr, m, n, c.rs, c.cs, beta);
    else
        hipLaunchKernelGGL(scalm_kernel<float>, dim3(g), dim3(256), 0, st,
                           (float*)c.ptr, m, n, c.rs, c.cs, (float)beta);
    HIP_CHECK(hipGetLastError());
}

void pack_run(Dt dt, const void* src, int64_t m, int64_t n, int64_t rs, int64_t cs,
              void* dst, hipStream_t st) {
    const int g = grid_for(m * n);
    if (dt == Dt::F64)
        hipLaunchKernelGGL(pack_kernel<double>, dim3(g), dim3(256), 0, st,
                           (const double*)src, m, n, rs, cs, (double*)dst);
    else
        hipLaunchKernelGGL(pack_kernel<float>, dim3(g), dim3(256), 0, st,
                           (const float*)src, m, n, rs, cs, (float*)dst);
    HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Launch plumbing
// ---------------------------------------------------------------------------

// Split-k launch parameters: `parts` partitions of `kchunk` (the last one
// shorter) into the workspace `ws`.
struct SplitK {
    int parts = 1;
    int64_t kchunk = 0;
    void* ws = nullptr;
};

template <typename T, class TL>
static void launch_tl(const TileProblem<T>& p, bool a_k, bool b_k, const GridInfo& g,
                      int nblocks, const SplitK& sk, hipStream_t st) {
    dim3 grid(nblocks), block(TL::NT);
    if (sk.parts > 1) {
        const dim3 grid2(nblocks, sk.parts);
        T* ws = (T*)sk.ws;
        if (!a_k && !b_k)
            hipLaunchKernelGGL((gemm_splitk_kernel<T, TL, false, false>), grid2, block, 0, st, p, g,
                               sk.kchunk, ws);
        else if (!a_k && b_k)
            hipLaunchKernelGGL((gemm_splitk_kernel<T, TL, false, true>), grid2, block, 0, st, p, g,
                               sk.kchunk, ws);
        else if (a_k && !b_k)
            hipLaunchKernelGGL((gemm_splitk_kernel<T, TL, true, false>), grid2, block, 0, st, p, g,
                               sk.kchunk, ws);
        else
            hipLaunchKernelGGL((gemm_splitk_kernel<T, TL, true, true>), grid2, block, 0, st, p, g,
                               sk.kchunk, ws);
        HIP_CHECK(hipGetLastError());
        return;
    }
    if (!a_k && !b_k)
        hipLaunchKernelGGL((gemm_kernel<T, TL, false, false>), grid, block, 0, st, p, g);
    else if (!a_k && b_k)
        hipLaunchKernelGGL((gemm_kernel<T, TL, false, true>), grid, block, 0, st, p, g);
    else if (a_k && !b_k)
        hipLaunchKernelGGL((gemm_kernel<T, TL, true, false>), grid, block, 0, st, p, g);
    else
        hipLaunchKernelGGL((gemm_kernel<T, TL, true, true>), grid, block, 0, st, p, g);
    HIP_CHECK(hipGetLastError());
}

template <typename T>
static void launch_tile(int tile, const TileProblem<T>& p, bool a_k, bool b_k,
                        GridInfo g, const SplitK& sk, hipStream_t st) {
    const TileShape& ts = tile_shape(dt_of<T>(), tile);
    g.grid_m = (int)((p.m + ts.bm - 1) / ts.bm);
    g.grid_n = (int)((p.n + ts.bn - 1) / ts.bn);
    if (g.group_m < 1) g.group_m = 1;
    const int nblocks = g.grid_m * g.grid_n;
    if constexpr (std::is_same<T, double>::value) {
        switch (tile) {
#define BLIS_AMD_CASE(idx, BM, BN, BK, WM, WN) \
        case idx: launch_tl<double, TileCfg<double, BM, BN, BK, WM, WN>>(p, a_k, b_k, g, nblocks, sk, st); return;
        BLIS_AMD_TILES_F64(BLIS_AMD_CASE)
#undef BLIS_AMD_CASE
        }
    } else {
        switch (tile) {
#define BLIS_AMD_CASE(idx, BM, BN, BK, WM, WN) \
        case idx: launch_tl<float, TileCfg<float, BM, BN, BK, WM, WN>>(p, a_k, b_k, g, nblocks, sk, st); return;
        BLIS_AMD_TILES_F32(BLIS_AMD_CASE)
#undef BLIS_AMD_CASE
        }
    }
    abort_with(BLIS_INVALID_UKR_ID, "gemm: tile index out of range");
}

static void transpose_run(Dt dt, const void* src, int64_t rows, int64_t cols, int64_t lds,
                          void* dst, int64_t ldd, hipStream_t st) {
    const int64_t tiles = ((rows + 63) / 64) * ((cols + 63) / 64);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 16384));
    if (dt == Dt::F64)
        hipLaunchKernelGGL(transpose_kernel<double>, dim3(g), dim3(256), 0, st, (const double*)src,
                           rows, cols, lds, (double*)dst, ldd);
    else
        hipLaunchKernelGGL(transpose_kernel<float>, dim3(g), dim3(256), 0, st, (const float*)src,
                           rows, cols, lds, (float*)dst, ldd);
    HIP_CHECK(hipGetLastError());
}

// The FP64 128x128 tile of sixteen 32x32 waves (tile 13) stages a
// k-contiguous operand into LDS by transposing scalar stores, measured
// about 4% slower (8192^3 NN vs NT) than its rows-contiguous path. For large
// problems on that tile it pays to transpose the operand once (rows x k,
// reused `reuse` times) and take the faster path; the arithmetic is the same
// either way. The default large FP64 tile (6, eight 64x32 waves) stages it
// within 1.2% of its rows-contiguous path, less than the transpose costs
// (8192^3 NN: 16.11 ms as is, 16.23 ms transposed), so it takes the operand
// as it comes.
static bool transpose_pays(Dt dt, int64_t reuse, int64_t rows, int64_t k) {
    const double es = (double)dt_size(dt);
    const double gemm_s = 2.0 * (double)reuse * rows * k / (dt == Dt::F64 ? 60e12 : 120e12);
    const double tr_s = 2.0 * rows * k * es / 4e12 + 8e-6;
    return 0.04 * gemm_s > tr_s;
}

// k-contiguous operands of a large product to the rows-contiguous path, one
// transpose each into workspace (operands already packed are left alone).
static void transpose_k_operands(Plan& p, Dt dt, Scratch& sc, hipStream_t st) {
    const size_t es = dt_size(dt);
    if (p.a_k && !p.a_packed && transpose_pays(dt, p.n, p.m, p.k)) {
        const int64_t ldw = (p.m + 3) / 4 * 4;
        void* w = sc.take((size_t)ldw * p.k * es, 0);
        transpose_run(dt, p.a, p.m, p.k, p.lda, w, ldw, st);
        p.a = w; p.lda = ldw; p.a_k = false;
    }
    if (p.b_k && !p.b_packed && transpose_pays(dt, p.m, p.n, p.k)) {
        const int64_t ldw = (p.n + 3) / 4 * 4;
        void* w = sc.take((size_t)ldw * p.k * es, 1);
        transpose_run(dt, p.b, p.n, p.k, p.ldb, w, ldw, st);
        p.b = w; p.ldb = ldw; p.b_k = false;
    }
}

int gemm_default_tile(Dt dt, int64_t m, int64_t n, int64_t k) {
    // Size heuristic from a sweep of the tile table on MI355X: the best
    // large-problem tiles are the ones with high wave occupancy (several
    // waves per SIMD hide LDS and barrier latency); small problems use small
    // tiles so that enough workgroups exist to fill the 256 CUs.
    // Large shapes re-measured with exhaustive autotuning at m, n in
    // 4096..16384 and k in 128..8192: FP64 128x128 tiles of eight 64x32
    // waves beat sixteen 32x32 waves by 4-8% (8192^3 NN 16.09 vs 16.82 ms);
    // FP32 short-k products prefer 128x64 tiles. Mid-size bands (256 to
    // 1023 tiles of 128x128) take the tile with the least total time over
    // the tuning database's shapes in that band.
    auto tiles = [&](int bm, int bn) { return ((m + bm - 1) / bm) * ((n + bn - 1) / bn); };
    if (dt == Dt::F64) {
        if (tiles(128, 128) >= 1024) return 6;   // 128x128 (BK 16), 8 waves of 64x32
        if (tiles(128, 128) >= 256) return 10;   // 64x128 (BK 16), 8 waves of 32x32
        if (tiles(64, 64) >= 512) return 3;      // 64x64 (BK 16), 4 waves of 32x32
        return 4;                                // 32x32 (BK 16), 4 waves of 16x16
    }
    if (k <= 256 && tiles(128, 64) >= 512) return 1;   // 128x64 (BK 32), 4 waves of 64x32
    if (tiles(256, 128) >= 512) return 12;       // 256x128 (BK 16), 8 waves of 64x64
    if (tiles(128, 128) >= 512) return 9;        // 128x64 (BK 32), 8 waves of 32x32
    if (tiles(128, 128) >= 256) return 10;       // 64x128 (BK 32), 8 waves of 32x32
    if (tiles(64, 64) >= 512) return 3;
    return 4;
}

// Number of k partitions for a product whose output is small against its
// depth (m, n <= k/4, k >= 2048): such products have too few output tiles to
// fill 256 CUs and a long serial k loop in each. The count aims at about 512
// workgroups of the tile gemm_split_tile gives (128x128 when m, n >= 1024,
// else 64x64), keeps every partition at least 1024 deep and uses at most 64.
// It depends on the shape only, so every tile gives the same partition.
// Measured on MI355X (device time, default path, against the earlier rule of
// about 2048 workgroups of 32x32 outputs on the selector's tile):
//   FP64 1024x1024x8192 293 against 378-420 us (8 partitions of 128x128
//   tiles against 2), 512x512x16384 185 against 202 us, 128x128x65536 75
//   against 82 us;
//   FP32 1024x1024x8192 160 against 189 us, 256x256x32768 64 against 72 us,
//   128x128x65536 53 against 57 us.
static int gemm_default_split(int64_t m, int64_t n, int64_t k) {
    if (k < 2048 || 4 * std::max(m, n) > k) return 1;
    const int64_t bt = (m >= 1024 && n >= 1024) ? 128 : 64;
    int64_t s = std::max<int64_t>(1, 512 / (((m + bt - 1) / bt) * ((n + bt - 1) / bt)));
    s = std::min<int64_t>(s, k / 1024);
    return (int)std::min<int64_t>(s, 64);
}

// Tile for a product that gemm_default_split splits: 128x128 tiles when
// m, n >= 1024 (FP64 table tile 6, eight 64x32 waves; FP32 tile 0, four
// 64x64 waves); 64x64 tiles when there are at least four of them (FP64 tile
// 8, BK 32; FP32 tile 3); else 32x32 tiles (tile 4). A single 64x64 tile
// leaves too few workgroups: FP64 64x64x262144 took 222 us on it against
// 182 us on 32x32 tiles.
static int gemm_split_tile(Dt dt, int64_t m, int64_t n) {
    if (m >= 1024 && n >= 1024) return dt == Dt::F64 ? 6 : 0;
    if (((m + 63) / 64) * ((n + 63) / 64) >= 4) return dt == Dt::F64 ? 8 : 3;
    return 4;
}

// Process-wide override used when a caller passes no configuration
// (set through bli_amd_gemm_set_config; tile -1 restores the heuristic).
static GemmConfig g_override;

// Split-k launch parameters for `split` requested partitions of whole
// 64-deep k blocks (every tile's BK divides 64, so the partition, and with
// it the rounding, is the same for every tile).
static SplitK split_params(int64_t k, int split) {
    SplitK sk;
    if (split > 1) {
        const int64_t blocks = (k + 63) / 64;
        const int64_t want = std::min(split, 64);
        sk.kchunk = (blocks + want - 1) / want * 64;
        sk.parts = (int)((k + sk.kchunk - 1) / sk.kchunk);
    }
    return sk;
}

// Launch a normalised, packed plan with configuration cfg: the tile kernel,
// and for split-k the partition products into scratch and one reduction.
static void launch_plan(Dt dt, Plan& pl, const GemmConfig& cfg, double alpha, double beta,
                        Scratch& sc, hipStream_t st) {
    GridInfo g{};
    g.group_m = cfg.group_m;
    g.xcd_remap = cfg.xcd_remap;
    const int tile = cfg.tile >= 0 ? cfg.tile : gemm_default_tile(dt, pl.m, pl.n, pl.k);
    if (tile < 0 || tile >= tile_count(dt)) abort_with(BLIS_INVALID_UKR_ID, "gemm: bad tile");
    if (dt == Dt::F64 && tile == 13) transpose_k_operands(pl, dt, sc, st);
    // Split-k: each partition a full product into its own m x n slice of
    // scratch, then one reduction into C.
    SplitK sk = split_params(pl.k, cfg.split_k);
    if (gemm_sup_deep_k(tile, sk.parts, pl.k)) {
        gemm_sup_run(dt, pl, alpha, beta, st);
        return;
    }
    const size_t es = dt_size(dt);
    if (sk.parts > 1) sk.ws = sc.take((size_t)sk.parts * pl.m * pl.n * es, 17);
    if (dt == Dt::F64) {
        TileProblem<double> p = make_problem<double>(pl, alpha, beta);
        if (sk.parts > 1) p.vec_c = pl.m % Mma<double>::V == 0;
        launch_tile<double>(tile, p, pl.a_k, pl.b_k, g, sk, st);
    } else {
        TileProblem<float> p = make_problem<float>(pl, alpha, beta);
        if (sk.parts > 1) p.vec_c = pl.m % Mma<float>::V == 0;
        launch_tile<float>(tile, p, pl.a_k, pl.b_k, g, sk, st);
    }
    if (sk.parts > 1) {
        const int gr = grid_for(pl.m * pl.n);
        if (dt == Dt::F64)
            hipLaunchKernelGGL(splitk_reduce_kernel<double>, dim3(gr), dim3(256), 0, st,
                               (const double*)sk.ws, sk.parts, pl.m, pl.n, (double*)pl.c, pl.c_si,
                               pl.c_sj, beta);
        else
            hipLaunchKernelGGL(splitk_reduce_kernel<float>, dim3(gr), dim3(256), 0, st,
                               (const float*)sk.ws, sk.parts, pl.m, pl.n, (float*)pl.c, pl.c_si,
                               pl.c_sj, (float)beta);
        HIP_CHECK(hipGetLastError());
    }
}

void gemm_run(const GemmDesc& d, const GemmConfig* cfg, hipStream_t st) {
    if (d.m <= 0 || d.n <= 0) return;
    if (cfg == nullptr) cfg = &g_override;
    if (d.k <= 0 || d.alpha == 0.0) {
        scalm_run(d.dt, d.m, d.n, d.beta, d.c, st);
        return;
    }
    Plan pl = layout(d);
    // Small problems with no configuration given take the sup path, ahead
    // of the heuristic and the learned selector (exhaustive selection times
    // the table's tiles instead).
    if (cfg == &g_override && g_override.tile < 0 && g_override.split_k <= 1 &&
        !pl.a_packed && !pl.b_packed && gemm_sup_applies(pl.m, pl.n, pl.k) &&
        selector_mode() != SELECT_EXHAUSTIVE) {
        gemm_sup_run(d.dt, pl, d.alpha, d.beta, st);
        return;
    }
    GemmConfig picked;
    if (cfg == &g_override && g_override.tile < 0) {
        // No configuration given: the selector (heuristic, learned or
        // exhaustive) picks one from the normalised shape, before anything
        // is packed into scratch.
        const bool va = vec_ok(pl.a, pl.a_packed, pl.lda, d.dt);
        const bool vb = vec_ok(pl.b, pl.b_packed, pl.ldb, d.dt);
        const bool vc = aligned16(pl.c) && pl.c_si == 1 && pl.c_sj % (d.dt == Dt::F64 ? 2 : 4) == 0;
        picked = config_of(selector_pick(d.dt, pl.m, pl.n, pl.k, pl.a_k, pl.b_k, va, vb, vc));
        picked.split_k = g_override.split_k > 1 ? g_override.split_k
                                                : gemm_default_split(pl.m, pl.n, pl.k);
        // A product split by the shape rule takes the tile measured for it,
        // unless the selector timed every configuration itself.
        if (g_override.split_k <= 1 && picked.split_k > 1 && selector_mode() != SELECT_EXHAUSTIVE)
            picked.tile = gemm_split_tile(d.dt, pl.m, pl.n);
        cfg = &picked;
    }
    Scratch sc;
    pack_operands(pl, d.dt, sc, st);
    launch_plan(d.dt, pl, *cfg, d.alpha, d.beta, sc, st);
}

template <typename T, class TL>
static void launch_strided(const TileProblem<T>& p, bool a_k, bool b_k, const GridInfo& g,
                           int count, int64_t sa, int64_t sb, int64_t sc, hipStream_t st) {
    const dim3 grid(g.grid_m * g.grid_n, count), block(TL::NT);
    if (!a_k && !b_k)
        hipLaunchKernelGGL((gemm_strided_kernel<T, TL, false, false>), grid, block, 0, st, p, g,
                           sa, sb, sc);
    else if (!a_k && b_k)
        hipLaunchKernelGGL((gemm_strided_kernel<T, TL, false, true>), grid, block, 0, st, p, g,
                           sa, sb, sc);
    else if (a_k && !b_k)
        hipLaunchKernelGGL((gemm_strided_kernel<T, TL, true, false>), grid, block, 0, st, p, g,
                           sa, sb, sc);
    else
        hipLaunchKernelGGL((gemm_strided_kernel<T, TL, true, true>), grid, block, 0, st, p, g,
                           sa, sb, sc);
    HIP_CHECK(hipGetLastError());
}

void gemm_strided_run(const GemmDesc& d, int count, int64_t sa, int64_t sb, int64_t sc,
                      hipStream_t st) {
    if (d.m <= 0 || d.n <= 0 || count <= 0) return;
    if (d.c.rs != 1 || d.k <= 0 || d.alpha == 0.0)
        abort_with(BLIS_INVALID_UKR_ID, "gemm strided: C must be column-stored, k > 0, alpha != 0");
    Plan pl = layout(d);
    if (pl.a_packed || pl.b_packed)
        abort_with(BLIS_INVALID_UKR_ID, "gemm strided: operands need a unit stride");
    // The 32x32 tile (tile 4 of both tables): these batches are of small
    // products, several hundred tiles in all.
    GridInfo g{};
    g.group_m = 8;
    g.xcd_remap = 1;
    g.grid_m = (int)((pl.m + 31) / 32);
    g.grid_n = (int)((pl.n + 31) / 32);
    if (d.dt == Dt::F64) {
        constexpr int V = Mma<double>::V;
        TileProblem<double> p = make_problem<double>(pl, d.alpha, d.beta);
        p.vec_a = p.vec_a && sa % V == 0;
        p.vec_b = p.vec_b && sb % V == 0;
        p.vec_c = p.vec_c && sc % V == 0;
        launch_strided<double, TileCfg<double, 32, 32, 16, 16, 16>>(p, pl.a_k, pl.b_k, g, count,
                                                                   sa, sb, sc, st);
    } else {
        constexpr int V = Mma<float>::V;
        TileProblem<float> p = make_problem<float>(pl, d.alpha, d.beta);
        p.vec_a = p.vec_a && sa % V == 0;
        p.vec_b = p.vec_b && sb % V == 0;
        p.vec_c = p.vec_c && sc % V == 0;
        launch_strided<float, TileCfg<float, 32, 32, 32, 16, 16>>(p, pl.a_k, pl.b_k, g, count,
                                                                 sa, sb, sc, st);
    }
}

// ---------------------------------------------------------------------------
// GEMM without allocation (bli_amd_gemm_noalloc)
// ---------------------------------------------------------------------------
//
// The device memory a GEMM may take beyond its operands (packed copies of
// general-stride operands, split-k partitions) comes from one fixed array
// per GPU instead of the workspace arena. In host code `scratch` only
// names the array; its device address on the current GPU comes from
// hipGetSymbolAddress, taken once per device and kept in g_scratch_ptr.

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mutex;

// Device address of `scratch` on the current GPU (nullptr on failure).
static void* scratch_address(int dev) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

} // namespace blis_amd

extern "C" __attribute__((visibility("default")))
void bli_amd_gemm_set_config(int tile, int group_m, int xcd_remap, int split_k) {
    blis_amd::init_once();
    std::lock_guard<std::mutex> lock(blis_amd::api_mutex());
    blis_amd::g_override.tile = tile;
    blis_amd::g_override.group_m = group_m > 0 ? group_m : 8;
    blis_amd::g_override.xcd_remap = xcd_remap;
    blis_amd::g_override.split_k = split_k > 0 ? split_k : 1;
}

extern "C" __attribute__((visibility("default")))
int bli_amd_gemm_tile_count(int is_double) {
    return blis_amd::tile_count(is_double ? blis_amd::Dt::F64 : blis_amd::Dt::F32);
}

extern "C" __attribute__((visibility("default")))
int bli_amd_gemm_default_tile(int is_double, long long m, long long n, long long k) {
    return blis_amd::gemm_default_tile(is_double ? blis_amd::Dt::F64 : blis_amd::Dt::F32, m, n, k);
}

// C := beta*C + alpha*op(A)*op(B) for FP32/FP64 device operands, with no
// device allocation: scratch comes from the library's fixed per-GPU array
// (64 MiB). The kernels and their configuration are those of bli_?gemm with
// the default (heuristic) selector, so the results are the same bits.
extern "C" __attribute__((visibility("default")))
int bli_amd_gemm_noalloc(num_t dt, trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                         const void* alpha, const void* a, inc_t rs_a, inc_t cs_a,
                         const void* b, inc_t rs_b, inc_t cs_b, const void* beta, void* c,
                         inc_t rs_c, inc_t cs_c, void* stream) {
    using namespace blis_amd;
    if (dt != BLIS_FLOAT && dt != BLIS_DOUBLE) return BLIS_INVALID_DATATYPE;
    auto valid = [](trans_t t) {
        return t == BLIS_NO_TRANSPOSE || t == BLIS_TRANSPOSE || t == BLIS_CONJ_NO_TRANSPOSE ||
               t == BLIS_CONJ_TRANSPOSE;
    };
    if (!valid(transa) || !valid(transb)) return BLIS_INVALID_TRANS;
    if (m < 0 || n < 0 || k < 0) return BLIS_NEGATIVE_DIMENSION;
    if (alpha == nullptr || beta == nullptr) return BLIS_NULL_POINTER;
    if (m == 0 || n == 0) return BLIS_SUCCESS;
    const Dt d = dt == BLIS_DOUBLE ? Dt::F64 : Dt::F32;
    const double al = d == Dt::F64 ? *(const double*)alpha : (double)*(const float*)alpha;
    const double be = d == Dt::F64 ? *(const double*)beta : (double)*(const float*)beta;
    const bool touch_ab = k > 0 && al != 0.0;
    // Operands must already be in device memory (host staging allocates).
    if (c == nullptr || !is_device_pointer(c)) return BLIS_FAILURE;
    if (touch_ab && (a == nullptr || b == nullptr || !is_device_pointer(a) ||
                     !is_device_pointer(b)))
        return BLIS_FAILURE;
    const bool ta = (unsigned)transa & BLIS_TRANS_BIT, tb = (unsigned)transb & BLIS_TRANS_BIT;
    GemmDesc g;
    g.dt = d; g.m = m; g.n = n; g.k = k;
    g.a = { const_cast<void*>(a), ta ? cs_a : rs_a, ta ? rs_a : cs_a };
    g.b = { const_cast<void*>(b), tb ? cs_b : rs_b, tb ? rs_b : cs_b };
    g.c = { c, rs_c, cs_c };
    g.alpha = al; g.beta = be;
    hipStream_t st = (hipStream_t)stream;
    if (!touch_ab) {
        scalm_run(d, m, n, be, g.c, st);
        return BLIS_SUCCESS;
    }
    Plan pl = layout(g);
    GemmConfig cfg;
    cfg.tile = gemm_default_tile(d, pl.m, pl.n, pl.k);
    cfg.split_k = gemm_default_split(pl.m, pl.n, pl.k);
    if (cfg.split_k > 1) cfg.tile = gemm_split_tile(d, pl.m, pl.n);
    // Scratch this call needs, carved as Scratch::take does (256-byte
    // aligned pieces): packed general-stride operands and split-k slices.
    const size_t es = dt_size(d);
    size_t need = 0;
    if (pl.a_packed) need += align256((size_t)pl.m * pl.k * es);
    if (pl.b_packed) need += align256((size_t)pl.k * pl.n * es);
    const SplitK sk = split_params(pl.k, cfg.split_k);
    if (sk.parts > 1) need += align256((size_t)sk.parts * pl.m * pl.n * es);
    if (need > NOALLOC_SCRATCH_BYTES) return BLIS_INSUFFICIENT_STACK_BUF_SIZE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return BLIS_FAILURE;
    Scratch sc;
    sc.base = (char*)scratch_address(dev);
    sc.cap = NOALLOC_SCRATCH_BYTES;
    if (sc.base == nullptr) return BLIS_FAILURE;
    pack_operands(pl, d, sc, st);
    launch_plan(d, pl, cfg, al, be, sc, st);
    return BLIS_SUCCESS;
}
