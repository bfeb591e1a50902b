// gemm_tiles.hpp — the gfx950 tile table (the GPU analogue of BLIS's
// per-architecture blocksize table, config/zen3/bli_cntx_init_zen3.c:204-220).
//
// Each entry is (index, BM, BN, BK, WM, WN): workgroup tile BM x BN of C, a
// BK-deep k slab staged through LDS, and waves owning WM x WN sub-tiles made
// of 16x16x4 MFMA blocks. Entries are ordered from largest to smallest tile;
// the "heuristic" flag marks the ones the default planner may pick (the rest
// are reachable through the autotuner and the learned selector).
#pragma once

#include <cstdint>
#include <type_traits>

namespace blis_amd {

struct GridInfo {
    int grid_m, grid_n;   // tiles along m and n
    int group_m;          // raster group height (tiles)
    int xcd_remap;        // 1: contiguous logical tiles per XCD
};

struct TileShape {
    int bm, bn, bk, wm, wn;
    int threads;
};

//                  idx  BM   BN  BK  WM  WN   (threads = 64 * (BM/WM) * (BN/WN))
#define BLIS_AMD_TILES_F64(X) \
    X(0, 128, 128, 16, 64, 64) \
    X(1, 128,  64, 16, 64, 32) \
    X(2,  64, 128, 16, 32, 64) \
    X(3,  64,  64, 16, 32, 32) \
    X(4,  32,  32, 16, 16, 16) \
    X(5, 128, 128, 16, 32, 64) \
    X(6, 128, 128, 16, 64, 32) \
    X(7, 128,  64, 16, 32, 32) \
    X(8,  64,  64, 32, 32, 32) \
    X(9, 128, 128, 32, 32, 64) \
    X(10, 64, 128, 16, 32, 32) \
    X(11, 128, 64, 32, 32, 32) \
    X(12, 128, 128, 16, 32, 32) \
    X(13, 128, 128, 32, 32, 32)

#define BLIS_AMD_TILES_F32(X) \
    X(0, 128, 128, 32, 64, 64) \
    X(1, 128,  64, 32, 64, 32) \
    X(2,  64, 128, 32, 32, 64) \
    X(3,  64,  64, 32, 32, 32) \
    X(4,  32,  32, 32, 16, 16) \
    X(5, 128, 128, 32, 32, 64) \
    X(6, 256, 128, 32, 64, 64) \
    X(7, 128, 256, 32, 64, 64) \
    X(8, 128, 128, 16, 64, 64) \
    X(9, 128,  64, 32, 32, 32) \
    X(10, 64, 128, 32, 32, 32) \
    X(11, 128, 128, 64, 64, 64) \
    X(12, 256, 128, 16, 64, 64) \
    X(13, 128, 128, 16, 32, 64)

#define BLIS_AMD_ROW(i, BM, BN, BK, WM, WN) { BM, BN, BK, WM, WN, 64 * (BM / WM) * (BN / WN) },

inline const TileShape& tile_shape(Dt dt, int idx) {
    static const TileShape f64[] = { BLIS_AMD_TILES_F64(BLIS_AMD_ROW) };
    static const TileShape f32[] = { BLIS_AMD_TILES_F32(BLIS_AMD_ROW) };
    return dt == Dt::F64 ? f64[idx] : f32[idx];
}

inline int tile_count(Dt dt) {
#define BLIS_AMD_ONE(i, BM, BN, BK, WM, WN) +1
    return dt == Dt::F64 ? (0 BLIS_AMD_TILES_F64(BLIS_AMD_ONE)) : (0 BLIS_AMD_TILES_F32(BLIS_AMD_ONE));
#undef BLIS_AMD_ONE
}

template <typename T> constexpr Dt dt_of() {
    return std::is_same<T, double>::value ? Dt::F64 : Dt::F32;
}

int gemm_default_tile(Dt dt, int64_t m, int64_t n, int64_t k);

} // namespace blis_amd
