// gemm_plan.hpp — host-side GEMM planning shared by the plain GEMM
// (gemm.hip) and the structured level-3 operations (struc.hip): storage
// normalisation (BLIS's operation transpose, frame/3/gemm/bli_gemm_cntl.c:
// 104-161), general-stride packing (the packm case of ref_kernels/1m/
// bli_packm_cxk_ref.c:78-147), the tile problem handed to the kernels, and
// the XCD-aware tile raster.
#pragma once

#include "internal.hpp"
#include "gemm_tile.hpp"
#include "gemm_tiles.hpp"

namespace blis_amd {

using namespace dev;

// dst (column-major m x n, ld = m) := src with arbitrary strides (gemm.hip).
void pack_run(Dt dt, const void* src, int64_t m, int64_t n, int64_t rs, int64_t cs,
              void* dst, hipStream_t st);

__device__ inline void tile_coords(int bid, int nblocks, const GridInfo& g, int& tm, int& tn) {
    if (g.xcd_remap) {
        // Workgroups are dispatched round-robin over the 8 XCDs. Give each
        // XCD a contiguous range of logical tile ids so that tiles sharing
        // rows of A / columns of B hit the same L2.
        const int X = 8;
        const int q = nblocks / X, r = nblocks % X;
        const int x = bid % X, l = bid / X;
        bid = x * q + (x < r ? x : r) + l;
    }
    const int per_group = g.group_m * g.grid_n;
    const int group = bid / per_group;
    const int first_m = group * g.group_m;
    const int gsz = min(g.grid_m - first_m, g.group_m);
    const int in = bid - group * per_group;
    tm = first_m + in % gsz;
    tn = in / gsz;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Normalised GEMM ready to launch: A mode, B mode, leading dimensions.
struct Plan {
    int64_t m, n, k;
    const void* a; int64_t lda; bool a_k;
    const void* b; int64_t ldb; bool b_k;
    void* c; int64_t c_si, c_sj;
    bool a_packed = false, b_packed = false;   // operand copied into scratch
    // Source of an operand that still has to be packed (general strides).
    View a_src, b_src;
};

// Device memory a GEMM takes beyond its operands: the library's grow-only
// workspace arena, or, when `base` is set, a fixed region handed in by the
// caller, carved in 256-byte aligned pieces (nothing is allocated then).
struct Scratch {
    char*  base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    void* take(size_t bytes, int slot) {
        if (base == nullptr) return workspace(bytes, slot);
        const size_t start = (used + 255) & ~(size_t)255;
        if (start + bytes > cap)
            abort_with(BLIS_INSUFFICIENT_STACK_BUF_SIZE, "gemm: fixed scratch region too small");
        used = start + bytes;
        return base + start;
    }
};

// Storage normalisation without side effects: row-stored C is recast as
// C^T = B^T A^T, each operand is classified as rows- or k-contiguous, and an
// operand with neither unit stride is marked for packing into a contiguous
// column-major copy (A: m x k, ld m; B: k x n, ld k). pack_operands() then
// makes those copies. Splitting the two lets the configuration be chosen
// before any scratch memory is written (the exhaustive autotuner runs GEMMs
// of its own that use the same scratch).
inline Plan layout(const GemmDesc& d0) {
    GemmDesc d = d0;
    // Row-stored C: compute C^T = B^T A^T instead.
    if (d.c.rs != 1 && d.c.cs == 1) {
        GemmDesc t = d;
        t.m = d.n; t.n = d.m;
        t.a.ptr = d.b.ptr; t.a.rs = d.b.cs; t.a.cs = d.b.rs;
        t.b.ptr = d.a.ptr; t.b.rs = d.a.cs; t.b.cs = d.a.rs;
        t.c.rs = d.c.cs; t.c.cs = d.c.rs;
        d = t;
    }
    Plan p;
    p.m = d.m; p.n = d.n; p.k = d.k;
    p.c = d.c.ptr; p.c_si = d.c.rs; p.c_sj = d.c.cs;
    if (d.a.rs == 1) { p.a = d.a.ptr; p.lda = d.a.cs; p.a_k = false; }
    else if (d.a.cs == 1) { p.a = d.a.ptr; p.lda = d.a.rs; p.a_k = true; }
    else { p.a = nullptr; p.a_src = d.a; p.lda = d.m; p.a_k = false; p.a_packed = true; }
    if (d.b.rs == 1) { p.b = d.b.ptr; p.ldb = d.b.cs; p.b_k = true; }
    else if (d.b.cs == 1) { p.b = d.b.ptr; p.ldb = d.b.rs; p.b_k = false; }
    else { p.b = nullptr; p.b_src = d.b; p.ldb = d.k; p.b_k = true; p.b_packed = true; }
    return p;
}

inline void pack_operands(Plan& p, Dt dt, Scratch& sc, hipStream_t st) {
    const size_t es = dt_size(dt);
    if (p.a_packed && p.a == nullptr) {
        void* w = sc.take((size_t)p.m * p.k * es, 0);
        pack_run(dt, p.a_src.ptr, p.m, p.k, p.a_src.rs, p.a_src.cs, w, st);
        p.a = w;
    }
    if (p.b_packed && p.b == nullptr) {
        void* w = sc.take((size_t)p.k * p.n * es, 1);
        pack_run(dt, p.b_src.ptr, p.k, p.n, p.b_src.rs, p.b_src.cs, w, st);
        p.b = w;
    }
}

// 16-byte vector access to an operand: packed copies start 256-byte aligned.
inline bool vec_ok(const void* ptr, bool packed, int64_t ld, Dt dt) {
    const int64_t V = dt == Dt::F64 ? 2 : 4;
    return (packed || aligned16(ptr)) && ld % V == 0;
}

template <typename T>
inline TileProblem<T> make_problem(const Plan& pl, double alpha, double beta) {
    constexpr int V = Mma<T>::V;
    TileProblem<T> p;
    p.a = (const T*)pl.a; p.b = (const T*)pl.b; p.c = (T*)pl.c;
    p.m = pl.m; p.n = pl.n; p.k = pl.k;
    p.lda = pl.lda; p.ldb = pl.ldb;
    p.c_si = pl.c_si; p.c_sj = pl.c_sj;
    p.alpha = (T)alpha; p.beta = (T)beta;
    p.vec_a = aligned16(pl.a) && (pl.lda % V == 0);
    p.vec_b = aligned16(pl.b) && (pl.ldb % V == 0);
    p.vec_c = aligned16(pl.c) && pl.c_si == 1 && (pl.c_sj % V == 0);
    return p;
}

// The single-call small-problem path (sup.hip): applies to m, n, k all at
// most SMALL_GEMM_MAX_DIM with k > 64; the plan's operands must have a unit
// stride. gemm_sup_deep_k: the 32 x 32 table tile over k > 64 in one
// partition runs as the same kernel. BLIS_AMD_SUP=0 in the environment
// turns both off (read once).
bool gemm_sup_applies(int64_t m, int64_t n, int64_t k);
bool gemm_sup_deep_k(int tile, int parts, int64_t k);
void gemm_sup_run(Dt dt, const Plan& pl, double alpha, double beta, hipStream_t st);

} // namespace blis_amd
