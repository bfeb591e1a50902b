// select.hip — GEMM configuration space, autotuner and tile selector.
//
// BLIS fixes its blocksizes per architecture (config/zen3/bli_cntx_init_zen3.c:
// 204-220, the MC/KC/NC/MR/NR table consulted by bli_gemm_cntl_init,
// frame/3/gemm/bli_gemm_cntl.c:75-565). On the GPU the analogous choice is the
// workgroup tile and the tile schedule. This file owns the space of those
// choices and three ways to pick one per shape:
//   heuristic   the fixed size rule of gemm.hip (gemm_default_tile);
//   learned     the MLP of mlp.hip, scoring all configurations from 16 shape
//               features (one inference per new shape, then cached);
//   exhaustive  time every configuration on scratch operands of the same
//               shape and keep the fastest (cached).
// The learned selector is the default (weights shipped beside the library);
// it runs on the host, a few microseconds per new shape. Single calls with
// m, n, k <= 512 take the small-problem path before any selector but the
// exhaustive one.
// Every configuration computes bit-identical results (same MFMA k order, one
// fixed epilogue), so the selector changes speed only.
//
// Configuration c: tile = c / GEMM_NSCHED, schedule = c % GEMM_NSCHED, where a
// schedule is a (raster group height, XCD remap) pair.
#include "internal.hpp"
#include "gemm_tiles.hpp"
#include "mlp.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>

namespace blis_amd {

namespace detail_select {

struct Sched { int group_m, xcd_remap; };
constexpr Sched SCHEDS[GEMM_NSCHED] = {{8, 1}, {4, 0}, {16, 1}};

#define BLIS_AMD_ONE(i, BM, BN, BK, WM, WN) +1
static_assert((0 BLIS_AMD_TILES_F64(BLIS_AMD_ONE)) == GEMM_NTILE, "f64 tile table size");
static_assert((0 BLIS_AMD_TILES_F32(BLIS_AMD_ONE)) == GEMM_NTILE, "f32 tile table size");
#undef BLIS_AMD_ONE

__global__ void fill_kernel(double* d, float* f, int64_t n, uint64_t seed) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)i * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
        if (d) d[i] = u;
        else f[i] = (float)u;
    }
}

struct Key {
    int dt, a_k, b_k, va, vb, vc;
    int64_t m, n, k;
    bool operator==(const Key& o) const {
        return dt == o.dt && a_k == o.a_k && b_k == o.b_k && va == o.va && vb == o.vb &&
               vc == o.vc && m == o.m && n == o.n && k == o.k;
    }
};
struct KeyHash {
    size_t operator()(const Key& q) const {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
        mix(q.dt); mix(q.a_k); mix(q.b_k); mix(q.va); mix(q.vb); mix(q.vc);
        mix((uint64_t)q.m); mix((uint64_t)q.n); mix((uint64_t)q.k);
        return (size_t)h;
    }
};

struct SelectorState {
    bool initialised = false;
    int mode = SELECT_HEURISTIC;
    MlpHost net;   // the learned selector's network (host copy)
    std::unordered_map<Key, int, KeyHash> cache;
    bool warmed[2] = {false, false};
};

SelectorState& SS() {
    static SelectorState s;
    return s;
}

// <directory of libblis_amd.so>/../data/selector_gfx950.bin
std::string default_weights_path() {
    Dl_info info;
    if (dladdr((void*)&default_weights_path, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t slash = p.find_last_of('/');
        p = (slash == std::string::npos) ? std::string(".") : p.substr(0, slash);
        return p + "/../data/selector_gfx950.bin";
    }
    return "selector_gfx950.bin";
}

} // namespace detail_select
using namespace detail_select;

GemmConfig config_of(int cfg) {
    GemmConfig c;
    c.tile = cfg / GEMM_NSCHED;
    c.group_m = SCHEDS[cfg % GEMM_NSCHED].group_m;
    c.xcd_remap = SCHEDS[cfg % GEMM_NSCHED].xcd_remap;
    c.split_k = 1;
    return c;
}

// The 16 shape features of the learned selector, from the normalised
// problem (column-stored C; a_k: A stored k-contiguous, b_k: B stored
// k-contiguous; v*: 16-byte vector access possible for A, B, C).
void gemm_features(Dt dt, int64_t m, int64_t n, int64_t k, bool a_k, bool b_k, bool va, bool vb,
                   bool vc, float* f) {
    auto lg = [](double x) { return (float)std::log2(std::max(1.0, x)); };
    auto tiles = [&](int64_t bm, int64_t bn) {
        return (double)((m + bm - 1) / bm) * (double)((n + bn - 1) / bn);
    };
    f[0] = lg((double)m);
    f[1] = lg((double)n);
    f[2] = lg((double)k);
    f[3] = dt == Dt::F64 ? 1.f : 0.f;
    f[4] = a_k ? 1.f : 0.f;
    f[5] = b_k ? 1.f : 0.f;
    f[6] = va ? 1.f : 0.f;
    f[7] = vb ? 1.f : 0.f;
    f[8] = vc ? 1.f : 0.f;
    f[9] = lg(tiles(128, 128));
    f[10] = lg(tiles(64, 64));
    f[11] = lg(tiles(32, 32));
    f[12] = (m % 64 == 0) ? 1.f : 0.f;
    f[13] = (n % 64 == 0) ? 1.f : 0.f;
    f[14] = (k % 32 == 0) ? 1.f : 0.f;
    f[15] = lg((double)std::min(m, std::min(n, k)));
}

// Default: the learned selector with the weights shipped beside the
// library (or BLIS_AMD_SELECTOR_WEIGHTS). BLIS_AMD_SELECTOR=heuristic or
// exhaustive chooses another mode; an explicit "mlp" whose weights cannot be
// read is an error, while missing default weights leave the heuristic in
// place with one warning.
static void selector_init() {
    SelectorState& s = SS();
    if (s.initialised) return;
    s.initialised = true;
    const char* mode = std::getenv("BLIS_AMD_SELECTOR");
    const char* w = std::getenv("BLIS_AMD_SELECTOR_WEIGHTS");
    const std::string path = w ? std::string(w) : default_weights_path();
    if (mode == nullptr) {
        if (mlp_read(path.c_str(), s.net.p, s.net.musd)) s.mode = SELECT_LEARNED;
        else std::fprintf(stderr, "libblis_amd: selector weights '%s' not found; using the size "
                                  "heuristic\n", path.c_str());
        return;
    }
    if (!std::strcmp(mode, "mlp") || !std::strcmp(mode, "learned")) {
        if (!mlp_read(path.c_str(), s.net.p, s.net.musd)) {
            std::fprintf(stderr, "libblis_amd: cannot load selector weights '%s'\n", path.c_str());
            std::abort();
        }
        s.mode = SELECT_LEARNED;
    } else if (!std::strcmp(mode, "exhaustive")) {
        s.mode = SELECT_EXHAUSTIVE;
    } else if (std::strcmp(mode, "heuristic") && std::strcmp(mode, "default")) {
        std::fprintf(stderr, "libblis_amd: BLIS_AMD_SELECTOR='%s' is not one of "
                             "heuristic, mlp, exhaustive\n", mode);
        std::abort();
    }
}

// Compile every kernel variant once (code objects load lazily on first
// launch, which must not be timed).
static void warm_all(Dt dt, void* a, void* b, void* c, hipStream_t st) {
    for (int cfg = 0; cfg < GEMM_NCFG; ++cfg)
        for (int t = 0; t < 4; ++t) {
            GemmDesc d;
            d.dt = dt; d.m = 64; d.n = 64; d.k = 64;
            d.a = {a, (t & 1) ? 64 : 1, (t & 1) ? 1 : 64};
            d.b = {b, (t & 2) ? 64 : 1, (t & 2) ? 1 : 64};
            d.c = {c, 1, 64};
            d.alpha = 1.0; d.beta = 0.5;
            const GemmConfig gc = config_of(cfg);
            gemm_run(d, &gc, st);
        }
    HIP_CHECK(hipStreamSynchronize(st));
}

int gemm_autotune(Dt dt, bool ta, bool tb, int64_t m, int64_t n, int64_t k, int64_t lda,
                  int64_t ldb, int64_t ldc, float* ms_out) {
    hipStream_t st = stream();
    const size_t es = dt_size(dt);
    const int64_t a_elems = lda * (ta ? m : k), b_elems = ldb * (tb ? k : n), c_elems = ldc * n;
    const int64_t need = std::max<int64_t>({a_elems, b_elems, c_elems, 64 * 64});
    void* A = workspace((size_t)need * es, 8);
    void* B = workspace((size_t)need * es, 9);
    void* Cm = workspace((size_t)need * es, 10);
    const bool dbl = dt == Dt::F64;
    void* bufs[3] = {A, B, Cm};
    const int64_t cnt[3] = {a_elems, b_elems, c_elems};
    for (int i = 0; i < 3; ++i)
        hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, st,
                           dbl ? (double*)bufs[i] : nullptr, dbl ? nullptr : (float*)bufs[i],
                           std::max<int64_t>(cnt[i], 64 * 64), (uint64_t)(i + 1) * 7919u);
    HIP_CHECK(hipGetLastError());
    SelectorState& s = SS();
    if (!s.warmed[dbl]) {
        warm_all(dt, A, B, Cm, st);
        s.warmed[dbl] = true;
    }
    GemmDesc d;
    d.dt = dt; d.m = m; d.n = n; d.k = k;
    // op(A)(i,p) = A[i*rs + p*cs]: A is m x k (lda >= m) or, transposed,
    // k x m (lda >= k). op(B)(p,j) likewise with B k x n or n x k.
    d.a = {A, ta ? lda : 1, ta ? 1 : lda};
    d.b = {B, tb ? ldb : 1, tb ? 1 : ldb};
    d.c = {Cm, 1, ldc};
    d.alpha = 1.0; d.beta = 0.5;
    // Every element the GEMM touches must lie inside its buffer.
    auto last = [](const View& v, int64_t rows, int64_t cols) {
        return (rows - 1) * v.rs + (cols - 1) * v.cs;
    };
    if (last(d.a, m, k) >= need || last(d.b, k, n) >= need || last(d.c, m, n) >= need ||
        last(d.a, m, k) >= a_elems || last(d.b, k, n) >= b_elems || last(d.c, m, n) >= c_elems)
        abort_with(BLIS_INVALID_DIM_STRIDE_COMBINATION, "autotune: operand span exceeds buffer");
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    for (int cfg = 0; cfg < GEMM_NCFG; ++cfg) {
        const GemmConfig gc = config_of(cfg);
        HIP_CHECK(hipEventRecord(e0, st));
        gemm_run(d, &gc, st);
        HIP_CHECK(hipEventRecord(e1, st));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        // Short kernels: average further launches (about 0.2 ms in total).
        if (ms < 0.05f) {
            const int reps = std::min(50, std::max(2, (int)std::ceil(0.2f / std::max(ms, 0.002f))));
            HIP_CHECK(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r) gemm_run(d, &gc, st);
            HIP_CHECK(hipEventRecord(e1, st));
            HIP_CHECK(hipEventSynchronize(e1));
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= (float)reps;
        }
        ms_out[cfg] = ms;
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    return 0;
}

static int heuristic_cfg(Dt dt, int64_t m, int64_t n, int64_t k) {
    return gemm_default_tile(dt, m, n, k) * GEMM_NSCHED;
}

int selector_mode() {
    selector_init();
    return SS().mode;
}

int selector_pick(Dt dt, int64_t m, int64_t n, int64_t k, bool a_k, bool b_k, bool va, bool vb,
                  bool vc) {
    selector_init();
    SelectorState& s = SS();
    if (s.mode == SELECT_HEURISTIC) return heuristic_cfg(dt, m, n, k);
    const Key key{(int)dt, a_k, b_k, va, vb, vc, m, n, k};
    auto it = s.cache.find(key);
    if (it != s.cache.end()) return it->second;
    int cfg;
    if (s.mode == SELECT_LEARNED) {
        // FP64 products of at least 1024 output tiles of 128 x 128 whose
        // operands all allow 16-byte access keep the heuristic's large
        // tile: the band its rule was tuned on by exhaustive search, where
        // it measured within 1-2% of the best configuration on shapes with
        // dimensions divisible by 64 (and 88% of the FP64 peak at 8192^3),
        // while a network's pick there cost 5% at 8192^3. (The FP32 rule
        // is no better than the network in that band.)
        const int64_t t128 = ((m + 127) / 128) * ((n + 127) / 128);
        if (dt == Dt::F64 && t128 >= 1024 && va && vb && vc) {
            cfg = heuristic_cfg(dt, m, n, k);
        } else {
            float f[MLP_IN];
            gemm_features(dt, m, n, k, a_k, b_k, va, vb, vc, f);
            cfg = s.net.argmin(f);
        }
    } else {
        // Scratch operands with the leading dimensions the caller's alignment
        // implies (aligned: rows rounded to 16 bytes; unaligned: rows + 1).
        const int64_t V = (dt == Dt::F64) ? 2 : 4;
        auto ld = [&](int64_t rows, bool vec) {
            return vec ? (rows + V - 1) / V * V : (rows % V ? rows : rows + 1);
        };
        const int64_t lda = ld(a_k ? k : m, va), ldb = ld(b_k ? k : n, vb), ldc = ld(m, vc);
        std::vector<float> ms(GEMM_NCFG);
        gemm_autotune(dt, a_k, !b_k, m, n, k, lda, ldb, ldc, ms.data());
        cfg = (int)(std::min_element(ms.begin(), ms.end()) - ms.begin());
    }
    s.cache.emplace(key, cfg);
    return cfg;
}

} // namespace blis_amd

using namespace blis_amd;

extern "C" {

__attribute__((visibility("default"))) int bli_amd_gemm_config_count(int is_double) {
    (void)is_double;
    return GEMM_NCFG;
}

__attribute__((visibility("default")))
int bli_amd_gemm_config_describe(int is_double, int cfg, int* tile, int* group_m, int* xcd_remap,
                                 int* bm, int* bn, int* bk) {
    if (cfg < 0 || cfg >= GEMM_NCFG) return -1;
    const GemmConfig c = config_of(cfg);
    const TileShape& ts = tile_shape(is_double ? Dt::F64 : Dt::F32, c.tile);
    if (tile) *tile = c.tile;
    if (group_m) *group_m = c.group_m;
    if (xcd_remap) *xcd_remap = c.xcd_remap;
    if (bm) *bm = ts.bm;
    if (bn) *bn = ts.bn;
    if (bk) *bk = ts.bk;
    return 0;
}

__attribute__((visibility("default")))
void bli_amd_gemm_features(int is_double, trans_t transa, trans_t transb, dim_t m, dim_t n,
                           dim_t k, inc_t lda, inc_t ldb, inc_t ldc, float* out) {
    const int64_t V = is_double ? 2 : 4;
    const bool ta = ((unsigned)transa & BLIS_TRANS_BIT) != 0;
    const bool tb = ((unsigned)transb & BLIS_TRANS_BIT) != 0;
    gemm_features(is_double ? Dt::F64 : Dt::F32, m, n, k, ta, !tb, lda % V == 0, ldb % V == 0,
                  ldc % V == 0, out);
}

__attribute__((visibility("default")))
int bli_amd_gemm_autotune(int is_double, trans_t transa, trans_t transb, dim_t m, dim_t n,
                          dim_t k, inc_t lda, inc_t ldb, inc_t ldc, float* ms) {
    init_once();
    const bool ta = ((unsigned)transa & BLIS_TRANS_BIT) != 0;
    const bool tb = ((unsigned)transb & BLIS_TRANS_BIT) != 0;
    if (m <= 0 || n <= 0 || k <= 0 || lda < (ta ? k : m) || ldb < (tb ? n : k) || ldc < m)
        return -1;
    std::lock_guard<std::mutex> lock(api_mutex());
    return gemm_autotune(is_double ? Dt::F64 : Dt::F32, ta, tb, m, n, k, lda, ldb, ldc, ms);
}

__attribute__((visibility("default"))) void bli_amd_gemm_set_selector(int mode, void* net) {
    init_once();
    std::lock_guard<std::mutex> lock(api_mutex());
    selector_init();
    if (mode < SELECT_HEURISTIC || mode > SELECT_EXHAUSTIVE)
        abort_with(BLIS_INVALID_UKR_ID, "bli_amd_gemm_set_selector: unknown mode");
    SelectorState& s = SS();
    if (mode == SELECT_LEARNED) {
        MlpHost h;
        if (net != nullptr) {
            // A caller's network: its current parameters are copied.
            h.p.resize(MLP_NPARAM);
            h.musd.resize(2 * MLP_IN);
            static_cast<MlpNet*>(net)->get_params(h.p.data(), h.musd.data());
        } else if (!mlp_read(default_weights_path().c_str(), h.p, h.musd)) {
            abort_with(BLIS_EXPECTED_NONNULL_OBJECT_BUFFER, "selector weights not found");
        }
        s.net = std::move(h);
    }
    s.mode = mode;
    s.cache.clear();
}

__attribute__((visibility("default"))) int bli_amd_gemm_get_selector(void) {
    std::lock_guard<std::mutex> lock(api_mutex());
    selector_init();
    return SS().mode;
}

__attribute__((visibility("default")))
int bli_amd_gemm_selected_config(int is_double, trans_t transa, trans_t transb, dim_t m, dim_t n,
                                 dim_t k, inc_t lda, inc_t ldb, inc_t ldc) {
    init_once();
    const int64_t V = is_double ? 2 : 4;
    const bool ta = ((unsigned)transa & BLIS_TRANS_BIT) != 0;
    const bool tb = ((unsigned)transb & BLIS_TRANS_BIT) != 0;
    std::lock_guard<std::mutex> lock(api_mutex());
    return selector_pick(is_double ? Dt::F64 : Dt::F32, m, n, k, ta, !tb, lda % V == 0,
                         ldb % V == 0, ldc % V == 0);
}

} // extern "C"
