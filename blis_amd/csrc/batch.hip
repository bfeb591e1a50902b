// batch.hip — uneven batched GEMM in one persistent launch.
//
// The reference's ?gemm_batch_ (frame/compat/extra/bla_gemm_batch.c:172-241)
// loops over groups and problems and calls bli_gemm_ex once per problem, so
// each problem pays the full front-end and thread-fork cost and small
// problems take the sup path (frame/3/bli_l3_sup.c:37-130). Here the whole
// batch is one kernel:
//   * the host turns every problem into a descriptor (operand layout folded
//     into flags, as the single-GEMM planner does) and every output tile into
//     a work item (problem index, tile index);
//   * work items are ordered longest-first by their k extent (an LPT
//     schedule) and a grid sized to the machine walks them round-robin;
//   * each problem uses either 64 x 64 or 32 x 32 tiles of the same MFMA tile
//     routine as the single GEMM (gemm_tile.hpp), so every problem's result is
//     the same as a plain launch of that tile configuration.
// A plan object keeps the descriptors on the device and can replay the
// launch as a captured HIP graph.
#include "internal.hpp"
#include "gemm_tile.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace blis_amd {

using namespace dev;

namespace detail_batch {

enum : int32_t {
    F_A_K = 1, F_B_K = 2, F_VEC_A = 4, F_VEC_B = 8, F_VEC_C = 16, F_SMALL = 32
};

struct BatchDesc {
    const void* a;
    const void* b;
    void*       c;
    int64_t     lda, ldb, c_si, c_sj;
    int32_t     m, n, k;
    int32_t     flags;
    double      alpha, beta;
    int32_t     tiles_m;     // tiles along m at the problem's own tile size
    int32_t     tiles_m32;   // tiles along m at 32 (wave-tile mode)
};

struct WorkItem {
    int32_t prob;
    int32_t tile;
};

template <typename T> struct BatchTiles;
template <> struct BatchTiles<double> {
    using Big = TileCfg<double, 64, 64, 16, 32, 32>;
    using BigHalfK = TileCfg<double, 64, 64, 8, 32, 32>;
    using Small = TileCfg<double, 32, 32, 16, 16, 16>;
};
template <> struct BatchTiles<float> {
    using Big = TileCfg<float, 64, 64, 32, 32, 32>;
    using BigHalfK = TileCfg<float, 64, 64, 16, 32, 32>;
    using Small = TileCfg<float, 32, 32, 32, 16, 16>;
};

// One persistent kernel per (tile shape, A layout, B layout) variant: each
// instantiation holds only its own tile code, so its register count (and
// with it the number of resident workgroups) is that of the plain GEMM
// kernel with the same tile, and its code fits the instruction cache.
// Work items of one variant, split into one list per XCD: list x is
// items[off[x] .. off[x+1]). Workgroups are dispatched round-robin over the
// 8 XCDs, so workgroup b (on XCD b % 8) walks list b % 8; every tile of a
// problem then runs on one XCD and its operands are fetched into one L2.
constexpr int NXCD = 8;
struct XcdSegs {
    int off[NXCD + 1];
};

// Kernel modes (a schedule choice the batch autotuner times; every mode
// computes bit-identical results):
//   0: one slab in flight, guarded element loads on edge tiles;
//   1: one slab in flight, precomputed edge addressing (EdgeState);
//   2: two slabs in flight, precomputed edge addressing;
//   3: two slabs in flight, guarded element loads;
//   4: wave tiles (gemm_batch_wave_kernel);
//   5: as 0, with registers capped for four workgroups per CU, the C tile
//      pulled into L2 one slab ahead and C written through LDS in whole
//      cache lines;
//   6: as 0, with registers capped for four workgroups per CU;
//   7, 8: half the k slab on 64 x 64 tiles (half the LDS and staging
//      registers), registers capped for five / six workgroups per CU; FP32
//      also pulls C into L2 and writes it through LDS as in mode 5.
// Modes other than 4 can differ between the variants of one plan.
constexpr int NMODE = 9;
constexpr int WAVE_MODE = 4;
// Mode used until the autotuner runs: the one measured fastest on uneven
// batches of 10,000 problems (dims log-uniform in [8, 512]), for both types.
constexpr int default_mode(Dt) { return 8; }
constexpr int mode_pf(int mode) { return (mode == 2 || mode == 3) ? 2 : 1; }
constexpr bool mode_edge(int mode) { return mode == 1 || mode == 2; }
constexpr int mode_waves(int mode) { return mode == 7 ? 5 : mode == 8 ? 6 : (mode == 5 || mode == 6) ? 4 : 1; }
// The C tile fits the half-k stages only for FP32 (64 x 68 floats = two stages).
template <typename T> constexpr bool mode_lds_epi(int mode) {
    return mode == 5 || (mode >= 7 && std::is_same<T, float>::value);
}
template <typename T> constexpr bool mode_c_pf(int mode) { return mode_lds_epi<T>(mode); }

template <typename T, class TL, bool A_K, bool B_K, int MODE>
__global__ __launch_bounds__(TL::NT)
__attribute__((amdgpu_waves_per_eu(mode_waves(MODE)))) void gemm_batch_kernel(const BatchDesc* __restrict__ descs,
                                                            const WorkItem* __restrict__ items,
                                                            XcdSegs seg) {
    __shared__ T smem[2 * TL::STAGE];
    __shared__ int sink[mode_c_pf<T>(MODE) ? TL::NT : 1];
    const int x = blockIdx.x % NXCD;
    const int slot = blockIdx.x / NXCD, nslot = gridDim.x / NXCD;   // gridDim.x % NXCD == 0
    const int end = seg.off[x + 1];
    int it = seg.off[x] + slot;
    if (it >= end) return;
    // The next work item and its descriptor are fetched while the current
    // tile runs, so a tile's operand loads do not wait behind two dependent
    // descriptor reads.
    WorkItem w = items[it];
    BatchDesc d = descs[w.prob];
    for (;;) {
        const int nit = it + nslot;
        const bool has_next = nit < end;
        WorkItem wn{};
        BatchDesc dn{};
        if (has_next) {
            wn = items[nit];
            dn = descs[wn.prob];
        }
        TileProblem<T> p;
        p.a = (const T*)d.a; p.b = (const T*)d.b; p.c = (T*)d.c;
        p.m = d.m; p.n = d.n; p.k = d.k;
        p.lda = d.lda; p.ldb = d.ldb; p.c_si = d.c_si; p.c_sj = d.c_sj;
        p.alpha = (T)d.alpha; p.beta = (T)d.beta;
        p.vec_a = (d.flags & F_VEC_A) != 0;
        p.vec_b = (d.flags & F_VEC_B) != 0;
        p.vec_c = (d.flags & F_VEC_C) != 0;
        const int64_t tm = w.tile % d.tiles_m, tn = w.tile / d.tiles_m;
        gemm_tile<T, TL, A_K, B_K, mode_pf(MODE), mode_edge(MODE), mode_lds_epi<T>(MODE),
                  mode_c_pf<T>(MODE)>(p, tm, tn, smem, sink);
        __syncthreads();
        if (!has_next) break;
        it = nit;
        w = wn;
        d = dn;
    }
}

// Variant v = 4*small + 2*a_k + b_k.
constexpr int NVARIANT = 8;
__host__ __device__ inline int variant_of(int32_t flags) {
    return ((flags & F_SMALL) ? 4 : 0) + ((flags & F_A_K) ? 2 : 0) + ((flags & F_B_K) ? 1 : 0);
}

// The whole batch in one dispatch (every mode but the wave mode). Work
// items of all variants sit in one list per XCD, longest first; workgroup b
// runs on XCD b % 8 and takes the next item of that list from a device
// counter (ctr[x]), so the variants share the machine and no workgroup
// idles while its XCD has work left. The index two items ahead is fetched,
// and the next item's descriptor loaded, while the current tile runs. The
// last workgroup to finish sets the counters back to zero for the next
// launch (ctr[NXCD] counts finished workgroups), so a plan's launches on
// one stream need no reset in between; two launches of one plan must not
// overlap.
template <typename T, int MODE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(mode_waves(MODE)))) void gemm_batch_dyn_kernel(
    const BatchDesc* __restrict__ descs, const WorkItem* __restrict__ items, XcdSegs seg,
    int* __restrict__ ctr) {
    using Big = typename std::conditional<(MODE >= 7), typename BatchTiles<T>::BigHalfK,
                                          typename BatchTiles<T>::Big>::type;
    using Small = typename BatchTiles<T>::Small;
    static_assert(Big::NT == 256 && Small::NT == 256, "one workgroup shape for every variant");
    constexpr int SMEM = 2 * (Big::STAGE > Small::STAGE ? Big::STAGE : Small::STAGE);
    __shared__ T smem[SMEM];
    __shared__ int sink[mode_c_pf<T>(MODE) ? 256 : 1];
    __shared__ int shared_idx[2];
    const int x = blockIdx.x % NXCD;
    const int beg = seg.off[x], end = seg.off[x + 1];
    if (threadIdx.x == 0) {
        shared_idx[0] = beg + atomicAdd(&ctr[x], 1);
        shared_idx[1] = beg + atomicAdd(&ctr[x], 1);
    }
    __syncthreads();
    // Uniform across the workgroup: keep the indices (and so the item and
    // descriptor loads) scalar.
    int cur = __builtin_amdgcn_readfirstlane(shared_idx[0]);
    int nxt = __builtin_amdgcn_readfirstlane(shared_idx[1]);
    WorkItem w{};
    BatchDesc d{};
    if (cur < end) {
        w = items[cur];
        d = descs[w.prob];
    }
    while (cur < end) {
        int nn = end;
        if (threadIdx.x == 0 && nxt < end) nn = beg + atomicAdd(&ctr[x], 1);
        WorkItem wn{};
        BatchDesc dn{};
        if (nxt < end) {
            wn = items[nxt];
            dn = descs[wn.prob];
        }
        TileProblem<T> p;
        p.a = (const T*)d.a; p.b = (const T*)d.b; p.c = (T*)d.c;
        p.m = d.m; p.n = d.n; p.k = d.k;
        p.lda = d.lda; p.ldb = d.ldb; p.c_si = d.c_si; p.c_sj = d.c_sj;
        p.alpha = (T)d.alpha; p.beta = (T)d.beta;
        p.vec_a = (d.flags & F_VEC_A) != 0;
        p.vec_b = (d.flags & F_VEC_B) != 0;
        p.vec_c = (d.flags & F_VEC_C) != 0;
        const int64_t tm = w.tile % d.tiles_m, tn = w.tile / d.tiles_m;
        constexpr int PF = mode_pf(MODE);
        constexpr bool EG = mode_edge(MODE), LE = mode_lds_epi<T>(MODE), CP = mode_c_pf<T>(MODE);
        switch (variant_of(d.flags)) {
        case 0: gemm_tile<T, Big, false, false, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        case 1: gemm_tile<T, Big, false, true, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        case 2: gemm_tile<T, Big, true, false, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        case 3: gemm_tile<T, Big, true, true, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        case 4: gemm_tile<T, Small, false, false, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        case 5: gemm_tile<T, Small, false, true, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        case 6: gemm_tile<T, Small, true, false, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        default: gemm_tile<T, Small, true, true, PF, EG, LE, CP>(p, tm, tn, smem, sink); break;
        }
        if (threadIdx.x == 0) shared_idx[0] = nn;
        __syncthreads();   // smem free for the next tile; nn visible
        cur = nxt;
        w = wn;
        d = dn;
        nxt = __builtin_amdgcn_readfirstlane(shared_idx[0]);
        __syncthreads();   // everyone has read shared_idx before it is rewritten
    }
    // All of this workgroup's counter fetches are done. The last workgroup
    // to get here resets the counters (vector atomics).
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctr[NXCD], 1) == (int)gridDim.x - 1) {
            for (int i = 0; i < NXCD; ++i) atomicExch(&ctr[i], 0);
            atomicExch(&ctr[NXCD], 0);
        }
    }
}

// Wave-tile kernel (mode 4): one wave computes one 32 x 32 tile of C as 2 x 2
// MFMA blocks of 16 x 16, loading its MFMA operand fragments straight from
// global memory (no LDS, no barriers), one k-step ahead. For the small
// problems of a batch this wastes less work on padding than 64 x 64 tiles and
// needs few registers, so many waves per SIMD hide the memory latency.
// Lane l supplies A(row 16i + l%16, k 4s + l/16) and B(k 4s + l/16,
// column 16j + l%16) to k-step s: the same MFMA sequence per output element
// as the LDS-staged tiles, hence bit-identical results.
template <typename T, bool A_K, bool B_K>
__global__ __launch_bounds__(256) void gemm_batch_wave_kernel(const BatchDesc* __restrict__ descs,
                                                              const WorkItem* __restrict__ items,
                                                              XcdSegs seg) {
    using acc_t = typename Mma<T>::acc_t;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = blockIdx.x % NXCD;
    const int slot = (blockIdx.x / NXCD) * 4 + wave, nslot = (gridDim.x / NXCD) * 4;
    const int lr = lane & 15, lk = lane >> 4;
    for (int it = seg.off[x] + slot; it < seg.off[x + 1]; it += nslot) {
        const WorkItem w = items[it];
        const BatchDesc& d = descs[w.prob];
        const T* __restrict__ a = (const T*)d.a;
        const T* __restrict__ b = (const T*)d.b;
        const int64_t m = d.m, n = d.n, K = d.k, lda = d.lda, ldb = d.ldb;
        const int64_t m0 = (int64_t)(w.tile % d.tiles_m32) * 32, n0 = (int64_t)(w.tile / d.tiles_m32) * 32;
        // Rows / columns this lane feeds, clamped into the matrix.
        int64_t ar[2], bc[2];
        bool aok[2], bok[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t r = m0 + 16 * i + lr, c = n0 + 16 * i + lr;
            aok[i] = r < m;
            bok[i] = c < n;
            ar[i] = (aok[i] ? r : m - 1) * (A_K ? lda : 1);
            bc[i] = (bok[i] ? c : n - 1) * (B_K ? ldb : 1);
        }
        acc_t acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{};
        const int64_t nks = (K + 3) / 4;
        auto frag = [&](int64_t s, T (&fa)[2], T (&fb)[2]) {
            const int64_t kk = 4 * s + lk;
            const bool kin = kk < K;
            const int64_t kc = kin ? kk : K - 1;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const T va = a[ar[i] + kc * (A_K ? 1 : lda)];
                const T vb = b[bc[i] + kc * (B_K ? 1 : ldb)];
                fa[i] = (kin && aok[i]) ? va : T(0);
                fb[i] = (kin && bok[i]) ? vb : T(0);
            }
        };
        if (nks > 0) {
            T fa[2], fb[2], ga[2], gb[2];
            frag(0, fa, fb);
            for (int64_t s = 0; s < nks; s += 2) {
                if (s + 1 < nks) frag(s + 1, ga, gb);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = Mma<T>::mma(fa[i], fb[j], acc[i][j]);
                if (s + 1 < nks) {
                    if (s + 2 < nks) frag(s + 2, fa, fb);
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j) acc[i][j] = Mma<T>::mma(ga[i], gb[j], acc[i][j]);
                }
            }
        }
        // Store: register r of block (i, j) holds row 16i + acc_row(r, lk),
        // column 16j + lr.
        T* c = (T*)d.c;
        const T alpha = (T)d.alpha, beta = (T)d.beta;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t col = n0 + 16 * j + lr;
            if (col >= n) continue;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t row = m0 + 16 * i + Mma<T>::acc_row(r, lk);
                    if (row >= m) continue;
                    T* dst = c + row * d.c_si + col * d.c_sj;
                    *dst = (beta == T(0)) ? alpha * acc[i][j][r]
                                          : epilogue(alpha, acc[i][j][r], beta, *dst);
                }
        }
    }
}

template <typename T, int MODE>
inline const void* variant_fn(int v) {
    using Big = typename std::conditional<(MODE >= 7), typename BatchTiles<T>::BigHalfK,
                                          typename BatchTiles<T>::Big>::type;
    using Small = typename BatchTiles<T>::Small;
    switch (v) {
    case 0: return (const void*)&gemm_batch_kernel<T, Big, false, false, MODE>;
    case 1: return (const void*)&gemm_batch_kernel<T, Big, false, true, MODE>;
    case 2: return (const void*)&gemm_batch_kernel<T, Big, true, false, MODE>;
    case 3: return (const void*)&gemm_batch_kernel<T, Big, true, true, MODE>;
    case 4: return (const void*)&gemm_batch_kernel<T, Small, false, false, MODE>;
    case 5: return (const void*)&gemm_batch_kernel<T, Small, false, true, MODE>;
    case 6: return (const void*)&gemm_batch_kernel<T, Small, true, false, MODE>;
    case 7: return (const void*)&gemm_batch_kernel<T, Small, true, true, MODE>;
    }
    abort_with(BLIS_INVALID_UKR_ID, "batch: variant");
    return nullptr;
}

template <typename T>
inline void variant_info(int mode, int v, const void*& fn, int& threads) {
    threads = (v < 4) ? BatchTiles<T>::Big::NT : BatchTiles<T>::Small::NT;
    if (mode == WAVE_MODE) {
        // v = 2*a_k + b_k
        threads = 256;
        switch (v) {
        case 0: fn = (const void*)&gemm_batch_wave_kernel<T, false, false>; return;
        case 1: fn = (const void*)&gemm_batch_wave_kernel<T, false, true>; return;
        case 2: fn = (const void*)&gemm_batch_wave_kernel<T, true, false>; return;
        case 3: fn = (const void*)&gemm_batch_wave_kernel<T, true, true>; return;
        }
        abort_with(BLIS_INVALID_UKR_ID, "batch: wave variant");
    }
    switch (mode) {
    case 0: fn = variant_fn<T, 0>(v); return;
    case 1: fn = variant_fn<T, 1>(v); return;
    case 2: fn = variant_fn<T, 2>(v); return;
    case 3: fn = variant_fn<T, 3>(v); return;
    case 5: fn = variant_fn<T, 5>(v); return;
    case 6: fn = variant_fn<T, 6>(v); return;
    case 7: fn = variant_fn<T, 7>(v); return;
    case 8: fn = variant_fn<T, 8>(v); return;
    }
    abort_with(BLIS_INVALID_UKR_ID, "batch: mode");
}

template <typename T>
inline const void* dyn_fn(int mode) {
    switch (mode) {
    case 0: return (const void*)&gemm_batch_dyn_kernel<T, 0>;
    case 1: return (const void*)&gemm_batch_dyn_kernel<T, 1>;
    case 2: return (const void*)&gemm_batch_dyn_kernel<T, 2>;
    case 3: return (const void*)&gemm_batch_dyn_kernel<T, 3>;
    case 5: return (const void*)&gemm_batch_dyn_kernel<T, 5>;
    case 6: return (const void*)&gemm_batch_dyn_kernel<T, 6>;
    case 7: return (const void*)&gemm_batch_dyn_kernel<T, 7>;
    case 8: return (const void*)&gemm_batch_dyn_kernel<T, 8>;
    }
    abort_with(BLIS_INVALID_UKR_ID, "batch: single-dispatch mode");
    return nullptr;
}

// C of every problem to / from a contiguous scratch copy (the autotuner
// runs the batch several times and restores C afterwards).
template <typename T>
__global__ void batch_c_copy_kernel(const BatchDesc* __restrict__ descs, int ndesc,
                                    const int64_t* __restrict__ offs, T* scratch, int to_scratch) {
    for (int i = blockIdx.x; i < ndesc; i += gridDim.x) {
        const BatchDesc& d = descs[i];
        T* c = (T*)d.c;
        T* s = scratch + offs[i];
        const int64_t cnt = (int64_t)d.m * d.n;
        for (int64_t e = threadIdx.x; e < cnt; e += blockDim.x) {
            const int64_t r = e % d.m, col = e / d.m;
            T* cp = c + r * d.c_si + col * d.c_sj;
            if (to_scratch) s[e] = *cp;
            else *cp = s[e];
        }
    }
}

inline bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int cu_count() {
    static int n = 0;
    if (n == 0) {
        int dev = device_id();
        HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        if (n <= 0) n = 256;
    }
    return n;
}

} // namespace detail_batch
using namespace detail_batch;

// Device-resident batch ready to launch (also the plan object of the
// bli_amd_?gemm_batch_plan_* extension).
struct BatchPlan {
    Dt dt;
    BatchDesc* d_descs = nullptr;
    WorkItem* d_items = nullptr;
    int nitems = 0;
    XcdSegs seg[NVARIANT] = {};       // per variant: absolute item offsets per XCD list
    int grid[NMODE][NVARIANT] = {};   // persistent grid per kernel mode and variant
    WorkItem* d_witems = nullptr;     // wave-tile mode: 32 x 32 tiles of every problem
    XcdSegs wseg[4] = {};             // per layout variant (2*a_k + b_k)
    int mode = 0;                     // kernel mode used by execute
    int vmode[NVARIANT] = {};         // per-variant kernel mode (mode != WAVE_MODE)
    // Single dispatch (gemm_batch_dyn_kernel): every variant's items in one
    // list per XCD, the workgroup count per mode, the per-XCD counters. Off
    // until the autotuner picks it: on the 10,000-problem batch (dims
    // log-uniform in [8, 512]) the kernel holding all eight tile paths
    // needs more registers (FP64 mode 8: 320 bytes per lane of spills
    // against 20-64 for the per-variant kernels) and measured slower
    // (d 1.83 ms at best against 1.74 ms, s 1.24 against 1.03 ms).
    bool dyn = false;
    WorkItem* d_ditems = nullptr;
    XcdSegs dseg = {};
    int dgrid[NMODE] = {};
    int* d_ctr = nullptr;
    int ndesc = 0;
    int64_t c_elems = 0;              // sum of m*n (autotuner scratch)
    std::vector<int64_t> c_offs;
    std::vector<BatchProblem> fallback;   // general-stride problems (plain GEMMs)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    ~BatchPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_descs) (void)hipFree(d_descs);
        if (d_items) (void)hipFree(d_items);
        if (d_witems) (void)hipFree(d_witems);
        if (d_ditems) (void)hipFree(d_ditems);
        if (d_ctr) (void)hipFree(d_ctr);
    }
};

static BatchPlan* plan_build(Dt dt, const std::vector<BatchProblem>& probs, hipStream_t st) {
    const int V = dt == Dt::F64 ? 2 : 4;
    auto* plan = new BatchPlan();
    plan->dt = dt;
    std::vector<BatchDesc> descs;
    descs.reserve(probs.size());
    std::vector<WorkItem> items;
    std::vector<int32_t> item_k;
    for (const BatchProblem& q0 : probs) {
        BatchProblem q = q0;
        if (q.m <= 0 || q.n <= 0) continue;
        // Row-stored C: solve C^T = B^T A^T (see gemm.hip).
        if (q.c.rs != 1 && q.c.cs == 1) {
            BatchProblem t = q;
            t.m = q.n; t.n = q.m;
            t.a = { q.b.ptr, q.b.cs, q.b.rs };
            t.b = { q.a.ptr, q.a.cs, q.a.rs };
            t.c = { q.c.ptr, q.c.cs, q.c.rs };
            q = t;
        }
        BatchDesc d{};
        int32_t flags = 0;
        bool general = false;
        if (q.k > 0) {
            if (q.a.rs == 1) { d.lda = q.a.cs; }
            else if (q.a.cs == 1) { d.lda = q.a.rs; flags |= F_A_K; }
            else general = true;
            if (q.b.rs == 1) { d.ldb = q.b.cs; flags |= F_B_K; }
            else if (q.b.cs == 1) { d.ldb = q.b.rs; }
            else general = true;
        } else {
            d.lda = 1; d.ldb = 1;
        }
        if (general || q.m > INT32_MAX || q.n > INT32_MAX || q.k > INT32_MAX) {
            plan->fallback.push_back(q0);
            continue;
        }
        d.a = q.a.ptr; d.b = q.b.ptr; d.c = q.c.ptr;
        d.c_si = q.c.rs; d.c_sj = q.c.cs;
        d.m = (int32_t)q.m; d.n = (int32_t)q.n; d.k = (int32_t)q.k;
        d.alpha = q.alpha; d.beta = q.beta;
        if (al16(d.a) && d.lda % V == 0) flags |= F_VEC_A;
        if (al16(d.b) && d.ldb % V == 0) flags |= F_VEC_B;
        if (al16(d.c) && d.c_si == 1 && d.c_sj % V == 0) flags |= F_VEC_C;
        // 32 x 32 tiles where they cover the problem at lower cost: a 32 x 32
        // tile was measured at about 0.4 of a 64 x 64 tile's time at equal k
        // (both latency-bound at these sizes), so it wins where it cuts the
        // padded area by more than that.
        const bool small = (double)((q.m + 31) / 32) * (double)((q.n + 31) / 32) * 0.4 <
                           (double)((q.m + 63) / 64) * (double)((q.n + 63) / 64);
        if (small) flags |= F_SMALL;
        const int tb = small ? 32 : 64;
        d.flags = flags;
        d.tiles_m = (int32_t)((q.m + tb - 1) / tb);
        d.tiles_m32 = (int32_t)((q.m + 31) / 32);
        const int64_t tiles_n = (q.n + tb - 1) / tb;
        const int32_t pi = (int32_t)descs.size();
        descs.push_back(d);
        for (int64_t t = 0; t < (int64_t)d.tiles_m * tiles_n; ++t) {
            items.push_back(WorkItem{ pi, (int32_t)t });
            item_k.push_back(d.k);
        }
    }
    // Items are grouped by kernel variant, then by XCD list, then
    // longest-first by k (ties keep submission order). Problems are assigned
    // to XCD lists greedily, largest first, to the least-loaded list of their
    // variant (work = tile area x k summed over the problem's tiles).
    const size_t ni = items.size();
    std::vector<int32_t> item_v(ni), prob_x(descs.size(), 0);
    for (size_t i = 0; i < ni; ++i) item_v[i] = variant_of(descs[items[i].prob].flags);
    {
        std::vector<double> pwork(descs.size(), 0.0);
        for (size_t i = 0; i < ni; ++i) {
            const BatchDesc& d = descs[items[i].prob];
            const double area = (d.flags & F_SMALL) ? 32.0 * 32.0 : 64.0 * 64.0;
            pwork[items[i].prob] += area * std::max<int32_t>(d.k, 1);
        }
        std::vector<int32_t> porder(descs.size());
        std::iota(porder.begin(), porder.end(), 0);
        std::stable_sort(porder.begin(), porder.end(),
                         [&](int32_t x, int32_t y) { return pwork[x] > pwork[y]; });
        std::vector<double> load((size_t)NVARIANT * NXCD, 0.0);
        for (int32_t pi : porder) {
            if (pwork[pi] == 0.0) continue;
            const int v = variant_of(descs[pi].flags);
            int best = 0;
            for (int x = 1; x < NXCD; ++x)
                if (load[v * NXCD + x] < load[v * NXCD + best]) best = x;
            prob_x[pi] = best;
            load[v * NXCD + best] += pwork[pi];
        }
    }
    std::vector<int32_t> order(ni);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        if (item_v[x] != item_v[y]) return item_v[x] < item_v[y];
        const int32_t px = prob_x[items[x].prob], py = prob_x[items[y].prob];
        if (px != py) return px < py;
        return item_k[x] > item_k[y];
    });
    std::vector<int> count((size_t)NVARIANT * NXCD, 0);
    for (size_t i = 0; i < ni; ++i) count[item_v[i] * NXCD + prob_x[items[i].prob]]++;
    int off = 0;
    for (int v = 0; v < NVARIANT; ++v)
        for (int x = 0; x <= NXCD; ++x) {
            plan->seg[v].off[x] = off;
            if (x < NXCD) off += count[v * NXCD + x];
        }
    std::vector<WorkItem> sorted(ni);
    for (size_t i = 0; i < ni; ++i) sorted[i] = items[order[i]];
    plan->nitems = (int)sorted.size();
    if (!descs.empty()) {
        HIP_CHECK(hipMalloc(&plan->d_descs, descs.size() * sizeof(BatchDesc)));
        HIP_CHECK(hipMemcpyAsync(plan->d_descs, descs.data(), descs.size() * sizeof(BatchDesc),
                                 hipMemcpyHostToDevice, st));
    }
    if (!sorted.empty()) {
        HIP_CHECK(hipMalloc(&plan->d_items, sorted.size() * sizeof(WorkItem)));
        HIP_CHECK(hipMemcpyAsync(plan->d_items, sorted.data(), sorted.size() * sizeof(WorkItem),
                                 hipMemcpyHostToDevice, st));
    }
    // Wave-tile items: every 32 x 32 tile of every problem, grouped by layout
    // variant, then XCD list (same problem placement as above), then
    // longest-first by k.
    {
        std::vector<WorkItem> witems;
        std::vector<int32_t> wk, wv, wx;
        for (size_t pi = 0; pi < descs.size(); ++pi) {
            const BatchDesc& d = descs[pi];
            const int64_t t = (int64_t)d.tiles_m32 * ((d.n + 31) / 32);
            const int lv = variant_of(d.flags) & 3;
            for (int64_t ti = 0; ti < t; ++ti) {
                witems.push_back(WorkItem{(int32_t)pi, (int32_t)ti});
                wk.push_back(d.k);
                wv.push_back(lv);
                wx.push_back(prob_x[pi]);
            }
        }
        std::vector<int32_t> word(witems.size());
        std::iota(word.begin(), word.end(), 0);
        std::stable_sort(word.begin(), word.end(), [&](int32_t x, int32_t y) {
            if (wv[x] != wv[y]) return wv[x] < wv[y];
            if (wx[x] != wx[y]) return wx[x] < wx[y];
            return wk[x] > wk[y];
        });
        std::vector<int> wcnt(4 * NXCD, 0);
        for (size_t i = 0; i < witems.size(); ++i) wcnt[wv[i] * NXCD + wx[i]]++;
        int woff = 0;
        for (int v = 0; v < 4; ++v)
            for (int x = 0; x <= NXCD; ++x) {
                plan->wseg[v].off[x] = woff;
                if (x < NXCD) woff += wcnt[v * NXCD + x];
            }
        std::vector<WorkItem> wsorted(witems.size());
        for (size_t i = 0; i < witems.size(); ++i) wsorted[i] = witems[word[i]];
        if (!wsorted.empty()) {
            HIP_CHECK(hipMalloc(&plan->d_witems, wsorted.size() * sizeof(WorkItem)));
            HIP_CHECK(hipMemcpyAsync(plan->d_witems, wsorted.data(),
                                     wsorted.size() * sizeof(WorkItem), hipMemcpyHostToDevice, st));
        }
    }
    // Single-dispatch lists: the items of every variant, placed on the XCDs
    // by problem (greedy, largest first, on an estimate of each tile's time:
    // k slabs of 16 plus two for the epilogue, a 32 x 32 tile at 0.4 of a
    // 64 x 64 one), then longest tile first within each XCD (ties keep the
    // submission order, so a problem's tiles stay together).
    std::vector<int32_t> dx(descs.size(), 0);
    {
        auto tile_cost = [&](const BatchDesc& d) {
            return ((d.flags & F_SMALL) ? 0.4 : 1.0) * (double)((std::max<int32_t>(d.k, 0) + 15) / 16 + 2);
        };
        std::vector<double> pwork(descs.size(), 0.0);
        for (size_t i = 0; i < ni; ++i) pwork[items[i].prob] += tile_cost(descs[items[i].prob]);
        std::vector<int32_t> porder(descs.size());
        std::iota(porder.begin(), porder.end(), 0);
        std::stable_sort(porder.begin(), porder.end(),
                         [&](int32_t x, int32_t y) { return pwork[x] > pwork[y]; });
        double load[NXCD] = {};
        for (int32_t pi : porder) {
            int best = 0;
            for (int x = 1; x < NXCD; ++x)
                if (load[x] < load[best]) best = x;
            dx[pi] = best;
            load[best] += pwork[pi];
        }
        std::vector<int32_t> dord(ni);
        std::iota(dord.begin(), dord.end(), 0);
        std::stable_sort(dord.begin(), dord.end(), [&](int32_t x, int32_t y) {
            const int32_t px = dx[items[x].prob], py = dx[items[y].prob];
            if (px != py) return px < py;
            return tile_cost(descs[items[x].prob]) > tile_cost(descs[items[y].prob]);
        });
        int dcnt[NXCD] = {};
        for (size_t i = 0; i < ni; ++i) dcnt[dx[items[i].prob]]++;
        int doff = 0;
        for (int x = 0; x <= NXCD; ++x) {
            plan->dseg.off[x] = doff;
            if (x < NXCD) doff += dcnt[x];
        }
        if (ni > 0) {
            std::vector<WorkItem> dsorted(ni);
            for (size_t i = 0; i < ni; ++i) dsorted[i] = items[dord[i]];
            HIP_CHECK(hipMalloc(&plan->d_ditems, ni * sizeof(WorkItem)));
            HIP_CHECK(hipMemcpyAsync(plan->d_ditems, dsorted.data(), ni * sizeof(WorkItem),
                                     hipMemcpyHostToDevice, st));
        }
        static const int zeros[NXCD + 1] = {};
        HIP_CHECK(hipMalloc(&plan->d_ctr, sizeof(zeros)));
        HIP_CHECK(hipMemcpyAsync(plan->d_ctr, zeros, sizeof(zeros), hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
    // Single dispatch: one resident wave of workgroups, a multiple of the
    // XCD count, and no more per XCD than its longest list.
    {
        int longest = 0;
        for (int x = 0; x < NXCD; ++x)
            longest = std::max(longest, plan->dseg.off[x + 1] - plan->dseg.off[x]);
        for (int mode = 0; mode < NMODE; ++mode) {
            if (mode == WAVE_MODE || longest == 0) continue;
            const void* fn = dt == Dt::F64 ? dyn_fn<double>(mode) : dyn_fn<float>(mode);
            int per_cu = 0;
            HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0));
            per_cu = std::max(1, per_cu);
            plan->dgrid[mode] = std::max(1, std::min(longest, cu_count() * per_cu / NXCD)) * NXCD;
        }
    }
    // One resident wave of workgroups per variant and mode (persistent
    // kernels), a multiple of the XCD count, and no more per XCD than its
    // longest list.
    for (int mode = 0; mode < NMODE; ++mode)
        for (int v = 0; v < NVARIANT; ++v) {
            const bool wavem = mode == WAVE_MODE;
            if (wavem && v >= 4) continue;
            const XcdSegs& sg = wavem ? plan->wseg[v] : plan->seg[v];
            int longest = 0;
            for (int x = 0; x < NXCD; ++x) longest = std::max(longest, sg.off[x + 1] - sg.off[x]);
            if (longest == 0) continue;
            if (wavem) longest = (longest + 3) / 4;     // four waves per workgroup
            const void* fn;
            int threads;
            if (dt == Dt::F64) variant_info<double>(mode, v, fn, threads);
            else variant_info<float>(mode, v, fn, threads);
            int per_cu = 0;
            HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
            per_cu = std::max(1, per_cu);
            const int slots = std::max(1, std::min(longest, cu_count() * per_cu / NXCD));
            plan->grid[mode][v] = slots * NXCD;
        }
    plan->mode = default_mode(dt);
    for (int v = 0; v < NVARIANT; ++v) plan->vmode[v] = plan->mode;
    plan->ndesc = (int)descs.size();
    plan->c_offs.resize(descs.size());
    for (size_t i = 0; i < descs.size(); ++i) {
        plan->c_offs[i] = plan->c_elems;
        plan->c_elems += (int64_t)descs[i].m * descs[i].n;
    }
    return plan;
}

// Launch variant v of the plan in kernel mode `mode` (nothing if the variant
// has no work in that mode).
static void launch_variant(const BatchPlan* plan, int mode, int v, hipStream_t st) {
    if (plan->grid[mode][v] == 0) return;
    const void* fn;
    int threads;
    if (plan->dt == Dt::F64) variant_info<double>(mode, v, fn, threads);
    else variant_info<float>(mode, v, fn, threads);
    const BatchDesc* descs = plan->d_descs;
    const WorkItem* items = (mode == WAVE_MODE) ? plan->d_witems : plan->d_items;
    XcdSegs seg = (mode == WAVE_MODE) ? plan->wseg[v] : plan->seg[v];
    void* args[] = {(void*)&descs, (void*)&items, (void*)&seg};
    HIP_CHECK(hipLaunchKernel(fn, dim3(plan->grid[mode][v]), dim3(threads), args, 0, st));
}

// The single dispatch in kernel mode `mode` (not the wave mode).
static void launch_dyn(const BatchPlan* plan, int mode, hipStream_t st) {
    if (plan->dgrid[mode] == 0) return;
    const void* fn = plan->dt == Dt::F64 ? dyn_fn<double>(mode) : dyn_fn<float>(mode);
    const BatchDesc* descs = plan->d_descs;
    const WorkItem* items = plan->d_ditems;
    XcdSegs seg = plan->dseg;
    int* ctr = plan->d_ctr;
    void* args[] = {(void*)&descs, (void*)&items, (void*)&seg, (void*)&ctr};
    HIP_CHECK(hipLaunchKernel(fn, dim3(plan->dgrid[mode]), dim3(256), args, 0, st));
}

static void plan_launch(const BatchPlan* plan, hipStream_t st, bool with_fallback = true) {
    if (plan->dyn && plan->mode != WAVE_MODE)
        launch_dyn(plan, plan->mode, st);
    else
        for (int v = 0; v < NVARIANT; ++v)
            launch_variant(plan, plan->mode == WAVE_MODE ? WAVE_MODE : plan->vmode[v], v, st);
    if (!with_fallback) return;
    for (const BatchProblem& q : plan->fallback) {
        GemmDesc d;
        d.dt = plan->dt; d.m = q.m; d.n = q.n; d.k = q.k;
        d.a = q.a; d.b = q.b; d.c = q.c; d.alpha = q.alpha; d.beta = q.beta;
        gemm_run(d, nullptr, st);
    }
}

namespace detail_batch {

bool same_view(const View& x, const View& y) {
    return x.ptr == y.ptr && x.rs == y.rs && x.cs == y.cs;
}
bool same_problem(const BatchProblem& x, const BatchProblem& y) {
    return x.m == y.m && x.n == y.n && x.k == y.k && same_view(x.a, y.a) &&
           same_view(x.b, y.b) && same_view(x.c, y.c) && x.alpha == y.alpha && x.beta == y.beta;
}
uint64_t batch_hash(Dt dt, const std::vector<BatchProblem>& probs) {
    uint64_t h = 1469598103934665603ull ^ (uint64_t)dt;
    auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
    for (const BatchProblem& q : probs) {
        mix((uint64_t)q.m); mix((uint64_t)q.n); mix((uint64_t)q.k);
        mix((uint64_t)(uintptr_t)q.a.ptr); mix((uint64_t)q.a.rs); mix((uint64_t)q.a.cs);
        mix((uint64_t)(uintptr_t)q.b.ptr); mix((uint64_t)q.b.rs); mix((uint64_t)q.b.cs);
        mix((uint64_t)(uintptr_t)q.c.ptr); mix((uint64_t)q.c.rs); mix((uint64_t)q.c.cs);
        uint64_t u;
        std::memcpy(&u, &q.alpha, 8); mix(u);
        std::memcpy(&u, &q.beta, 8); mix(u);
    }
    return h;
}

// Plans of recent drop-in calls. A plan is a pure function of the call's
// arguments (dimensions, pointers, strides, scalars), so a call whose
// arguments equal an earlier one's reuses its device descriptors, work
// lists and captured graph instead of rebuilding them. A plan reused for
// the second time is autotuned first (all kernel modes timed on the plan
// itself, C saved and restored), unless BLIS_AMD_BATCH_AUTOTUNE=0.
struct CachedPlan {
    Dt dt;
    uint64_t hash;
    std::vector<BatchProblem> probs;
    BatchPlan* plan;
    int uses;
    uint64_t last;
};
struct PlanCache {
    static constexpr size_t CAP = 4;
    std::vector<CachedPlan> e;
    uint64_t clock = 0;
};
// Never destroyed: the plans' device memory and graphs are released with
// the process, not by static destructors that may run after the HIP
// runtime has shut down.
PlanCache& plan_cache() {
    static PlanCache* pc = new PlanCache();
    return *pc;
}
bool batch_autotune_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("BLIS_AMD_BATCH_AUTOTUNE");
        return !(v && v[0] == '0');
    }();
    return on;
}

} // namespace detail_batch

static void plan_tune(BatchPlan* plan, int reps);

// The drop-in batch (bli_?gemm_batch, ?gemm_batch_): launch a cached plan
// when one matches the call, else build one (and cache it). Returns when the
// batch is done (the caller synchronises as BLIS semantics require).
void gemm_batch_run(Dt dt, const std::vector<BatchProblem>& probs, hipStream_t st) {
    PlanCache& pc = plan_cache();
    const uint64_t h = batch_hash(dt, probs);
    ++pc.clock;
    for (CachedPlan& c : pc.e) {
        if (c.dt != dt || c.hash != h || c.probs.size() != probs.size()) continue;
        bool same = true;
        for (size_t i = 0; i < probs.size() && same; ++i) same = same_problem(c.probs[i], probs[i]);
        if (!same) continue;
        c.last = pc.clock;
        ++c.uses;
        BatchPlan* plan = c.plan;
        if (c.uses == 2 && batch_autotune_enabled()) plan_tune(plan, 3);
        if (plan->fallback.empty()) {
            if (!plan->exec) {
                HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                plan_launch(plan, st);
                HIP_CHECK(hipStreamEndCapture(st, &plan->graph));
                HIP_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
            }
            HIP_CHECK(hipGraphLaunch(plan->exec, st));
        } else {
            plan_launch(plan, st);
        }
        HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    BatchPlan* plan = plan_build(dt, probs, st);
    plan_launch(plan, st);
    HIP_CHECK(hipStreamSynchronize(st));
    if (pc.e.size() >= PlanCache::CAP) {
        size_t old = 0;
        for (size_t i = 1; i < pc.e.size(); ++i)
            if (pc.e[i].last < pc.e[old].last) old = i;
        delete pc.e[old].plan;
        pc.e.erase(pc.e.begin() + (long)old);
    }
    pc.e.push_back(CachedPlan{dt, h, probs, plan, 1, pc.clock});
}

} // namespace blis_amd

using namespace blis_amd;

// ---- extension: reusable batch plans with hipGraph replay --------------------

template <typename T>
static void* plan_create_t(const trans_t* transa, const trans_t* transb, const dim_t* m,
                           const dim_t* n, const dim_t* k, const T* alpha,
                           const T* const* a, const inc_t* rs_a, const inc_t* cs_a,
                           const T* const* b, const inc_t* rs_b, const inc_t* cs_b,
                           const T* beta, T* const* c, const inc_t* rs_c, const inc_t* cs_c,
                           dim_t group_count, const dim_t* group_size) {
    init_once();
    std::vector<BatchProblem> probs;
    dim_t idx = 0;
    for (dim_t g = 0; g < group_count; ++g) {
        const bool ta = ((unsigned)transa[g] & BLIS_TRANS_BIT) != 0;
        const bool tb = ((unsigned)transb[g] & BLIS_TRANS_BIT) != 0;
        for (dim_t s = 0; s < group_size[g]; ++s, ++idx) {
            BatchProblem p;
            p.m = m[g]; p.n = n[g]; p.k = k[g];
            p.alpha = (double)alpha[g]; p.beta = (double)beta[g];
            if (p.alpha == 0.0) p.k = 0;
            p.a = { (void*)a[idx], ta ? cs_a[g] : rs_a[g], ta ? rs_a[g] : cs_a[g] };
            p.b = { (void*)b[idx], tb ? cs_b[g] : rs_b[g], tb ? rs_b[g] : cs_b[g] };
            p.c = { (void*)c[idx], rs_c[g], cs_c[g] };
            probs.push_back(p);
        }
    }
    std::lock_guard<std::mutex> lock(api_mutex());
    return plan_build(std::is_same<T, double>::value ? Dt::F64 : Dt::F32, probs, stream());
}

extern "C" {

// Build a device-resident plan for a batch whose operands are already in
// device memory (same argument meaning as bli_?gemm_batch).
__attribute__((visibility("default")))
void* bli_amd_dgemm_batch_plan_create(const trans_t* transa, const trans_t* transb,
                                      const dim_t* m, const dim_t* n, const dim_t* k,
                                      const double* alpha, const double* const* a,
                                      const inc_t* rs_a, const inc_t* cs_a,
                                      const double* const* b, const inc_t* rs_b,
                                      const inc_t* cs_b, const double* beta,
                                      double* const* c, const inc_t* rs_c, const inc_t* cs_c,
                                      dim_t group_count, const dim_t* group_size) {
    return plan_create_t<double>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                                 beta, c, rs_c, cs_c, group_count, group_size);
}

__attribute__((visibility("default")))
void* bli_amd_sgemm_batch_plan_create(const trans_t* transa, const trans_t* transb,
                                      const dim_t* m, const dim_t* n, const dim_t* k,
                                      const float* alpha, const float* const* a,
                                      const inc_t* rs_a, const inc_t* cs_a,
                                      const float* const* b, const inc_t* rs_b,
                                      const inc_t* cs_b, const float* beta,
                                      float* const* c, const inc_t* rs_c, const inc_t* cs_c,
                                      dim_t group_count, const dim_t* group_size) {
    return plan_create_t<float>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                                beta, c, rs_c, cs_c, group_count, group_size);
}

// Execute a plan. use_graph != 0 replays the launch as a HIP graph that is
// captured on first use (the plan's problems and pointers are fixed).
__attribute__((visibility("default")))
void bli_amd_gemm_batch_plan_execute(void* handle, int use_graph) {
    auto* plan = (BatchPlan*)handle;
    std::lock_guard<std::mutex> lock(api_mutex());
    hipStream_t st = stream();
    if (use_graph && plan->fallback.empty()) {
        if (!plan->exec) {
            HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            plan_launch(plan, st);
            HIP_CHECK(hipStreamEndCapture(st, &plan->graph));
            HIP_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
        }
        HIP_CHECK(hipGraphLaunch(plan->exec, st));
    } else {
        plan_launch(plan, st);
    }
    sync_if_needed();
}

// Choose the plan's kernel mode by timing every mode on the plan itself.
// C is saved before and restored after, so the call leaves the caller's
// data unchanged. Returns the chosen mode.
__attribute__((visibility("default")))
int bli_amd_gemm_batch_plan_autotune(void* handle, int reps) {
    auto* plan = (BatchPlan*)handle;
    std::lock_guard<std::mutex> lock(api_mutex());
    plan_tune(plan, reps);
    return plan->mode;
}

} // extern "C"

namespace blis_amd {

// The batch autotuner (the caller holds the API mutex).
static void plan_tune(BatchPlan* plan, int reps) {
    if (plan->nitems == 0) return;
    hipStream_t st = stream();
    reps = std::max(1, reps);
    const size_t es = dt_size(plan->dt);
    void* scratch = workspace((size_t)std::max<int64_t>(plan->c_elems, 1) * es, 11);
    int64_t* d_offs = (int64_t*)workspace(sizeof(int64_t) * std::max<size_t>(plan->c_offs.size(), 1), 12);
    HIP_CHECK(hipMemcpyAsync(d_offs, plan->c_offs.data(), sizeof(int64_t) * plan->c_offs.size(),
                             hipMemcpyHostToDevice, st));
    auto copy_c = [&](int to_scratch) {
        const unsigned g = (unsigned)std::max(1, std::min(plan->ndesc, 4096));
        if (plan->dt == Dt::F64)
            hipLaunchKernelGGL(batch_c_copy_kernel<double>, dim3(g), dim3(256), 0, st, plan->d_descs,
                               plan->ndesc, d_offs, (double*)scratch, to_scratch);
        else
            hipLaunchKernelGGL(batch_c_copy_kernel<float>, dim3(g), dim3(256), 0, st, plan->d_descs,
                               plan->ndesc, d_offs, (float*)scratch, to_scratch);
        HIP_CHECK(hipGetLastError());
    };
    copy_c(1);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    // The general-stride problems run as plain GEMMs whatever the mode; they
    // are left out of the timing (and so are not re-applied).
    auto time_ms = [&](auto&& launch) {
        launch();                                   // warm-up (code load)
        HIP_CHECK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) launch();
        HIP_CHECK(hipEventRecord(e1, st));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return ms / reps;
    };
    // Every variant is its own kernel: time each one in each LDS-tile mode
    // and keep its fastest; the wave-tile mode covers all variants at once
    // and is timed as a whole.
    int vbest[NVARIANT];
    float vbest_ms[NVARIANT];
    float sum_ms = 0.f;
    for (int v = 0; v < NVARIANT; ++v) {
        vbest[v] = default_mode(plan->dt);   // never the wave mode (it spans variants)
        vbest_ms[v] = 0.f;
        for (int mode = 0; mode < NMODE; ++mode) {
            if (mode == WAVE_MODE || plan->grid[mode][v] == 0) continue;
            const float ms = time_ms([&] { launch_variant(plan, mode, v, st); });
            if (mode == 0 || ms < vbest_ms[v]) { vbest_ms[v] = ms; vbest[v] = mode; }
        }
        sum_ms += vbest_ms[v];
    }
    const float wave_ms = time_ms([&] {
        for (int v = 0; v < NVARIANT; ++v) launch_variant(plan, WAVE_MODE, v, st);
    });
    // The single dispatch in every LDS-tile mode.
    int dbest = default_mode(plan->dt);
    float dbest_ms = 0.f;
    for (int mode = 0; mode < NMODE; ++mode) {
        if (mode == WAVE_MODE || plan->dgrid[mode] == 0) continue;
        const float ms = time_ms([&] { launch_dyn(plan, mode, st); });
        if (dbest_ms == 0.f || ms < dbest_ms) { dbest_ms = ms; dbest = mode; }
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    copy_c(0);
    HIP_CHECK(hipStreamSynchronize(st));
    bool changed = false;
    const bool was_dyn = plan->dyn;
    plan->dyn = dbest_ms > 0.f && dbest_ms < sum_ms && dbest_ms < wave_ms;
    changed = was_dyn != plan->dyn;
    if (plan->dyn) {
        changed = changed || plan->mode != dbest;
        plan->mode = dbest;
        for (int v = 0; v < NVARIANT; ++v) plan->vmode[v] = dbest;
    } else if (wave_ms < sum_ms) {
        changed = changed || plan->mode != WAVE_MODE;
        plan->mode = WAVE_MODE;
        for (int v = 0; v < NVARIANT; ++v) plan->vmode[v] = WAVE_MODE;
    } else {
        // Reported mode: the one of the variant that takes longest.
        int top = 0;
        for (int v = 1; v < NVARIANT; ++v)
            if (vbest_ms[v] > vbest_ms[top]) top = v;
        changed = changed || plan->mode != vbest[top];
        plan->mode = vbest[top];
        for (int v = 0; v < NVARIANT; ++v) {
            changed = changed || plan->vmode[v] != vbest[v];
            plan->vmode[v] = vbest[v];
        }
    }
    if (changed && plan->exec) {            // captured graph belongs to the old modes
        HIP_CHECK(hipGraphExecDestroy(plan->exec));
        HIP_CHECK(hipGraphDestroy(plan->graph));
        plan->exec = nullptr;
        plan->graph = nullptr;
    }
}

} // namespace blis_amd

extern "C" {

__attribute__((visibility("default")))
void bli_amd_gemm_batch_plan_set_mode(void* handle, int mode) {
    auto* plan = (BatchPlan*)handle;
    std::lock_guard<std::mutex> lock(api_mutex());
    if (mode < 0 || mode >= NMODE) abort_with(BLIS_INVALID_UKR_ID, "batch plan mode");
    bool changed = mode != plan->mode;
    for (int v = 0; v < NVARIANT; ++v) changed = changed || plan->vmode[v] != mode;
    if (changed && plan->exec) {
        HIP_CHECK(hipGraphExecDestroy(plan->exec));
        HIP_CHECK(hipGraphDestroy(plan->graph));
        plan->exec = nullptr;
        plan->graph = nullptr;
    }
    plan->mode = mode;
    for (int v = 0; v < NVARIANT; ++v) plan->vmode[v] = mode;
}

// on != 0: run the plan as one dispatch (every mode but the wave mode);
// 0: one kernel per variant, one after another.
__attribute__((visibility("default")))
void bli_amd_gemm_batch_plan_set_single_dispatch(void* handle, int on) {
    auto* plan = (BatchPlan*)handle;
    std::lock_guard<std::mutex> lock(api_mutex());
    if ((on != 0) != plan->dyn && plan->exec) {
        HIP_CHECK(hipGraphExecDestroy(plan->exec));
        HIP_CHECK(hipGraphDestroy(plan->graph));
        plan->exec = nullptr;
        plan->graph = nullptr;
    }
    plan->dyn = on != 0;
}

__attribute__((visibility("default")))
int bli_amd_gemm_batch_plan_get_single_dispatch(void* handle) {
    return ((BatchPlan*)handle)->dyn ? 1 : 0;
}

__attribute__((visibility("default")))
int bli_amd_gemm_batch_plan_get_mode(void* handle) {
    return ((BatchPlan*)handle)->mode;
}

__attribute__((visibility("default")))
int bli_amd_gemm_batch_plan_work_items(void* handle) {
    return ((BatchPlan*)handle)->nitems;
}

__attribute__((visibility("default")))
void bli_amd_gemm_batch_plan_destroy(void* handle) {
    delete (BatchPlan*)handle;
}

} // extern "C"
