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
    int32_t     tiles_m;
    int32_t     pad;
};

struct WorkItem {
    int32_t prob;
    int32_t tile;
};

template <typename T> struct BatchTiles;
template <> struct BatchTiles<double> {
    using Big = TileCfg<double, 64, 64, 16, 32, 32>;
    using Small = TileCfg<double, 32, 32, 16, 16, 16>;
};
template <> struct BatchTiles<float> {
    using Big = TileCfg<float, 64, 64, 32, 32, 32>;
    using Small = TileCfg<float, 32, 32, 32, 16, 16>;
};

template <typename T, class TL>
__device__ inline void run_tile(const TileProblem<T>& p, int32_t flags, int64_t tm, int64_t tn,
                                T* smem) {
    const bool a_k = flags & F_A_K, b_k = flags & F_B_K;
    if (!a_k && !b_k) gemm_tile<T, TL, false, false>(p, tm, tn, smem);
    else if (!a_k && b_k) gemm_tile<T, TL, false, true>(p, tm, tn, smem);
    else if (a_k && !b_k) gemm_tile<T, TL, true, false>(p, tm, tn, smem);
    else gemm_tile<T, TL, true, true>(p, tm, tn, smem);
}

template <typename T>
__global__ __launch_bounds__(256) void gemm_batch_kernel(const BatchDesc* __restrict__ descs,
                                                         const WorkItem* __restrict__ items,
                                                         int nitems) {
    using Big = typename BatchTiles<T>::Big;
    using Small = typename BatchTiles<T>::Small;
    constexpr int SM = 2 * (Big::STAGE > Small::STAGE ? Big::STAGE : Small::STAGE);
    __shared__ T smem[SM];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const WorkItem w = items[it];
        const BatchDesc& d = descs[w.prob];
        TileProblem<T> p;
        p.a = (const T*)d.a; p.b = (const T*)d.b; p.c = (T*)d.c;
        p.m = d.m; p.n = d.n; p.k = d.k;
        p.lda = d.lda; p.ldb = d.ldb; p.c_si = d.c_si; p.c_sj = d.c_sj;
        p.alpha = (T)d.alpha; p.beta = (T)d.beta;
        p.vec_a = (d.flags & F_VEC_A) != 0;
        p.vec_b = (d.flags & F_VEC_B) != 0;
        p.vec_c = (d.flags & F_VEC_C) != 0;
        const int64_t tm = w.tile % d.tiles_m, tn = w.tile / d.tiles_m;
        if (d.flags & F_SMALL) run_tile<T, Small>(p, d.flags, tm, tn, smem);
        else run_tile<T, Big>(p, d.flags, tm, tn, smem);
        __syncthreads();
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
    int grid = 0;
    std::vector<BatchProblem> fallback;   // general-stride problems (plain GEMMs)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    ~BatchPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_descs) (void)hipFree(d_descs);
        if (d_items) (void)hipFree(d_items);
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
        const bool small = q.m <= 48 && q.n <= 48;
        if (small) flags |= F_SMALL;
        const int tb = small ? 32 : 64;
        d.flags = flags;
        d.tiles_m = (int32_t)((q.m + tb - 1) / tb);
        const int64_t tiles_n = (q.n + tb - 1) / tb;
        const int32_t pi = (int32_t)descs.size();
        descs.push_back(d);
        for (int64_t t = 0; t < (int64_t)d.tiles_m * tiles_n; ++t) {
            items.push_back(WorkItem{ pi, (int32_t)t });
            item_k.push_back(d.k);
        }
    }
    // Longest-first (by k extent; ties keep submission order).
    std::vector<int32_t> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return item_k[x] > item_k[y]; });
    std::vector<WorkItem> sorted(items.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
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
    HIP_CHECK(hipStreamSynchronize(st));
    const int per_cu = 4;
    plan->grid = std::max(1, std::min(plan->nitems, cu_count() * per_cu));
    return plan;
}

static void plan_launch(const BatchPlan* plan, hipStream_t st) {
    if (plan->nitems > 0) {
        if (plan->dt == Dt::F64)
            hipLaunchKernelGGL(gemm_batch_kernel<double>, dim3(plan->grid), dim3(256), 0, st,
                               plan->d_descs, plan->d_items, plan->nitems);
        else
            hipLaunchKernelGGL(gemm_batch_kernel<float>, dim3(plan->grid), dim3(256), 0, st,
                               plan->d_descs, plan->d_items, plan->nitems);
        HIP_CHECK(hipGetLastError());
    }
    for (const BatchProblem& q : plan->fallback) {
        GemmDesc d;
        d.dt = plan->dt; d.m = q.m; d.n = q.n; d.k = q.k;
        d.a = q.a; d.b = q.b; d.c = q.c; d.alpha = q.alpha; d.beta = q.beta;
        gemm_run(d, nullptr, st);
    }
}

void gemm_batch_run(Dt dt, const std::vector<BatchProblem>& probs, hipStream_t st) {
    BatchPlan* plan = plan_build(dt, probs, st);
    plan_launch(plan, st);
    HIP_CHECK(hipStreamSynchronize(st));
    delete plan;
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

__attribute__((visibility("default")))
int bli_amd_gemm_batch_plan_work_items(void* handle) {
    return ((BatchPlan*)handle)->nitems;
}

__attribute__((visibility("default")))
void bli_amd_gemm_batch_plan_destroy(void* handle) {
    delete (BatchPlan*)handle;
}

} // extern "C"
