// sup.hip — the small-problem path for single GEMM calls (BLIS's "sup":
// frame/3/bli_l3_sup.c:37-130, taken below the MT/NT/KT thresholds of
// frame/base/bli_cntx.h:186-193).
//
// On the CPU, sup skips packing for small operands. On the GPU the cost of a
// small GEMM is latency: the launch and the chain of k slabs each
// workgroup walks (a 32 x 32 tile of a 512^3 product walks 32 slabs of 16,
// each a global load, an LDS store, a barrier and four MFMAs). The sup path
// keeps the 16 x 16 wave tiles of the smallest table tile (enough
// workgroups to spread a small output over the CUs) but stages k slabs up
// to 64 deep with two slabs in flight per workgroup, so the chain is up to
// a quarter as long and its loads overlap. Products with k <= 64 keep the
// table tile. Every output element still
// takes the same MFMA k-steps in the same order as every other tile
// configuration, so the results are the same bits.
#include "gemm_plan.hpp"

namespace blis_amd {

namespace {

// 32 x 32 workgroup tiles of four 16 x 16 waves; the k slab depth follows
// k (a slab deeper than k only loads zeros): 32 up to k = 256, 64 beyond.
template <typename T, int BK> using SupTL = TileCfg<T, 32, 32, BK, 16, 16>;
constexpr int SUP_THREADS = 256;   // four waves

} // namespace

template <typename T, int BK, bool A_K, bool B_K>
__global__ __launch_bounds__(SUP_THREADS) void gemm_sup_kernel(TileProblem<T> p, int grid_m) {
    using TL = SupTL<T, BK>;
    static_assert(TL::NT == SUP_THREADS, "sup tile shape");
    __shared__ T smem[2 * TL::STAGE];
    const int tm = blockIdx.x % grid_m, tn = blockIdx.x / grid_m;
    gemm_tile<T, TL, A_K, B_K, 2>(p, tm, tn, smem);
}

namespace {

template <typename T, int BK>
void launch_sup_bk(const Plan& pl, double alpha, double beta, hipStream_t st) {
    using TL = SupTL<T, BK>;
    const TileProblem<T> p = make_problem<T>(pl, alpha, beta);
    const int gm = (int)((pl.m + TL::BM - 1) / TL::BM);
    const int gn = (int)((pl.n + TL::BN - 1) / TL::BN);
    const dim3 grid(gm * gn), block(TL::NT);
    if (!pl.a_k && !pl.b_k)
        hipLaunchKernelGGL((gemm_sup_kernel<T, BK, false, false>), grid, block, 0, st, p, gm);
    else if (!pl.a_k && pl.b_k)
        hipLaunchKernelGGL((gemm_sup_kernel<T, BK, false, true>), grid, block, 0, st, p, gm);
    else if (pl.a_k && !pl.b_k)
        hipLaunchKernelGGL((gemm_sup_kernel<T, BK, true, false>), grid, block, 0, st, p, gm);
    else
        hipLaunchKernelGGL((gemm_sup_kernel<T, BK, true, true>), grid, block, 0, st, p, gm);
    HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_sup(const Plan& pl, double alpha, double beta, hipStream_t st) {
    if (pl.k <= 256) launch_sup_bk<T, 32>(pl, alpha, beta, st);
    else launch_sup_bk<T, 64>(pl, alpha, beta, st);
}

} // namespace

// Measured on MI355X (FP64, device time per call): 512^3 12.9 us against
// 16.3 us on the table's 32 x 32 tile, 256^3 7.6 against 9.6, 128^3 5.8
// against 6.4. Up to k = 64 the table tile (one 16-deep slab in flight) is
// as fast or faster (64^3: 4.6 against 4.9 us), so those calls keep it.
static bool sup_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("BLIS_AMD_SUP");
        return !(e && e[0] == '0');
    }();
    return on;
}

bool gemm_sup_applies(int64_t m, int64_t n, int64_t k) {
    return sup_enabled() && m <= SMALL_GEMM_MAX_DIM && n <= SMALL_GEMM_MAX_DIM &&
           k <= SMALL_GEMM_MAX_DIM && k > 64;
}

// Table tile 4 is the 32 x 32 tile of four 16 x 16 waves: the sup kernel
// computes the same tiles with deeper slabs. Measured (FP64, column-major,
// ld 4096): 512 x 1024 x 512 20.1 us against 21.3 us; at k = 1024 the
// 64-deep slabs allow only two workgroups per CU and lose (1024^3: 64.2
// against 57.8 us), so the substitution stops at k = 512.
bool gemm_sup_deep_k(int tile, int parts, int64_t k) {
    return sup_enabled() && tile == 4 && parts <= 1 && k > 64 && k <= SMALL_GEMM_MAX_DIM;
}

void gemm_sup_run(Dt dt, const Plan& pl, double alpha, double beta, hipStream_t st) {
    if (dt == Dt::F64) launch_sup<double>(pl, alpha, beta, st);
    else launch_sup<float>(pl, alpha, beta, st);
}

} // namespace blis_amd
