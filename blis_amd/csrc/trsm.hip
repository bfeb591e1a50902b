// trsm.hip — triangular solve B := alpha * inv(op(A)) * B (left) or
// B := alpha * B * inv(op(A)) (right) on gfx950.
//
// Reference structure (paths relative to the BLIS tree):
//   * right side is solved as a left-side problem on transposed operands
//     (frame/3/bli_l3_oapi_ex.c:753-758);
//   * blocked algorithm: diagonal blocks are solved by a triangular kernel and
//     the rest of B is updated by GEMM (frame/3/trsm/bli_trsm_blk_var1.c:37-189,
//     frame/3/trsm/bli_trsm_ll_ker_var2.c:240-335);
//   * the diagonal kernel multiplies by the reciprocal of each diagonal
//     element, which packing stores in place of the element
//     (BLIS_ENABLE_TRSM_PREINVERSION; ref_kernels/1m/bli_packm_cxc_diag_ref.c:
//     99-107, ref_kernels/3/bli_trsm_ref.c:42-231).
//
// GPU formulation. Every case is first reduced to a left-side solve
// L_eff * X = alpha * B' with L_eff lower or upper, by folding op(A) and the
// side into strides. The solve recurses on halves: solve the leading half,
// update the trailing half with one GEMM (the bulk of the m^2 n flops, on the
// MFMA GEMM path), solve the trailing half. Leaves of at most NB = 256 rows
// are solved by trsm_leaf_kernel in native precision: a wave owns CPW
// columns of B with one row per lane per 64-row register block, and solves
// by broadcast substitution against rows of the triangle held in registers,
// multiplying by the reciprocal of each diagonal element.
#include "internal.hpp"

namespace blis_amd {

namespace detail_trsm {

constexpr int BR = 64;          // rows per register block (one per lane)
constexpr int P = 4;            // register blocks per leaf
constexpr int NB = BR * P;      // leaf rows
constexpr int WPB = 4;          // waves per workgroup
// Right-hand sides from which the inverted-diagonal path pays: inverting a
// block costs about as much as solving NB right-hand sides by substitution.
constexpr int64_t TRSM_INV_MIN_N = 512;
// Columns per wave when inverting the diagonal blocks: each wave reads the
// triangle's rows once for all of its columns.
constexpr int INV_CPW = 2;

__device__ inline double lane_bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float lane_bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Solve a leaf of nrows <= NB rows for all n columns of B:
// L * X = alpha * B with L(i,j) = a[i*ars + j*acs] (lower or upper, only the
// stored triangle is read), B(i,j) = b[i*brs + j*bcs], X overwrites B.
// Each wave owns CPW columns; its slab of B sits in LDS as P blocks of 64
// rows (lane l holds row p*64 + l). Blocks are solved in order: the block in
// work is loaded into registers, updated by the blocks solved before it
// (x -= L_pq x_q, the gemm part of gemmtrsm, reading x_q as LDS broadcasts),
// then solved by 64 broadcast steps: the solved row i, scaled by the
// reciprocal of the diagonal, is read from lane i and every other lane
// subtracts its L element times it.
// Several independent leaves in one launch: leaf y = blockIdx.y starts at
// a + y*a_step, b + y*b_step, and the last one has nrows_last rows.
// INV: the right-hand side is alpha times the identity (b is written only),
// so the leaf computes alpha inv(L). A wave's CPW columns lie in one 64-row
// block p0, and the solution is zero in the blocks on the far side of p0
// (above it for a lower triangle, below it for an upper one): those are
// neither solved nor used in updates.
template <typename T, bool LOWER, int CPW, bool INV = false>
__global__ __launch_bounds__(64 * WPB) void trsm_leaf_kernel(const T* __restrict__ a, int64_t ars,
                                                             int64_t acs, T* __restrict__ b,
                                                             int64_t brs, int64_t bcs,
                                                             int nrows, int64_t n, T alpha,
                                                             int unit_diag, int64_t a_step,
                                                             int64_t b_step, int nrows_last,
                                                             int zgroup) {
    __shared__ T xs[WPB][P][CPW][BR];
    a += (int64_t)blockIdx.y * a_step;
    b += (int64_t)blockIdx.y * b_step;
    if (blockIdx.y == gridDim.y - 1) nrows = nrows_last;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t col0 = ((int64_t)blockIdx.x * WPB + wave) * CPW;
    if (col0 >= n) return;      // whole wave leaves together (no block barriers below)
    const int nblk = (nrows + BR - 1) / BR;

    for (int p = 0; p < nblk; ++p)
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int r = p * BR + lane;
            const int64_t col = col0 + c;
            T v = T(0);
            if (INV) v = (r < nrows && col < n && r == col) ? alpha : T(0);
            else if (r < nrows && col < n) v = alpha * b[r * brs + col * bcs];
            xs[wave][p][c][lane] = v;
        }
    __builtin_amdgcn_wave_barrier();

    // First block in solve order that can be nonzero.
    int t0 = 0;
    if (INV) {
        const int p0 = (int)(col0 / BR);
        t0 = LOWER ? p0 : nblk - 1 - p0;
        if (t0 < 0) t0 = 0;
    }
    T Lr[BR];
#pragma unroll 1
    for (int t = t0; t < nblk; ++t) {
        const int pb = LOWER ? t : nblk - 1 - t;   // block being solved
        const int row = pb * BR + lane;
        T x[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) x[c] = xs[wave][pb][c][lane];
        // Updates from the blocks solved earlier.
#pragma unroll 1
        for (int u = t0; u < t; ++u) {
            const int q = LOWER ? u : nblk - 1 - u;
#pragma unroll
            for (int i = 0; i < BR; ++i) {
                const int colA = q * BR + i;
                Lr[i] = (row < nrows && colA < nrows) ? a[row * ars + colA * acs] : T(0);
            }
#pragma unroll
            for (int i = 0; i < BR; ++i)
#pragma unroll
                for (int c = 0; c < CPW; ++c) x[c] -= Lr[i] * xs[wave][q][c][i];
        }
        // Diagonal block: stored triangle only, reciprocal of the diagonal.
#pragma unroll
        for (int i = 0; i < BR; ++i) {
            const int colA = pb * BR + i;
            const bool stored = LOWER ? (i < lane) : (i > lane);
            Lr[i] = (stored && row < nrows && colA < nrows) ? a[row * ars + colA * acs] : T(0);
        }
        T inv = T(1);
        if (!unit_diag && row < nrows) inv = T(1) / a[row * ars + row * acs];
#pragma unroll
        for (int s = 0; s < BR; ++s) {
            const int i = LOWER ? s : BR - 1 - s;
            const T inv_i = lane_bcast(inv, i);
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const T xi = lane_bcast(x[c], i) * inv_i;
                // Rows solved earlier are left as they are (an Inf in xi
                // must not turn them into NaN through a zero of L).
                if (lane == i) x[c] = xi;
                else if (LOWER ? lane > i : lane < i) x[c] -= Lr[i] * xi;
            }
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c) xs[wave][pb][c][lane] = x[c];
        // LDS is in order within a wave; this keeps the compiler from moving
        // the reads of the next block above the writes.
        __builtin_amdgcn_wave_barrier();
    }

    for (int p = 0; p < nblk; ++p)
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int r = p * BR + lane;
            const int64_t col = col0 + c;
            if (r < nrows && col < n) b[r * brs + col * bcs] = xs[wave][p][c][lane];
        }
    // Inverses of nrows-row leaves inside a block of zgroup leaves (INV
    // only): zero the leaf's columns of that block on the far side of the
    // diagonal, rows above the leaf (lower) or below it (upper).
    if (INV && zgroup > 0) {
        const int d = (int)(blockIdx.y % zgroup);
        const int z0 = LOWER ? -d * nrows : nrows;
        const int zn = LOWER ? d * nrows : (zgroup - 1 - d) * nrows;
        for (int r = lane; r < zn; r += 64)
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int64_t col = col0 + c;
                if (col < n) b[(int64_t)(z0 + r) * brs + col * bcs] = T(0);
            }
    }
}

// Which diagonal blocks may be solved by multiplying with their inverse.
// Block q (rows q*NB .. q*NB + nb - 1, one workgroup of NB threads, thread
// t on row q*NB + t) passes when every row is diagonally dominant by a
// factor of two within the block, |l_ii| > 2 * sum_{j != i} |l_ij|, and
// the diagonal magnitudes stay within a factor of 16 of each other. Then
// L = D (I + D^-1 N) with ||D^-1 N||_inf < 1/2, so ||inv(L)||_inf <=
// 2 / min|l_ii| and cond_inf(L) <= 3 * max|l_ii| / min|l_ii| <= 48: the
// product with the FP64 (FP32) inverse keeps the residual of substitution
// to within that factor.
template <typename T, bool LOWER>
__global__ __launch_bounds__(NB) void trsm_diag_check_kernel(const T* __restrict__ a, int64_t ars,
                                                            int64_t acs, int64_t m,
                                                            int unit_diag, int* __restrict__ ok) {
    __shared__ int bad;
    // Largest and smallest |l_ii| as IEEE bit patterns: for non-negative
    // doubles the unsigned order of the patterns is the numeric order, so
    // the extremes are exact whatever the diagonal's range.
    __shared__ unsigned long long dmax, dmin;
    const int t = threadIdx.x;
    const int64_t q0 = (int64_t)blockIdx.x * NB;
    const int64_t nb = (m - q0 < NB) ? m - q0 : NB;
    if (t == 0) { bad = 0; dmax = 0ull; dmin = ~0ull; }
    __syncthreads();
    if (t < nb) {
        const int64_t r = q0 + t;
        // Eight independent partial sums keep eight loads in flight.
        double part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int64_t j0 = LOWER ? q0 : r + 1, j1 = LOWER ? r : q0 + nb;
        int64_t j = j0;
        for (; j + 8 <= j1; j += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) part[u] += fabs((double)a[r * ars + (j + u) * acs]);
        }
        for (; j < j1; ++j) part[0] += fabs((double)a[r * ars + j * acs]);
        const double off = ((part[0] + part[1]) + (part[2] + part[3])) +
                           ((part[4] + part[5]) + (part[6] + part[7]));
        const double d = unit_diag ? 1.0 : fabs((double)a[r * ars + r * acs]);
        if (!(d > 2.0 * off) || !(d < 1.0e300)) atomicOr(&bad, 1);
        const unsigned long long bits = (unsigned long long)__double_as_longlong(d);
        atomicMax(&dmax, bits);
        atomicMin(&dmin, bits);
    }
    __syncthreads();
    if (t == 0) {
        const double hi = __longlong_as_double((long long)dmax);
        const double lo = __longlong_as_double((long long)dmin);
        ok[blockIdx.x] = (!bad && hi <= 16.0 * lo) ? 1 : 0;
    }
}

// dst(i, j) = src(i, j) for i < rows, j < cols, both general-strided, through
// 64 x 64 LDS tiles: reads run along i and writes along j, so a copy between
// column-major and row-major storage is coalesced on both sides.
template <typename T>
__global__ __launch_bounds__(256) void trsm_copy_kernel(const T* __restrict__ src, int64_t srs,
                                                        int64_t scs, T* __restrict__ dst,
                                                        int64_t drs, int64_t dcs, int64_t rows,
                                                        int64_t cols) {
    __shared__ T tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t ti = (rows + 63) / 64, tj = (cols + 63) / 64;
    for (int64_t t = blockIdx.x; t < ti * tj; t += gridDim.x) {
        const int64_t i0 = (t % ti) * 64, j0 = (t / ti) * 64;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t i = i0 + tx, j = j0 + ty + 4 * q;
            if (i < rows && j < cols) tile[tx][ty + 4 * q] = src[i * srs + j * scs];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int64_t i = i0 + ty + 4 * q, j = j0 + tx;
            if (i < rows && j < cols) dst[i * drs + j * dcs] = tile[ty + 4 * q][tx];
        }
        __syncthreads();
    }
}

struct Solve {
    Dt dt;
    bool lower;
    bool unit;
    View a;          // L_eff(i,j) = a.ptr[i*a.rs + j*a.cs]
    View b;          // B'(i,j)    = b.ptr[i*b.rs + j*b.cs]
    int64_t n;
    hipStream_t st;
    // Inverted diagonal blocks (NB x NB each, column-major, ld NB) and the
    // per-block flag that allows their use; null when not prepared.
    const void* linv = nullptr;
    const int* inv_ok = nullptr;
    void* wtmp = nullptr;   // NB x n staging for a product leaf solved in place
    // Where the updates modify the right-hand side and where the solved rows
    // live. Both are b for the in-place solve; when every block is solved by
    // product they are separate buffers and B is written once at the end.
    View rhs;
    View x;
    bool separate = false;
    // Solves with every block inverted: inverses of the first npair aligned
    // 512-row diagonal blocks (512 x 512 each, column-major, ld 512).
    const void* linv2 = nullptr;
    int64_t npair = 0;
    // Band store of the inverses (when set, linv and linv2 only mark that
    // the 256- and 512-row inverses exist): element (i, j) of the block
    // diagonal inverse at band[(i - j + BAND_KL) + j * BAND_LD], so every
    // diagonal block, and every block within BAND_KL of the diagonal, is a
    // column-major matrix with leading dimension BAND_LD - 1, and the
    // blocks along the diagonal are a constant stride apart.
    const void* band = nullptr;
};

constexpr int64_t BAND_KL = 2 * NB;
constexpr int64_t BAND_LD = 2 * BAND_KL + 1;

// Block (r, c) of the inverse in the band store (|r - c| + block size - 1 <= BAND_KL).
inline View band_view(const Solve& s, int64_t r, int64_t c) {
    return { (char*)const_cast<void*>(s.band) + ((r - c + BAND_KL) + c * BAND_LD) * (int64_t)dt_size(s.dt),
             1, BAND_LD - 1 };
}

// Diagonal quarters of the 512-row inverses from the 256-row ones (block
// 2p at the top left, 2p+1 at the bottom right of pair p) and zeros in the
// quarter outside the triangle; the other off-diagonal quarter is written by
// GEMMs.
template <typename T>
__global__ __launch_bounds__(256) void trsm_assemble_kernel(const T* __restrict__ linv,
                                                            T* __restrict__ linv2,
                                                            int64_t npair, int upper) {
    constexpr int64_t PB = 2 * NB;
    const int64_t total = npair * PB * PB;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / (PB * PB), e = t % (PB * PB), i = e % PB, j = e / PB;
        if (upper ? (i < NB && j >= NB) : (i >= NB && j < NB)) continue;
        T v = T(0);
        if (i < NB && j < NB) v = linv[(2 * p) * NB * NB + i + j * NB];
        else if (i >= NB && j >= NB) v = linv[(2 * p + 1) * NB * NB + (i - NB) + (j - NB) * NB];
        linv2[t] = v;
    }
}

inline void* offset(const View& v, int64_t i, int64_t j, size_t es) {
    return (char*)v.ptr + (i * v.rs + j * v.cs) * (int64_t)es;
}

// Launch `count` leaves of the same shape: leaf y solves the nb-row (the
// last one nb_last-row) triangle at a + y*a_step for the n columns at
// b + y*b_step with strides (brs, bcs).
template <typename T, int CPW, bool INV = false>
void leaf_launch(bool lower, bool unit, const void* a, int64_t ars, int64_t acs, void* b,
                 int64_t brs, int64_t bcs, int64_t nb, int64_t n, double alpha, int count,
                 int64_t a_step, int64_t b_step, int64_t nb_last, hipStream_t st,
                 int zgroup = 0) {
    static_assert(BR % CPW == 0, "a wave's columns must lie in one 64-row block");
    const int64_t waves = (n + CPW - 1) / CPW;
    const dim3 grid((unsigned)((waves + WPB - 1) / WPB), (unsigned)count), block(64 * WPB);
    if (lower)
        hipLaunchKernelGGL((trsm_leaf_kernel<T, true, CPW, INV>), grid, block, 0, st,
                           (const T*)a, ars, acs, (T*)b, brs, bcs, (int)nb, n, (T)alpha,
                           (int)unit, a_step, b_step, (int)nb_last, zgroup);
    else
        hipLaunchKernelGGL((trsm_leaf_kernel<T, false, CPW, INV>), grid, block, 0, st,
                           (const T*)a, ars, acs, (T*)b, brs, bcs, (int)nb, n, (T)alpha,
                           (int)unit, a_step, b_step, (int)nb_last, zgroup);
    HIP_CHECK(hipGetLastError());
}

template <typename T, int CPW>
void leaf_t(const Solve& s, const void* a, void* b, int64_t nb, double alpha) {
    leaf_launch<T, CPW>(s.lower, s.unit, a, s.a.rs, s.a.cs, b, s.b.rs, s.b.cs, nb, s.n, alpha, 1,
                        0, 0, nb, s.st);
}

// A leaf whose diagonal block passed the check: X_q = alpha * inv(L_qq) * R_q
// as one GEMM, straight into x when it is a separate buffer, else into
// staging and then copied over B_q.
void leaf_by_product(const Solve& s, int64_t r0, int64_t nb, double alpha) {
    const size_t es = dt_size(s.dt);
    const int64_t q = r0 / NB;
    GemmDesc d;
    d.dt = s.dt;
    d.m = nb; d.n = s.n; d.k = nb;
    d.a = s.band ? band_view(s, r0, r0) : View{ (char*)s.linv + (size_t)q * NB * NB * es, 1, NB };
    d.b = { offset(s.rhs, r0, 0, es), s.rhs.rs, s.rhs.cs };
    d.alpha = alpha;
    d.beta = 0.0;
    if (s.separate) {
        d.c = { offset(s.x, r0, 0, es), s.x.rs, s.x.cs };
        gemm_run(d, nullptr, s.st);
        return;
    }
    d.c = { s.wtmp, 1, nb };
    gemm_run(d, nullptr, s.st);
    const int64_t tiles = ((nb + 63) / 64) * ((s.n + 63) / 64);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 8192));
    if (s.dt == Dt::F64)
        hipLaunchKernelGGL(trsm_copy_kernel<double>, dim3(g), dim3(256), 0, s.st,
                           (const double*)s.wtmp, 1, nb, (double*)offset(s.b, r0, 0, es), s.b.rs,
                           s.b.cs, nb, s.n);
    else
        hipLaunchKernelGGL(trsm_copy_kernel<float>, dim3(g), dim3(256), 0, s.st,
                           (const float*)s.wtmp, 1, nb, (float*)offset(s.b, r0, 0, es), s.b.rs,
                           s.b.cs, nb, s.n);
    HIP_CHECK(hipGetLastError());
}

void leaf(const Solve& s, int64_t r0, int64_t nb, double alpha) {
    if (s.linv && r0 % NB == 0 && s.inv_ok[r0 / NB]) {
        leaf_by_product(s, r0, nb, alpha);
        return;
    }
    const size_t es = dt_size(s.dt);
    void* a = offset(s.a, r0, r0, es);
    void* b = offset(s.b, r0, 0, es);
    // Two columns per wave: measured faster than one at n = 1024 on MI355X
    // (each row of the triangle is read once for both columns).
    if (s.dt == Dt::F64) leaf_t<double, 2>(s, a, b, nb, alpha);
    else leaf_t<float, 2>(s, a, b, nb, alpha);
}

// R[dst rows] := beta * R[dst rows] - L_eff[dst rows, src rows] * X[src rows]
void update(const Solve& s, int64_t dst0, int64_t dstn, int64_t src0, int64_t srcn,
            double beta) {
    const size_t es = dt_size(s.dt);
    GemmDesc d;
    d.dt = s.dt;
    d.m = dstn; d.n = s.n; d.k = srcn;
    d.a = { offset(s.a, dst0, src0, es), s.a.rs, s.a.cs };
    d.b = { offset(s.x, src0, 0, es), s.x.rs, s.x.cs };
    d.c = { offset(s.rhs, dst0, 0, es), s.rhs.rs, s.rhs.cs };
    d.alpha = -1.0;
    d.beta = beta;
    // FP64 updates of 1024 and 2048 rows (n = 1024, k = rows, operands in
    // the solve's layout), measured over the tile table with 1, 2 and 4 k
    // partitions: 1024 rows are fastest on 128x64x16 tiles of eight waves
    // in two partitions (53.1 us against 59.7 on 64x64x32 tiles and 57.8 on
    // the size heuristic's 32x32), 2048 rows on 128x64x32 tiles in one
    // (158 against 173 us on the heuristic's 64x64). Every tile gives the
    // same result; the partitioned sum rounds in another order.
    GemmConfig cfg;
    if (s.dt == Dt::F64 && dstn >= 1536 && srcn >= 1536) cfg.tile = 11;
    else if (s.dt == Dt::F64 && dstn >= 768 && srcn >= 768) { cfg.tile = 7; cfg.split_k = 2; }
    gemm_run(d, cfg.tile >= 0 ? &cfg : nullptr, s.st);
}

// X[r0 .. r0+511] = alpha * inv(L_pp) * R[r0 .. r0+511] for an aligned
// 512-row block whose inverse was composed: one product in place of a
// leaf, an update and a leaf.
void leaf2_by_product(const Solve& s, int64_t r0, double alpha) {
    const size_t es = dt_size(s.dt);
    GemmDesc d;
    d.dt = s.dt;
    d.m = 2 * NB; d.n = s.n; d.k = 2 * NB;
    d.a = s.band ? band_view(s, r0, r0)
                 : View{ (char*)s.linv2 + (size_t)(r0 / (2 * NB)) * 4 * NB * NB * es, 1, 2 * NB };
    d.b = { offset(s.rhs, r0, 0, es), s.rhs.rs, s.rhs.cs };
    d.c = { offset(s.x, r0, 0, es), s.x.rs, s.x.cs };
    d.alpha = alpha;
    d.beta = 0.0;
    gemm_run(d, nullptr, s.st);
}

void solve(const Solve& s, int64_t r0, int64_t m, double alpha) {
    if (s.linv2 && m == 2 * NB && r0 % (2 * NB) == 0 && r0 / (2 * NB) < s.npair) {
        leaf2_by_product(s, r0, alpha);
        return;
    }
    if (m <= NB) {
        leaf(s, r0, m, alpha);
        return;
    }
    // Split into halves rounded to whole leaves.
    int64_t h = ((m / 2) + NB - 1) / NB * NB;
    if (h >= m) h = m - NB;
    const int64_t m1 = h, m2 = m - h;
    if (s.lower) {
        solve(s, r0, m1, alpha);
        update(s, r0 + m1, m2, r0, m1, alpha);
        solve(s, r0 + m1, m2, 1.0);
    } else {
        solve(s, r0 + m1, m2, alpha);
        update(s, r0, m1, r0 + m1, m2, alpha);
        solve(s, r0, m1, 1.0);
    }
}

// Inverses of the aligned 512-row diagonal blocks from the 256-row inverses
// at linv (ld NB): inv([[L11, 0], [L21, L22]]) =
// [[inv(L11), 0], [-inv(L22) L21 inv(L11), inv(L22)]] for a lower triangle,
// inv([[U11, U12], [0, U22]]) = [[inv(U11), -inv(U11) U12 inv(U22)],
// [0, inv(U22)]] for an upper one. Two strided launches of 256^3 products
// cover every pair. The blocks passed the dominance check, so the inverse
// applied first has condition number at most 48 and the product form of
// the 512-row leaf loses at most that factor against the leaf, update, leaf
// sequence it replaces.
void compose_pairs(Solve& s, const void* linv, int64_t rows) {
    const size_t es = dt_size(s.dt);
    const int64_t npair = rows / (2 * NB);
    if (npair < 1) return;
    void* linv2 = workspace((size_t)npair * 4 * NB * NB * es, 18);
    void* t = workspace((size_t)npair * NB * NB * es, 19);
    const int64_t total = npair * 4 * NB * NB;
    const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 8192);
    const int upper = s.lower ? 0 : 1;
    if (s.dt == Dt::F64)
        hipLaunchKernelGGL(trsm_assemble_kernel<double>, dim3(g), dim3(256), 0, s.st,
                           (const double*)linv, (double*)linv2, npair, upper);
    else
        hipLaunchKernelGGL(trsm_assemble_kernel<float>, dim3(g), dim3(256), 0, s.st,
                           (const float*)linv, (float*)linv2, npair, upper);
    HIP_CHECK(hipGetLastError());
    const size_t blk = (size_t)NB * NB * es;
    // Lower: T_p = L21_p * inv(L11_p). Upper: T_p = U12_p * inv(U22_p).
    GemmDesc d;
    d.dt = s.dt;
    d.m = NB; d.n = NB; d.k = NB;
    d.a = upper ? View{ offset(s.a, 0, NB, es), s.a.rs, s.a.cs }
                : View{ offset(s.a, NB, 0, es), s.a.rs, s.a.cs };
    d.b = { (char*)const_cast<void*>(linv) + (upper ? blk : 0), 1, NB };
    d.c = { t, 1, NB };
    d.alpha = 1.0;
    d.beta = 0.0;
    gemm_strided_run(d, (int)npair, 2 * NB * (s.a.rs + s.a.cs), 2 * NB * NB, NB * NB, s.st);
    // Lower: -inv(L22_p) * T_p into the bottom-left quarter of pair p.
    // Upper: -inv(U11_p) * T_p into the top-right quarter.
    d.a = { (char*)const_cast<void*>(linv) + (upper ? 0 : blk), 1, NB };
    d.b = { t, 1, NB };
    d.c = { (char*)linv2 + (upper ? (size_t)NB * 2 * NB * es : (size_t)NB * es), 1, 2 * NB };
    d.alpha = -1.0;
    gemm_strided_run(d, (int)npair, 2 * NB * NB, NB * NB, 4 * NB * NB, s.st);
    s.linv2 = linv2;
    s.npair = npair;
}

// The same composition in the band store, for every aligned pair of b-row
// inverses: 2b-row inverses from b-row ones (b = 64, 128, 256). Both
// products are one strided launch over the pairs: in the band the pairs are
// 2b * BAND_LD elements apart, in the triangle 2b (rs + cs).
void compose_band(const Solve& s, int64_t rows, int64_t b) {
    const size_t es = dt_size(s.dt);
    const int64_t npair = rows / (2 * b);
    if (npair < 1) return;
    void* t = workspace((size_t)npair * b * b * es, 19);
    const bool upper = !s.lower;
    const int64_t step = 2 * b * BAND_LD;
    GemmDesc d;
    d.dt = s.dt;
    d.m = b; d.n = b; d.k = b;
    // Lower: T_p = L21_p * inv(L11_p). Upper: T_p = U12_p * inv(U22_p).
    d.a = upper ? View{ offset(s.a, 0, b, es), s.a.rs, s.a.cs }
                : View{ offset(s.a, b, 0, es), s.a.rs, s.a.cs };
    d.b = upper ? band_view(s, b, b) : band_view(s, 0, 0);
    d.c = { t, 1, b };
    d.alpha = 1.0;
    d.beta = 0.0;
    gemm_strided_run(d, (int)npair, 2 * b * (s.a.rs + s.a.cs), step, b * b, s.st);
    // Lower: -inv(L22_p) * T_p below the diagonal. Upper: -inv(U11_p) * T_p
    // to its right.
    d.a = upper ? band_view(s, 0, 0) : band_view(s, b, b);
    d.b = { t, 1, b };
    d.c = upper ? band_view(s, 0, b) : band_view(s, b, 0);
    d.alpha = -1.0;
    gemm_strided_run(d, (int)npair, step, b * b, step, s.st);
}

} // namespace detail_trsm
using namespace detail_trsm;

// The stream of the diagonal-block check: created once, default priority,
// not ordered with the legacy default stream.
hipStream_t check_stream() {
    static hipStream_t cs = nullptr;
    if (!cs) HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    return cs;
}

void trsm_run(Dt dt, side_t side, uplo_t uplo, trans_t trans, diag_t diag, int64_t m,
              int64_t n, double alpha, View a, View b, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const bool t = ((unsigned)trans & BLIS_TRANS_BIT) != 0;
    const bool lower = (uplo == BLIS_LOWER);
    Solve s;
    s.dt = dt;
    s.unit = (diag == BLIS_UNIT_DIAG);
    s.st = st;
    int64_t rows;
    if (side == BLIS_LEFT) {
        // L_eff = op(A).
        s.a = { a.ptr, t ? a.cs : a.rs, t ? a.rs : a.cs };
        s.lower = t ? !lower : lower;
        s.b = b;
        s.n = n;
        rows = m;
    } else {
        // X op(A) = alpha B  <=>  op(A)^T X^T = alpha B^T.
        s.a = { a.ptr, t ? a.rs : a.cs, t ? a.cs : a.rs };
        s.lower = t ? lower : !lower;
        s.b = { b.ptr, b.cs, b.rs };
        s.n = m;
        rows = n;
    }
    s.rhs = s.b;
    s.x = s.b;
    // Many leaves over many columns: check the diagonal blocks, invert them
    // all in one launch, and solve the blocks that passed by GEMM products.
    std::vector<int> ok;
    if (rows > NB && s.n >= TRSM_INV_MIN_N) {
        const size_t es = dt_size(dt);
        const int nblk = (int)((rows + NB - 1) / NB);
        const int64_t last = rows - (int64_t)(nblk - 1) * NB;
        int* d_ok = (int*)workspace(sizeof(int) * nblk, 15);
        // Whole 512-row blocks and a unit-stride triangle: the inverses are
        // built in the band store from 64-row leaves; otherwise 256-row
        // leaves are inverted directly (ragged last block).
        const bool use_band = rows % (2 * NB) == 0 && (s.a.rs == 1 || s.a.cs == 1);
        const int64_t total = (int64_t)nblk * NB * NB;
        void* linv = use_band ? workspace((size_t)(rows + 1) * BAND_LD * es, 13)
                              : workspace((size_t)total * es, 13);
        const int64_t a_step = NB * (s.a.rs + s.a.cs);
        // The check runs on a side stream, beside the inversion on the main
        // stream; only the host reads its flags (pinned memory, one wait).
        hipStream_t cs = check_stream();
        hipEvent_t ready, checked;
        HIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&checked, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(ready, st));
        HIP_CHECK(hipStreamWaitEvent(cs, ready, 0));
        if (dt == Dt::F64) {
            if (s.lower)
                hipLaunchKernelGGL((trsm_diag_check_kernel<double, true>), dim3(nblk), dim3(NB), 0,
                                   cs, (const double*)s.a.ptr, s.a.rs, s.a.cs, rows, (int)s.unit,
                                   d_ok);
            else
                hipLaunchKernelGGL((trsm_diag_check_kernel<double, false>), dim3(nblk), dim3(NB),
                                   0, cs, (const double*)s.a.ptr, s.a.rs, s.a.cs, rows,
                                   (int)s.unit, d_ok);
        } else {
            if (s.lower)
                hipLaunchKernelGGL((trsm_diag_check_kernel<float, true>), dim3(nblk), dim3(NB), 0,
                                   cs, (const float*)s.a.ptr, s.a.rs, s.a.cs, rows, (int)s.unit,
                                   d_ok);
            else
                hipLaunchKernelGGL((trsm_diag_check_kernel<float, false>), dim3(nblk), dim3(NB), 0,
                                   cs, (const float*)s.a.ptr, s.a.rs, s.a.cs, rows, (int)s.unit,
                                   d_ok);
        }
        HIP_CHECK(hipGetLastError());
        ok.assign(nblk, 0);
        static int* pinned = nullptr;
        constexpr int PINNED_FLAGS = 1 << 16;
        if (!pinned) HIP_CHECK(hipHostMalloc((void**)&pinned, sizeof(int) * PINNED_FLAGS));
        int* flags = nblk <= PINNED_FLAGS ? pinned : ok.data();
        HIP_CHECK(hipMemcpyAsync(flags, d_ok, sizeof(int) * nblk, hipMemcpyDeviceToHost, cs));
        HIP_CHECK(hipEventRecord(checked, cs));
        // The inverses: leaves whose right-hand side is the identity, built
        // in the kernel, skipping each column's zero blocks.
        if (use_band) {
            // 64-row leaves straight into the band (each also zeroes its
            // columns of the 512-row block on the far side of the
            // diagonal), then 128- and 256-row inverses by composition.
            constexpr int64_t LB = 64;
            s.band = linv;
            const View b0 = band_view(s, 0, 0);
            const int count = (int)(rows / LB);
            const int zgroup = (int)(2 * NB / LB);
            if (dt == Dt::F64)
                leaf_launch<double, INV_CPW, true>(s.lower, s.unit, s.a.ptr, s.a.rs, s.a.cs,
                                                   b0.ptr, 1, b0.cs, LB, LB, 1.0, count,
                                                   LB * (s.a.rs + s.a.cs), LB * BAND_LD, LB, st,
                                                   zgroup);
            else
                leaf_launch<float, INV_CPW, true>(s.lower, s.unit, s.a.ptr, s.a.rs, s.a.cs,
                                                  b0.ptr, 1, b0.cs, LB, LB, 1.0, count,
                                                  LB * (s.a.rs + s.a.cs), LB * BAND_LD, LB, st,
                                                  zgroup);
            compose_band(s, rows, LB);
            compose_band(s, rows, 2 * LB);
            // The 512-row inverses are used only when every block passes;
            // they are composed before the wait for the check anyway, so
            // that the GPU stays busy while the host reads the flags.
            compose_band(s, rows, NB);
        } else if (dt == Dt::F64) {
            leaf_launch<double, INV_CPW, true>(s.lower, s.unit, s.a.ptr, s.a.rs, s.a.cs, linv, 1,
                                               NB, NB, NB, 1.0, nblk, a_step, (int64_t)NB * NB,
                                               last, st);
        } else {
            leaf_launch<float, INV_CPW, true>(s.lower, s.unit, s.a.ptr, s.a.rs, s.a.cs, linv, 1,
                                              NB, NB, NB, 1.0, nblk, a_step, (int64_t)NB * NB,
                                              last, st);
        }
        HIP_CHECK(hipEventSynchronize(checked));
        HIP_CHECK(hipEventDestroy(checked));
        HIP_CHECK(hipEventDestroy(ready));
        if (flags != ok.data()) std::copy(flags, flags + nblk, ok.begin());
        s.linv = linv;
        s.inv_ok = ok.data();
        auto copy = [&](View src, View dst) {
            if (src.rs == 1 && dst.rs == 1 && src.cs == rows && dst.cs == rows) {
                // Both column-stored without padding: one contiguous span,
                // which the runtime's copy moves at about 5.5 TB/s against
                // 2.3 TB/s for the tiled kernel (32 MB: 12 against 28 us).
                HIP_CHECK(hipMemcpyAsync(dst.ptr, src.ptr, (size_t)rows * s.n * es,
                                         hipMemcpyDeviceToDevice, st));
                return;
            }
            const int64_t tiles = ((rows + 63) / 64) * ((s.n + 63) / 64);
            const unsigned gc = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 8192));
            if (dt == Dt::F64)
                hipLaunchKernelGGL(trsm_copy_kernel<double>, dim3(gc), dim3(256), 0, st,
                                   (const double*)src.ptr, src.rs, src.cs, (double*)dst.ptr,
                                   dst.rs, dst.cs, rows, s.n);
            else
                hipLaunchKernelGGL(trsm_copy_kernel<float>, dim3(gc), dim3(256), 0, st,
                                   (const float*)src.ptr, src.rs, src.cs, (float*)dst.ptr,
                                   dst.rs, dst.cs, rows, s.n);
            HIP_CHECK(hipGetLastError());
        };
        bool all = true;
        for (int v : ok) all = all && v;
        if (all) {
            // Every block by product: the right-hand side is updated in place
            // and X is built in a separate column-major buffer, copied over B
            // once at the end (no per-leaf staging copies).
            s.x = { workspace((size_t)rows * s.n * es, 16), 1, rows };
            s.separate = true;
            // The composition multiplies off-diagonal blocks of the triangle
            // in place, which needs a unit stride; a general-stride triangle
            // keeps 256-row leaves with updates between them (the update
            // GEMMs pack it).
            if (use_band) {
                s.linv2 = linv;
                s.npair = rows / (2 * NB);
            } else if (rows >= 2 * NB && (s.a.rs == 1 || s.a.cs == 1)) {
                compose_pairs(s, linv, rows);
            }
            solve(s, 0, rows, alpha);
            copy(s.x, s.b);
            return;
        }
        // Some blocks by substitution: in place, product leaves through
        // staging.
        s.wtmp = workspace((size_t)NB * s.n * es, 14);
    }
    solve(s, 0, rows, alpha);
}

} // namespace blis_amd
