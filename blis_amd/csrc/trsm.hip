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
template <typename T, bool LOWER, int CPW>
__global__ __launch_bounds__(64 * WPB) void trsm_leaf_kernel(const T* __restrict__ a, int64_t ars,
                                                             int64_t acs, T* __restrict__ b,
                                                             int64_t brs, int64_t bcs,
                                                             int nrows, int64_t n, T alpha,
                                                             int unit_diag) {
    __shared__ T xs[WPB][P][CPW][BR];
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
            xs[wave][p][c][lane] = (r < nrows && col < n) ? alpha * b[r * brs + col * bcs] : T(0);
        }
    __builtin_amdgcn_wave_barrier();

    T Lr[BR];
#pragma unroll 1
    for (int t = 0; t < nblk; ++t) {
        const int pb = LOWER ? t : nblk - 1 - t;   // block being solved
        const int row = pb * BR + lane;
        T x[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) x[c] = xs[wave][pb][c][lane];
        // Updates from the blocks solved earlier.
#pragma unroll 1
        for (int u = 0; u < t; ++u) {
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
                if (lane == i) x[c] = xi;
                else x[c] -= Lr[i] * xi;
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
}

struct Solve {
    Dt dt;
    bool lower;
    bool unit;
    View a;          // L_eff(i,j) = a.ptr[i*a.rs + j*a.cs]
    View b;          // B'(i,j)    = b.ptr[i*b.rs + j*b.cs]
    int64_t n;
    hipStream_t st;
};

inline void* offset(const View& v, int64_t i, int64_t j, size_t es) {
    return (char*)v.ptr + (i * v.rs + j * v.cs) * (int64_t)es;
}

template <typename T, int CPW>
void leaf_t(const Solve& s, const void* a, void* b, int64_t nb, double alpha) {
    const int64_t waves = (s.n + CPW - 1) / CPW;
    const dim3 grid((unsigned)((waves + WPB - 1) / WPB)), block(64 * WPB);
    if (s.lower)
        hipLaunchKernelGGL((trsm_leaf_kernel<T, true, CPW>), grid, block, 0, s.st, (const T*)a,
                           s.a.rs, s.a.cs, (T*)b, s.b.rs, s.b.cs, (int)nb, s.n, (T)alpha,
                           (int)s.unit);
    else
        hipLaunchKernelGGL((trsm_leaf_kernel<T, false, CPW>), grid, block, 0, s.st, (const T*)a,
                           s.a.rs, s.a.cs, (T*)b, s.b.rs, s.b.cs, (int)nb, s.n, (T)alpha,
                           (int)s.unit);
    HIP_CHECK(hipGetLastError());
}

void leaf(const Solve& s, int64_t r0, int64_t nb, double alpha) {
    const size_t es = dt_size(s.dt);
    void* a = offset(s.a, r0, r0, es);
    void* b = offset(s.b, r0, 0, es);
    // Two columns per wave: measured faster than one at n = 1024 on MI355X
    // (each row of the triangle is read once for both columns).
    if (s.dt == Dt::F64) leaf_t<double, 2>(s, a, b, nb, alpha);
    else leaf_t<float, 2>(s, a, b, nb, alpha);
}

// B'[dst rows] := beta * B'[dst rows] - L_eff[dst rows, src rows] * X[src rows]
void update(const Solve& s, int64_t dst0, int64_t dstn, int64_t src0, int64_t srcn,
            double beta) {
    const size_t es = dt_size(s.dt);
    GemmDesc d;
    d.dt = s.dt;
    d.m = dstn; d.n = s.n; d.k = srcn;
    d.a = { offset(s.a, dst0, src0, es), s.a.rs, s.a.cs };
    d.b = { offset(s.b, src0, 0, es), s.b.rs, s.b.cs };
    d.c = { offset(s.b, dst0, 0, es), s.b.rs, s.b.cs };
    d.alpha = -1.0;
    d.beta = beta;
    gemm_run(d, nullptr, s.st);
}

void solve(const Solve& s, int64_t r0, int64_t m, double alpha) {
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

} // namespace detail_trsm
using namespace detail_trsm;

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
    solve(s, 0, rows, alpha);
}

} // namespace blis_amd
