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
// MFMA GEMM path), solve the trailing half. Leaves of at most NB = 64 rows
// are solved by trsm_diag_kernel in native precision: each wave owns 64
// columns of B, the diagonal block sits in LDS (its diagonal pre-inverted),
// and each lane runs forward (or backward) substitution for one column with
// the column held in registers.
#include "internal.hpp"

namespace blis_amd {

namespace detail_trsm {

constexpr int NB = 64;        // leaf block rows
constexpr int COLS = 64;      // columns of B per workgroup (one per lane)

// Solve a leaf: rows [0, nb) of B (rows offset already applied), all n
// columns. A(i,j) = a[i*ars + j*acs] restricted to the leaf's triangle.
template <typename T, bool LOWER>
__global__ __launch_bounds__(64) void trsm_diag_kernel(const T* __restrict__ a, int64_t ars,
                                                       int64_t acs, T* __restrict__ b,
                                                       int64_t brs, int64_t bcs, int nb,
                                                       int64_t n, T alpha, int unit_diag) {
    // Lc[j][i] = A(i,j) (column-major inside LDS), diagonal inverted; rows
    // and columns beyond nb are an identity so the padded solve is exact.
    __shared__ T Lc[NB][NB];
    __shared__ T Bt[NB][COLS + 1];
    const int lane = threadIdx.x;
    const int64_t col0 = (int64_t)blockIdx.x * COLS;

    for (int j = 0; j < NB; ++j) {
        const int i = lane;
        T v = T(0);
        if (i < nb && j < nb) {
            if (i == j) {
                v = unit_diag ? T(1) : T(1) / a[i * ars + j * acs];
            } else if (LOWER ? (i > j) : (i < j)) {
                v = a[i * ars + j * acs];
            }
        } else if (i == j) {
            v = T(1);
        }
        Lc[j][i] = v;
    }
    // Coalesced load of the nb x COLS slab of B (lanes walk rows when B is
    // column-stored).
    for (int c = 0; c < COLS; ++c) {
        const int64_t col = col0 + c;
        const int r = lane;
        T v = T(0);
        if (r < nb && col < n) v = alpha * b[r * brs + col * bcs];
        Bt[r][c] = v;
    }
    __syncthreads();

    T x[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) x[r] = Bt[r][lane];

    if constexpr (LOWER) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const T xi = x[i] * Lc[i][i];
            x[i] = xi;
#pragma unroll
            for (int r = i + 1; r < NB; ++r) x[r] -= Lc[i][r] * xi;
        }
    } else {
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            const T xi = x[i] * Lc[i][i];
            x[i] = xi;
#pragma unroll
            for (int r = 0; r < i; ++r) x[r] -= Lc[i][r] * xi;
        }
    }

    __syncthreads();
#pragma unroll
    for (int r = 0; r < NB; ++r) Bt[r][lane] = x[r];
    __syncthreads();
    for (int c = 0; c < COLS; ++c) {
        const int64_t col = col0 + c;
        const int r = lane;
        if (r < nb && col < n) b[r * brs + col * bcs] = Bt[r][c];
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

void leaf(const Solve& s, int64_t r0, int64_t nb, double alpha) {
    const size_t es = dt_size(s.dt);
    const dim3 grid((unsigned)((s.n + COLS - 1) / COLS)), block(64);
    void* a = offset(s.a, r0, r0, es);
    void* b = offset(s.b, r0, 0, es);
    if (s.dt == Dt::F64) {
        if (s.lower)
            hipLaunchKernelGGL((trsm_diag_kernel<double, true>), grid, block, 0, s.st,
                               (const double*)a, s.a.rs, s.a.cs, (double*)b, s.b.rs, s.b.cs,
                               (int)nb, s.n, alpha, (int)s.unit);
        else
            hipLaunchKernelGGL((trsm_diag_kernel<double, false>), grid, block, 0, s.st,
                               (const double*)a, s.a.rs, s.a.cs, (double*)b, s.b.rs, s.b.cs,
                               (int)nb, s.n, alpha, (int)s.unit);
    } else {
        if (s.lower)
            hipLaunchKernelGGL((trsm_diag_kernel<float, true>), grid, block, 0, s.st,
                               (const float*)a, s.a.rs, s.a.cs, (float*)b, s.b.rs, s.b.cs,
                               (int)nb, s.n, (float)alpha, (int)s.unit);
        else
            hipLaunchKernelGGL((trsm_diag_kernel<float, false>), grid, block, 0, s.st,
                               (const float*)a, s.a.rs, s.a.cs, (float*)b, s.b.rs, s.b.cs,
                               (int)nb, s.n, (float)alpha, (int)s.unit);
    }
    HIP_CHECK(hipGetLastError());
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
