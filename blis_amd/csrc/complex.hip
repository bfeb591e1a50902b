// complex.hip — complex GEMM by the 1m induced method, and the conversion
// and accumulation kernels that mixed-datatype GEMM needs.
//
// BLIS computes complex GEMM with real-domain microkernels through the 1m
// method (frame/3/bli_l3_ind.c; the packing formats of frame/1m/packm/
// bli_packm_cxk_1er.c): A is packed in the "1e" format, each complex
// element a = ar + i*ai becoming the real 2 x 2 block [[ar, -ai], [ai, ar]],
// and B in the "1r" format, each element becoming the real pair (br, bi)
// down a column. Then the real product of the 2m x 2k matrix A_1e and the
// 2k x n matrix B_1r is the 2m x n real matrix whose rows interleave the
// real and imaginary parts of C = A*B: row 2i is ar*br - ai*bi summed over
// k, row 2i+1 is ai*br + ar*bi. On the GPU:
//   * A_1e is written to workspace by pack_1e_kernel (64 x 16 complex tiles
//     through LDS, read along the source's unit stride, written down the
//     columns of A_1e); conjugation and a complex alpha are applied there;
//   * B_1r needs no copy when op(B) = B is column-stored (its memory read
//     as reals already is B_1r, with column stride 2*ldb); otherwise
//     pack_1r_kernel writes it;
//   * column-stored C read as reals is the interleaved 2m x n result, so the
//     real GEMM (gemm_run, any tile) writes it in place; row-stored C is
//     handled by the operation transpose C^T = op(B)^T op(A)^T, and C with
//     general strides goes through a column-major copy;
//   * a complex beta with a nonzero imaginary part scales C first
//     (cscalm_kernel); otherwise the real GEMM applies real(beta), and
//     beta == 0 does not read C.
#include "internal.hpp"

#include <algorithm>

namespace blis_amd {

namespace {

template <typename R> struct Cplx { R re, im; };

template <typename R>
__device__ inline Cplx<R> cmul(Cplx<R> a, Cplx<R> b) {
    return { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
}

int grid_for(int64_t total) {
    int64_t g = (total + 255) / 256;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

} // namespace

// A_1e (2m x 2k, column-major, ld lde) from op(A) scaled by alpha:
// x = alpha * (conj ? conj(A_eff(i, p)) : A_eff(i, p)), A_eff(i, p) at
// a[i*rs + p*cs] (complex elements), written as the block
// [[x.re, -x.im], [x.im, x.re]] at rows 2i..2i+1, columns 2p..2p+1.
// A workgroup moves 64 x 16 complex tiles through LDS: lanes run along the
// source's unit stride for the reads and along i for the writes.
template <typename R>
__global__ __launch_bounds__(256) void pack_1e_kernel(const Cplx<R>* __restrict__ a, int64_t m,
                                                      int64_t k, int64_t rs, int64_t cs,
                                                      int conj, Cplx<R> alpha,
                                                      R* __restrict__ ae, int64_t lde) {
    __shared__ Cplx<R> t[16][65];
    const int tid = threadIdx.x;
    const int64_t tm = (m + 63) / 64, tk = (k + 15) / 16;
    const bool along_p = (cs == 1 && rs != 1);
    for (int64_t tile = blockIdx.x; tile < tm * tk; tile += gridDim.x) {
        const int64_t i0 = (tile % tm) * 64, p0 = (tile / tm) * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;          // 1024 elements per tile
            const int ii = along_p ? e / 16 : e % 64;
            const int pp = along_p ? e % 16 : e / 64;
            const int64_t i = i0 + ii, p = p0 + pp;
            if (i < m && p < k) {
                Cplx<R> x = a[i * rs + p * cs];
                if (conj) x.im = -x.im;
                t[pp][ii] = cmul(alpha, x);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            const int ii = e % 64, pp = e / 64;
            const int64_t i = i0 + ii, p = p0 + pp;
            if (i < m && p < k) {
                const Cplx<R> x = t[pp][ii];
                R* col0 = ae + (2 * p) * lde + 2 * i;
                R* col1 = col0 + lde;
                col0[0] = x.re; col0[1] = x.im;
                col1[0] = -x.im; col1[1] = x.re;
            }
        }
        __syncthreads();
    }
}

// B_1r (2k x n, column-major, ld ldr): rows 2p and 2p+1 of column j hold the
// real and imaginary parts of (conj ? conj(B_eff(p, j)) : B_eff(p, j)),
// B_eff(p, j) at b[p*rs + j*cs].
template <typename R>
__global__ __launch_bounds__(256) void pack_1r_kernel(const Cplx<R>* __restrict__ b, int64_t k,
                                                      int64_t n, int64_t rs, int64_t cs,
                                                      int conj, R* __restrict__ br, int64_t ldr) {
    __shared__ Cplx<R> t[16][65];
    const int tid = threadIdx.x;
    const int64_t tk = (k + 63) / 64, tn = (n + 15) / 16;
    const bool along_j = (cs == 1 && rs != 1);
    for (int64_t tile = blockIdx.x; tile < tk * tn; tile += gridDim.x) {
        const int64_t p0 = (tile % tk) * 64, j0 = (tile / tk) * 16;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            const int pp = along_j ? e / 16 : e % 64;
            const int jj = along_j ? e % 16 : e / 64;
            const int64_t p = p0 + pp, j = j0 + jj;
            if (p < k && j < n) {
                Cplx<R> x = b[p * rs + j * cs];
                if (conj) x.im = -x.im;
                t[jj][pp] = x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            const int pp = e % 64, jj = e / 64;
            const int64_t p = p0 + pp, j = j0 + jj;
            if (p < k && j < n) {
                const Cplx<R> x = t[jj][pp];
                R* dst = br + j * ldr + 2 * p;
                dst[0] = x.re; dst[1] = x.im;
            }
        }
        __syncthreads();
    }
}

// C := beta*C for complex C (element (i, j) at c[i*rs + j*cs]); beta == 0
// stores zeros without reading C.
template <typename R>
__global__ void cscalm_kernel(Cplx<R>* c, int64_t m, int64_t n, int64_t rs, int64_t cs,
                              Cplx<R> beta) {
    const int64_t total = m * n;
    const bool zero = beta.re == R(0) && beta.im == R(0);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        Cplx<R>* d = c + i * rs + j * cs;
        *d = zero ? Cplx<R>{ R(0), R(0) } : cmul(beta, *d);
    }
}

// dst(i, j) := src(i, j) for complex matrices of any strides.
template <typename R>
__global__ void ccopy_kernel(const Cplx<R>* src, int64_t m, int64_t n, int64_t srs,
                             int64_t scs, Cplx<R>* dst, int64_t drs, int64_t dcs) {
    const int64_t total = m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        dst[i * drs + j * dcs] = src[i * srs + j * scs];
    }
}

// Mixed-datatype element conversion (the typecast BLIS folds into packing,
// docs/MixedDatatypes.md): dst(i, j) := src(i, j) in dst's type; a complex
// source written to a real destination keeps its real part, a real source
// written to a complex destination gets a zero imaginary part. src strides
// are in elements of its own type; dst is column-major with ld ldd.
template <typename TS, typename TD, bool SCX, bool DCX>
__global__ void md_convert_kernel(const TS* src, int64_t m, int64_t n, int64_t rs, int64_t cs,
                                  TD* dst, int64_t ldd) {
    const int64_t total = m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        const int64_t e = i * rs + j * cs;
        const TS re = SCX ? src[2 * e] : src[e];
        const TS im = SCX ? src[2 * e + 1] : TS(0);
        const int64_t d = i + j * ldd;
        if (DCX) { dst[2 * d] = (TD)re; dst[2 * d + 1] = (TD)im; }
        else dst[d] = (TD)re;
    }
}

// C := beta*C + P for P (column-major, ld ldp) in the computation type. See
// md_accum_run for the domain rules. beta == 0 does not read C.
template <typename TP, typename TC, bool PCX, bool CCX>
__global__ void md_accum_kernel(const TP* p, int64_t ldp, int64_t m, int64_t n, TC br, TC bi,
                                TC* c, int64_t rs, int64_t cs) {
    const int64_t total = m * n;
    const bool zero = br == TC(0) && bi == TC(0);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t % m, j = t / m;
        const int64_t e = i * rs + j * cs, q = i + j * ldp;
        const TC pr = (TC)(PCX ? p[2 * q] : p[q]);
        if (CCX && PCX) {
            const TC pi = (TC)p[2 * q + 1];
            TC* d = c + 2 * e;
            if (zero) { d[0] = pr; d[1] = pi; }
            else {
                const TC cr = d[0], ci = d[1];
                d[0] = br * cr - bi * ci + pr;
                d[1] = br * ci + bi * cr + pi;
            }
        } else {
            TC* d = c + (CCX ? 2 * e : e);
            *d = (br == TC(0)) ? pr : br * *d + pr;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

template <typename R>
void launch_cscalm(void* c, int64_t m, int64_t n, int64_t rs, int64_t cs, double br, double bi,
                   hipStream_t st) {
    hipLaunchKernelGGL(cscalm_kernel<R>, dim3(grid_for(m * n)), dim3(256), 0, st,
                       (Cplx<R>*)c, m, n, rs, cs, Cplx<R>{ (R)br, (R)bi });
    HIP_CHECK(hipGetLastError());
}

template <typename R>
void launch_ccopy(const void* src, int64_t m, int64_t n, int64_t srs, int64_t scs, void* dst,
                  int64_t drs, int64_t dcs, hipStream_t st) {
    hipLaunchKernelGGL(ccopy_kernel<R>, dim3(grid_for(m * n)), dim3(256), 0, st,
                       (const Cplx<R>*)src, m, n, srs, scs, (Cplx<R>*)dst, drs, dcs);
    HIP_CHECK(hipGetLastError());
}

// Column-stored C (rs == 1): the 1m product into C read as reals.
template <typename R>
void zgemm_colstored(int64_t m, int64_t n, int64_t k, double ar, double ai, View a, bool conja,
                     View b, bool conjb, double br, double bi, View c, hipStream_t st) {
    const Dt dt = std::is_same<R, double>::value ? Dt::F64 : Dt::F32;
    if (bi != 0.0) {
        launch_cscalm<R>(c.ptr, m, n, c.rs, c.cs, br, bi, st);
        br = 1.0;
    }
    // alpha: a real alpha stays in the real GEMM; a complex one is folded
    // into A_1e.
    const bool fold = ai != 0.0;
    const Cplx<R> alpha_pack = fold ? Cplx<R>{ (R)ar, (R)ai } : Cplx<R>{ R(1), R(0) };
    const int64_t lde = (2 * m + 3) / 4 * 4;
    R* ae = (R*)workspace((size_t)lde * 2 * k * sizeof(R), 22);
    {
        const int64_t tiles = ((m + 63) / 64) * ((k + 15) / 16);
        const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 16384));
        hipLaunchKernelGGL(pack_1e_kernel<R>, dim3(g), dim3(256), 0, st, (const Cplx<R>*)a.ptr,
                           m, k, a.rs, a.cs, conja ? 1 : 0, alpha_pack, ae, lde);
        HIP_CHECK(hipGetLastError());
    }
    View b1r;
    if (!conjb && b.rs == 1) {
        b1r = View{ b.ptr, 1, 2 * b.cs };          // B's memory read as reals is B_1r
    } else {
        const int64_t ldr = (2 * k + 3) / 4 * 4;
        R* w = (R*)workspace((size_t)ldr * n * sizeof(R), 23);
        const int64_t tiles = ((k + 63) / 64) * ((n + 15) / 16);
        const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 16384));
        hipLaunchKernelGGL(pack_1r_kernel<R>, dim3(g), dim3(256), 0, st, (const Cplx<R>*)b.ptr,
                           k, n, b.rs, b.cs, conjb ? 1 : 0, w, ldr);
        HIP_CHECK(hipGetLastError());
        b1r = View{ w, 1, ldr };
    }
    GemmDesc d;
    d.dt = dt;
    d.m = 2 * m; d.n = n; d.k = 2 * k;
    d.a = View{ ae, 1, lde };
    d.b = b1r;
    d.c = View{ c.ptr, 1, 2 * c.cs };
    d.alpha = fold ? 1.0 : ar;
    d.beta = br;
    gemm_run(d, nullptr, st);
}

template <typename R>
void zgemm_impl(const ZGemmDesc& z, hipStream_t st) {
    const int64_t m = z.m, n = z.n, k = z.k;
    if (m <= 0 || n <= 0) return;
    if (k <= 0 || (z.alpha_re == 0.0 && z.alpha_im == 0.0)) {
        launch_cscalm<R>(z.c.ptr, m, n, z.c.rs, z.c.cs, z.beta_re, z.beta_im, st);
        return;
    }
    if (z.c.rs == 1) {
        zgemm_colstored<R>(m, n, k, z.alpha_re, z.alpha_im, z.a, z.conja, z.b, z.conjb,
                           z.beta_re, z.beta_im, z.c, st);
    } else if (z.c.cs == 1) {
        // C^T = op(B)^T op(A)^T: transposition swaps the operands (and their
        // strides); conjugation stays with each operand.
        zgemm_colstored<R>(n, m, k, z.alpha_re, z.alpha_im, View{ z.b.ptr, z.b.cs, z.b.rs },
                           z.conjb, View{ z.a.ptr, z.a.cs, z.a.rs }, z.conja, z.beta_re,
                           z.beta_im, View{ z.c.ptr, 1, z.c.rs }, st);
    } else {
        // General strides: through a column-major copy of C.
        void* w = workspace((size_t)m * n * 2 * sizeof(R), 24);
        if (z.beta_re != 0.0 || z.beta_im != 0.0)
            launch_ccopy<R>(z.c.ptr, m, n, z.c.rs, z.c.cs, w, 1, m, st);
        zgemm_colstored<R>(m, n, k, z.alpha_re, z.alpha_im, z.a, z.conja, z.b, z.conjb,
                           z.beta_re, z.beta_im, View{ w, 1, m }, st);
        launch_ccopy<R>(w, m, n, 1, m, z.c.ptr, z.c.rs, z.c.cs, st);
    }
}

} // namespace

void zgemm_run(const ZGemmDesc& z, hipStream_t st) {
    if (z.dt == Dt::F64) zgemm_impl<double>(z, st);
    else zgemm_impl<float>(z, st);
}

void zscalm_run(Dt dt, int64_t m, int64_t n, double br, double bi, View c, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    if (dt == Dt::F64) launch_cscalm<double>(c.ptr, m, n, c.rs, c.cs, br, bi, st);
    else launch_cscalm<float>(c.ptr, m, n, c.rs, c.cs, br, bi, st);
}

template <typename TS, typename TD>
static void md_convert_t(bool scx, const void* src, int64_t m, int64_t n, int64_t rs,
                         int64_t cs, bool dcx, void* dst, int64_t ldd, hipStream_t st) {
    const dim3 g(grid_for(m * n)), b(256);
    const TS* s = (const TS*)src;
    TD* d = (TD*)dst;
    if (scx && dcx) hipLaunchKernelGGL((md_convert_kernel<TS, TD, true, true>), g, b, 0, st, s, m, n, rs, cs, d, ldd);
    else if (scx) hipLaunchKernelGGL((md_convert_kernel<TS, TD, true, false>), g, b, 0, st, s, m, n, rs, cs, d, ldd);
    else if (dcx) hipLaunchKernelGGL((md_convert_kernel<TS, TD, false, true>), g, b, 0, st, s, m, n, rs, cs, d, ldd);
    else hipLaunchKernelGGL((md_convert_kernel<TS, TD, false, false>), g, b, 0, st, s, m, n, rs, cs, d, ldd);
    HIP_CHECK(hipGetLastError());
}

void md_convert_run(Dt sdt, bool src_cx, const void* src, int64_t m, int64_t n, int64_t rs,
                    int64_t cs, Dt ddt, bool dst_cx, void* dst, int64_t ldd, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    if (sdt == Dt::F64 && ddt == Dt::F64)
        md_convert_t<double, double>(src_cx, src, m, n, rs, cs, dst_cx, dst, ldd, st);
    else if (sdt == Dt::F64)
        md_convert_t<double, float>(src_cx, src, m, n, rs, cs, dst_cx, dst, ldd, st);
    else if (ddt == Dt::F64)
        md_convert_t<float, double>(src_cx, src, m, n, rs, cs, dst_cx, dst, ldd, st);
    else
        md_convert_t<float, float>(src_cx, src, m, n, rs, cs, dst_cx, dst, ldd, st);
}

template <typename TP, typename TC>
static void md_accum_t(bool pcx, const void* p, int64_t ldp, int64_t m, int64_t n, double br,
                       double bi, bool ccx, View c, hipStream_t st) {
    const dim3 g(grid_for(m * n)), b(256);
    const TP* pp = (const TP*)p;
    TC* cp = (TC*)c.ptr;
    const TC r = (TC)br, i = (TC)bi;
    if (pcx && ccx) hipLaunchKernelGGL((md_accum_kernel<TP, TC, true, true>), g, b, 0, st, pp, ldp, m, n, r, i, cp, c.rs, c.cs);
    else if (pcx) hipLaunchKernelGGL((md_accum_kernel<TP, TC, true, false>), g, b, 0, st, pp, ldp, m, n, r, TC(0), cp, c.rs, c.cs);
    else if (ccx) hipLaunchKernelGGL((md_accum_kernel<TP, TC, false, true>), g, b, 0, st, pp, ldp, m, n, r, TC(0), cp, c.rs, c.cs);
    else hipLaunchKernelGGL((md_accum_kernel<TP, TC, false, false>), g, b, 0, st, pp, ldp, m, n, r, TC(0), cp, c.rs, c.cs);
    HIP_CHECK(hipGetLastError());
}

void md_accum_run(Dt pdt, bool p_cx, const void* p, int64_t ldp, int64_t m, int64_t n,
                  double br, double bi, Dt cdt, bool c_cx, View c, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    if (pdt == Dt::F64 && cdt == Dt::F64) md_accum_t<double, double>(p_cx, p, ldp, m, n, br, bi, c_cx, c, st);
    else if (pdt == Dt::F64) md_accum_t<double, float>(p_cx, p, ldp, m, n, br, bi, c_cx, c, st);
    else if (cdt == Dt::F64) md_accum_t<float, double>(p_cx, p, ldp, m, n, br, bi, c_cx, c, st);
    else md_accum_t<float, float>(p_cx, p, ldp, m, n, br, bi, c_cx, c, st);
}

} // namespace blis_amd
