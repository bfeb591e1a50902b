// level12.hip — bandwidth-bound level-1v/level-2 kernels for gfx950:
// axpyv, dotv, gemv (and the scalv used by gemv's early exits).
//
// Reference semantics (paths relative to the BLIS tree):
//   axpyv  y := y + alpha*x             ref_kernels/1/bli_axpyv_ref.c:38-121
//   dotv   rho := sum x_i*y_i           ref_kernels/1/bli_dotv_ref.c:38-122
//   gemv   y := beta*y + alpha*op(A)*x  frame/2/bli_l2_tapi.c:43-123 with
//          var2 (axpyf over column groups, frame/2/gemv/bli_gemv_unf_var2.c)
//          and var1 (dotxf over row groups, frame/2/gemv/bli_gemv_unf_var1.c)
//   scalv with beta == 0 stores zeros without reading y.
//
// All of these move each byte of their operands once, so the bound is HBM
// bandwidth. The kernels use 16-byte vector loads where the stride and
// alignment allow it, grid-stride loops sized to fill all CUs, and, for the
// reductions (dotv, gemv with many column chunks), a fixed-order two-pass
// reduction so results are deterministic from run to run.
#include "internal.hpp"

namespace blis_amd {

namespace detail_level12 {

constexpr int TPB = 256;

template <typename T> struct V16;
template <> struct V16<double> { typedef double type __attribute__((ext_vector_type(2))); static constexpr int N = 2; };
template <> struct V16<float>  { typedef float  type __attribute__((ext_vector_type(4))); static constexpr int N = 4; };

inline bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int blocks_for(int64_t work_items, int per_thread) {
    int64_t g = (work_items + (int64_t)TPB * per_thread - 1) / ((int64_t)TPB * per_thread);
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (int)g;
}

// ---- axpyv -----------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(TPB) void axpyv_vec_kernel(int64_t n, T alpha, const T* __restrict__ x,
                                                        T* __restrict__ y) {
    using vt = typename V16<T>::type;
    constexpr int V = V16<T>::N;
    constexpr int RUN = 4 * TPB;   // vectors per workgroup per pass
    const int64_t nv = n / V;
    const vt* xv = reinterpret_cast<const vt*>(x);
    vt* yv = reinterpret_cast<vt*>(y);
    // Each workgroup takes contiguous runs of RUN vectors (16 KiB of x and
    // of y), four independent 16-byte loads per operand in flight per
    // thread. Measured at n = 2^28 (FP64): 5.45 TB/s, against 5.15 TB/s
    // with each thread's four vectors a grid apart.
    for (int64_t base = blockIdx.x * (int64_t)RUN; base < nv; base += (int64_t)gridDim.x * RUN) {
        const int64_t i = base + threadIdx.x;
        if (base + RUN <= nv) {
            const vt a0 = xv[i], a1 = xv[i + TPB], a2 = xv[i + 2 * TPB], a3 = xv[i + 3 * TPB];
            const vt b0 = yv[i], b1 = yv[i + TPB], b2 = yv[i + 2 * TPB], b3 = yv[i + 3 * TPB];
            yv[i] = b0 + alpha * a0;
            yv[i + TPB] = b1 + alpha * a1;
            yv[i + 2 * TPB] = b2 + alpha * a2;
            yv[i + 3 * TPB] = b3 + alpha * a3;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (i + q * TPB < nv) yv[i + q * TPB] = yv[i + q * TPB] + alpha * xv[i + q * TPB];
        }
    }
    // Tail elements.
    const int64_t t = nv * V + blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (blockIdx.x == 0 && t < n) y[t] = y[t] + alpha * x[t];
}

template <typename T>
__global__ __launch_bounds__(TPB) void axpyv_gen_kernel(int64_t n, T alpha, const T* x, int64_t incx,
                                                        T* y, int64_t incy) {
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * TPB) {
        y[i * incy] = y[i * incy] + alpha * x[i * incx];
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void scalv_kernel(int64_t n, T beta, T* y, int64_t incy) {
    for (int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * TPB) {
        T* p = y + i * incy;
        *p = (beta == T(0)) ? T(0) : beta * *p;
    }
}

// ---- dotv -------------------------------------------------------------------

template <typename T>
__device__ inline T block_reduce(T v, T* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    T r = T(0);
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    }
    return r;
}

// Pass 1: block b reduces a fixed, contiguous chunk of the vectors.
template <typename T>
__global__ __launch_bounds__(TPB) void dotv_partial_kernel(int64_t n, const T* __restrict__ x,
                                                           int64_t incx, const T* __restrict__ y,
                                                           int64_t incy, int64_t chunk,
                                                           int vec, T* __restrict__ part) {
    __shared__ T sh[TPB / 64];
    const int64_t lo = blockIdx.x * chunk;
    const int64_t hi = (lo + chunk < n) ? lo + chunk : n;
    T acc0 = T(0), acc1 = T(0);
    if (vec) {
        using vt = typename V16<T>::type;
        constexpr int V = V16<T>::N;
        const vt* xv = reinterpret_cast<const vt*>(x + lo);
        const vt* yv = reinterpret_cast<const vt*>(y + lo);
        const int64_t nv = (hi - lo) / V;
        int64_t i = threadIdx.x;
        for (; i + TPB < nv; i += 2 * TPB) {
            vt a0 = xv[i], b0 = yv[i], a1 = xv[i + TPB], b1 = yv[i + TPB];
#pragma unroll
            for (int e = 0; e < V; ++e) { acc0 += a0[e] * b0[e]; acc1 += a1[e] * b1[e]; }
        }
        for (; i < nv; i += TPB) {
            vt a0 = xv[i], b0 = yv[i];
#pragma unroll
            for (int e = 0; e < V; ++e) acc0 += a0[e] * b0[e];
        }
        for (int64_t t = lo + nv * V + threadIdx.x; t < hi; t += TPB) acc1 += x[t] * y[t];
    } else {
        for (int64_t t = lo + threadIdx.x; t < hi; t += TPB) acc0 += x[t * incx] * y[t * incy];
    }
    const T r = block_reduce(acc0 + acc1, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <typename T>
__global__ __launch_bounds__(TPB) void dotv_final_kernel(int nparts, const T* __restrict__ part,
                                                         T* __restrict__ rho) {
    __shared__ T sh[TPB / 64];
    T acc = T(0);
    for (int i = threadIdx.x; i < nparts; i += TPB) acc += part[i];
    const T r = block_reduce(acc, sh);
    if (threadIdx.x == 0) *rho = r;
}

// ---- gemv, A column-stored (rs == 1): y := beta*y + alpha*A*x --------------
//
// Workgroup (rb, cb) owns ROWS rows and the columns of chunk cb. On the
// vector path each thread keeps two vectors of V consecutive rows, TPB * V
// rows apart, and streams the chunk's columns with 16-byte loads (one row per
// thread otherwise); x is staged in LDS GEMV_CW entries at a time. With a
// single chunk the result goes straight to y; otherwise partial sums go to
// part[cb][row] and a second kernel adds them in a fixed order.
constexpr int GEMV_CW = 256;

template <typename T, bool VEC>
__global__ __launch_bounds__(TPB) void gemv_n_kernel(int64_t m, int64_t n, const T* __restrict__ a,
                                                     int64_t lda, const T* __restrict__ x,
                                                     int64_t incx, T alpha, T beta, T* y,
                                                     int64_t incy, int64_t cols_per_chunk,
                                                     T* __restrict__ part) {
    constexpr int V = VEC ? V16<T>::N : 1;
    // Row vectors per thread, TPB * V rows apart: two give each thread eight
    // 16-byte loads in flight per four columns.
    constexpr int RV = VEC ? 2 : 1;
    constexpr int RS = TPB * V;               // rows between a thread's vectors
    constexpr int ROWS = RS * RV;
    __shared__ T xs[GEMV_CW];
    const int64_t rb = blockIdx.x * (int64_t)ROWS + threadIdx.x * V;
    const int64_t cb = blockIdx.y * cols_per_chunk;
    const int64_t ce = (cb + cols_per_chunk < n) ? cb + cols_per_chunk : n;
    T acc[RV][V];
#pragma unroll
    for (int r = 0; r < RV; ++r)
#pragma unroll
        for (int e = 0; e < V; ++e) acc[r][e] = T(0);
    for (int64_t c0 = cb; c0 < ce; c0 += GEMV_CW) {
        const int64_t c1 = (c0 + GEMV_CW < ce) ? c0 + GEMV_CW : ce;
        __syncthreads();
        for (int j = threadIdx.x; j < GEMV_CW; j += TPB) {
            const int64_t col = c0 + j;
            xs[j] = (col < c1) ? x[col * incx] : T(0);
        }
        __syncthreads();
        if (rb >= m) continue;
        if (VEC && rb + (RV - 1) * RS + V <= m) {
            using vt = typename V16<T>::type;
            const T* p = a + rb + c0 * lda;
            int64_t j = c0;
            for (; j + 3 < c1; j += 4, p += 4 * lda) {
                vt v[RV][4];
#pragma unroll
                for (int r = 0; r < RV; ++r)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        v[r][q] = *reinterpret_cast<const vt*>(p + r * RS + q * lda);
                const T x0 = xs[j - c0], x1 = xs[j - c0 + 1], x2 = xs[j - c0 + 2], x3 = xs[j - c0 + 3];
#pragma unroll
                for (int r = 0; r < RV; ++r)
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        acc[r][e] += v[r][0][e] * x0 + v[r][1][e] * x1 + v[r][2][e] * x2 + v[r][3][e] * x3;
            }
            for (; j < c1; ++j, p += lda) {
#pragma unroll
                for (int r = 0; r < RV; ++r) {
                    const vt v0 = *reinterpret_cast<const vt*>(p + r * RS);
#pragma unroll
                    for (int e = 0; e < V; ++e) acc[r][e] += v0[e] * xs[j - c0];
                }
            }
        } else {
            for (int64_t j = c0; j < c1; ++j) {
#pragma unroll
                for (int r = 0; r < RV; ++r)
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        const int64_t row = rb + r * RS + e;
                        if (row < m) acc[r][e] += a[row + j * lda] * xs[j - c0];
                    }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RV; ++r)
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int64_t row = rb + r * RS + e;
            if (row >= m) continue;
            if (gridDim.y == 1) {
                T* py = y + row * incy;
                T o = alpha * acc[r][e];
                if (beta != T(0)) o += beta * *py;
                *py = o;
            } else {
                part[blockIdx.y * m + row] = acc[r][e];
            }
        }
}

// Sum of the column chunks' partial results, 64 rows per workgroup: the
// four waves each add a quarter of the chunks in order, and the quarters
// are added in order (a fixed order, so the result is the same on every
// run). Measured for m = 16384 and 64 chunks: 5.4 us against 7.9 us with
// one thread per row over all chunks.
template <typename T>
__global__ __launch_bounds__(TPB) void gemv_n_reduce_kernel(int64_t m, int nchunks,
                                                            const T* __restrict__ part, T alpha,
                                                            T beta, T* y, int64_t incy) {
    static_assert(TPB == 256, "four waves of 64 rows");
    __shared__ T sq[4][64];
    const int lr = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int per = (nchunks + 3) / 4;
    const int c0 = q * per, c1 = min(nchunks, c0 + per);
    for (int64_t r0 = blockIdx.x * (int64_t)64; r0 < m; r0 += (int64_t)gridDim.x * 64) {
        const int64_t r = r0 + lr;
        T s = T(0);
        if (r < m)
            for (int c = c0; c < c1; ++c) s += part[c * m + r];
        sq[q][lr] = s;
        __syncthreads();
        if (q == 0 && r < m) {
            const T sum = ((sq[0][lr] + sq[1][lr]) + sq[2][lr]) + sq[3][lr];
            T* py = y + r * incy;
            T o = alpha * sum;
            if (beta != T(0)) o += beta * *py;
            *py = o;
        }
        __syncthreads();   // sq is rewritten by the next row group
    }
}

// ---- gemv, A^T with A column-stored: y_j := beta*y_j + alpha*A(:,j).x ------
// One wave per column of A (dot product along the contiguous dimension).
template <typename T, bool VEC>
__global__ __launch_bounds__(TPB) void gemv_t_kernel(int64_t m, int64_t n, const T* __restrict__ a,
                                                     int64_t lda, const T* __restrict__ x,
                                                     int64_t incx, T alpha, T beta, T* y,
                                                     int64_t incy) {
    const int lane = threadIdx.x & 63;
    const int64_t col = blockIdx.x * (int64_t)(TPB / 64) + (threadIdx.x >> 6);
    if (col >= n) return;
    const T* p = a + col * lda;
    T acc0 = T(0), acc1 = T(0);
    if (VEC) {
        using vt = typename V16<T>::type;
        constexpr int V = V16<T>::N;
        const int64_t nv = m / V;
        const vt* pv = reinterpret_cast<const vt*>(p);
        const vt* xv = reinterpret_cast<const vt*>(x);
        int64_t i = lane;
        // Four 16-byte loads of the column in flight per lane: 0.356 ms at
        // 16384^2 (FP64) against 0.370 ms with two, and 0.373 ms with eight.
        T acc2 = T(0), acc3 = T(0);
        for (; i + 192 < nv; i += 256) {
            const vt a0 = pv[i], a1 = pv[i + 64], a2 = pv[i + 128], a3 = pv[i + 192];
            const vt x0 = xv[i], x1 = xv[i + 64], x2 = xv[i + 128], x3 = xv[i + 192];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                acc0 += a0[e] * x0[e];
                acc1 += a1[e] * x1[e];
                acc2 += a2[e] * x2[e];
                acc3 += a3[e] * x3[e];
            }
        }
        acc0 += acc2;
        acc1 += acc3;
        for (; i < nv; i += 64) {
            vt a0 = pv[i], x0 = xv[i];
#pragma unroll
            for (int e = 0; e < V; ++e) acc0 += a0[e] * x0[e];
        }
        for (int64_t t = nv * V + lane; t < m; t += 64) acc1 += p[t] * x[t];
    } else {
        for (int64_t t = lane; t < m; t += 64) acc0 += p[t] * x[t * incx];
    }
    T v = acc0 + acc1;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) {
        T* py = y + col * incy;
        T o = alpha * v;
        if (beta != T(0)) o += beta * *py;
        *py = o;
    }
}

// General strides: one thread per output element.
template <typename T>
__global__ __launch_bounds__(TPB) void gemv_gen_kernel(int64_t my, int64_t nx, const T* a,
                                                       int64_t s_out, int64_t s_in, const T* x,
                                                       int64_t incx, T alpha, T beta, T* y,
                                                       int64_t incy) {
    for (int64_t r = blockIdx.x * (int64_t)TPB + threadIdx.x; r < my;
         r += (int64_t)gridDim.x * TPB) {
        T s = T(0);
        for (int64_t j = 0; j < nx; ++j) s += a[r * s_out + j * s_in] * x[j * incx];
        T* py = y + r * incy;
        T o = alpha * s;
        if (beta != T(0)) o += beta * *py;
        *py = o;
    }
}

template <typename T>
void axpyv_t(int64_t n, double alpha, const void* x, int64_t incx, void* y, int64_t incy,
             hipStream_t st) {
    constexpr int V = V16<T>::N;
    const bool vec = incx == 1 && incy == 1 && al16(x) && al16(y);
    if (vec) {
        // Up to 65536 workgroups of four 16-byte vectors per thread: fewer
        // workgroups looping over more runs measured slower (5.33 TB/s with
        // 16384, 5.14 with 4096, at n = 2^28).
        const int g = (int)std::max<int64_t>(
            1, std::min<int64_t>((n / V + 1 + 4 * TPB - 1) / (4 * TPB), 65536));
        hipLaunchKernelGGL(axpyv_vec_kernel<T>, dim3(g), dim3(TPB), 0, st, n, (T)alpha,
                           (const T*)x, (T*)y);
    } else {
        const int g = blocks_for(n, 4);
        hipLaunchKernelGGL(axpyv_gen_kernel<T>, dim3(g), dim3(TPB), 0, st, n, (T)alpha,
                           (const T*)x, incx, (T*)y, incy);
    }
    HIP_CHECK(hipGetLastError());
}

template <typename T>
void dotv_t(int64_t n, const void* x, int64_t incx, const void* y, int64_t incy, void* rho,
            hipStream_t st) {
    constexpr int V = V16<T>::N;
    // Fixed decomposition: chunk size depends only on n, so the summation
    // order (and the result) is reproducible.
    int nparts = (int)((n + 65535) / 65536);
    if (nparts < 1) nparts = 1;
    if (nparts > 2048) nparts = 2048;
    int64_t chunk = (n + nparts - 1) / nparts;
    chunk = (chunk + 7) / 8 * 8;   // keep chunk starts 64-byte aligned
    if (chunk < 8) chunk = 8;      // n == 0: one empty chunk writes rho = 0
    nparts = (int)((n + chunk - 1) / chunk);
    if (nparts < 1) nparts = 1;
    const bool vec = incx == 1 && incy == 1 && al16(x) && al16(y) && (chunk % V == 0);
    T* part = (T*)workspace(sizeof(T) * (size_t)nparts, 6);
    hipLaunchKernelGGL(dotv_partial_kernel<T>, dim3(nparts), dim3(TPB), 0, st, n, (const T*)x,
                       incx, (const T*)y, incy, chunk, (int)vec, part);
    hipLaunchKernelGGL(dotv_final_kernel<T>, dim3(1), dim3(TPB), 0, st, nparts,
                       (const T*)part, (T*)rho);
    HIP_CHECK(hipGetLastError());
}

template <typename T>
void gemv_t_(trans_t trans, int64_t m, int64_t n, double alpha, View a, const void* x,
             int64_t incx, double beta, void* y, int64_t incy, hipStream_t st) {
    constexpr int V = V16<T>::N;
    const bool t = ((unsigned)trans & BLIS_TRANS_BIT) != 0;
    // y := beta*y + alpha * op(A) * x with op(A) of size m_y x n_x. Express
    // op(A) as "rows of length n_x": op(A)(r, j) = a[r*s_out + j*s_in].
    const int64_t my = t ? n : m, nx = t ? m : n;
    const int64_t s_out = t ? a.cs : a.rs, s_in = t ? a.rs : a.cs;
    if (s_out == 1) {
        // Output index is the contiguous one: column-stored axpyf style.
        const int64_t lda = s_in;
        const bool vec = al16(a.ptr) && (lda % V == 0);
        // Row blocks of two vectors per thread (vector path) or one row per
        // thread. The columns are split into chunks only as far as needed to
        // give the GPU about 1024 (vector) or 2048 workgroups. Measured at
        // 16384^2 (FP64, main kernel): 0.359 ms with two vectors per thread
        // in 1024 workgroups, against 0.379 ms with one vector per thread in
        // 2048 and 0.404 ms with two vectors per thread in 2048.
        const int rows_per_block = vec ? 2 * TPB * V : TPB;
        const int64_t gx = (my + rows_per_block - 1) / rows_per_block;
        const int64_t max_chunks = (nx + GEMV_CW - 1) / GEMV_CW;
        const int64_t target = vec ? 1024 : 2048;
        int64_t gy = (target + gx - 1) / gx;
        if (gy > max_chunks) gy = max_chunks;
        if (gy < 1) gy = 1;
        int64_t cpc = (nx + gy - 1) / gy;
        cpc = (cpc + GEMV_CW - 1) / GEMV_CW * GEMV_CW;
        gy = (nx + cpc - 1) / cpc;
        T* part = nullptr;
        if (gy > 1) part = (T*)workspace(sizeof(T) * (size_t)gy * my, 5);
        if (vec)
            hipLaunchKernelGGL((gemv_n_kernel<T, true>), dim3((unsigned)gx, (unsigned)gy),
                               dim3(TPB), 0, st, my, nx, (const T*)a.ptr, lda, (const T*)x,
                               incx, (T)alpha, (T)beta, (T*)y, incy, cpc, part);
        else
            hipLaunchKernelGGL((gemv_n_kernel<T, false>), dim3((unsigned)gx, (unsigned)gy),
                               dim3(TPB), 0, st, my, nx, (const T*)a.ptr, lda, (const T*)x,
                               incx, (T)alpha, (T)beta, (T*)y, incy, cpc, part);
        if (gy > 1) {
            const int g = (int)std::min<int64_t>((my + 63) / 64, 65536);
            hipLaunchKernelGGL(gemv_n_reduce_kernel<T>, dim3(g), dim3(TPB), 0, st, my,
                               (int)gy, (const T*)part, (T)alpha, (T)beta, (T*)y, incy);
        }
    } else if (s_in == 1) {
        // Reduction index is contiguous: dotxf style, one wave per output.
        const int64_t lda = s_out;
        const bool vec = al16(a.ptr) && (lda % V == 0) && incx == 1 && al16(x);
        const unsigned g = (unsigned)((my + (TPB / 64) - 1) / (TPB / 64));
        if (vec)
            hipLaunchKernelGGL((gemv_t_kernel<T, true>), dim3(g), dim3(TPB), 0, st, nx, my,
                               (const T*)a.ptr, lda, (const T*)x, incx, (T)alpha, (T)beta,
                               (T*)y, incy);
        else
            hipLaunchKernelGGL((gemv_t_kernel<T, false>), dim3(g), dim3(TPB), 0, st, nx, my,
                               (const T*)a.ptr, lda, (const T*)x, incx, (T)alpha, (T)beta,
                               (T*)y, incy);
    } else {
        const int g = blocks_for(my, 1);
        hipLaunchKernelGGL(gemv_gen_kernel<T>, dim3(g), dim3(TPB), 0, st, my, nx,
                           (const T*)a.ptr, s_out, s_in, (const T*)x, incx, (T)alpha, (T)beta,
                           (T*)y, incy);
    }
    HIP_CHECK(hipGetLastError());
}

} // namespace detail_level12
using namespace detail_level12;

void axpyv_run(Dt dt, int64_t n, double alpha, const void* x, int64_t incx, void* y,
               int64_t incy, hipStream_t st) {
    if (n <= 0) return;
    if (dt == Dt::F64) axpyv_t<double>(n, alpha, x, incx, y, incy, st);
    else axpyv_t<float>(n, alpha, x, incx, y, incy, st);
}

void dotv_run(Dt dt, int64_t n, const void* x, int64_t incx, const void* y, int64_t incy,
              void* rho, hipStream_t st) {
    if (dt == Dt::F64) dotv_t<double>(n, x, incx, y, incy, rho, st);
    else dotv_t<float>(n, x, incx, y, incy, rho, st);
}

void scalv_run(Dt dt, int64_t n, double beta, void* y, int64_t incy, hipStream_t st) {
    if (n <= 0) return;
    const int g = blocks_for(n, 1);
    if (dt == Dt::F64)
        hipLaunchKernelGGL(scalv_kernel<double>, dim3(g), dim3(TPB), 0, st, n, beta, (double*)y,
                           incy);
    else
        hipLaunchKernelGGL(scalv_kernel<float>, dim3(g), dim3(TPB), 0, st, n, (float)beta,
                           (float*)y, incy);
    HIP_CHECK(hipGetLastError());
}

void gemv_run(Dt dt, trans_t trans, int64_t m, int64_t n, double alpha, View a, const void* x,
              int64_t incx, double beta, void* y, int64_t incy, hipStream_t st) {
    if (dt == Dt::F64) gemv_t_<double>(trans, m, n, alpha, a, x, incx, beta, y, incy, st);
    else gemv_t_<float>(trans, m, n, alpha, a, x, incx, beta, y, incy, st);
}

} // namespace blis_amd
