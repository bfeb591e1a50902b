// split.hip — split-precision GEMM: C := alpha*op(A)*op(B) + beta*C with the
// arithmetic done as exact INT8 slice products on the MFMA cores and the
// slice products summed in FP64.
//
// The reference has no such path. Each result must match what its FP64 (FP32)
// GEMM computes (ref_kernels/3/bli_gemm_ref.c:163-317) to within native
// precision. The method is the Ozaki scheme with integer slices: Ozaki, Ogita,
// Oishi, Rump, "Error-free transformations of matrix multiplication by using
// fast routines of matrix multiplication and its applications", Numer.
// Algorithms 59 (2012), in the integer form of Ootomo, Ozaki, Yokota, "DGEMM
// on integer matrix multiplication unit", IJHPCA (2024):
//
//   1. row i of op(A) is scaled by 2^-e_i, where e_i is the binary exponent
//      of its largest magnitude (max = f 2^e_i, f in [0.5, 1)). It is then
//      split exactly into s signed 7-bit digits:
//        a_ip 2^-e_i = sum_{t=1..s} A_t[i][p] 2^-7t + r,   |A_t| <= 127,   |r| < 2^-7s.
//      Column j of op(B) is split the same way (exponent f_j, digits B_u);
//   2. all products of weight d = t + u <= s + 1 are formed on the INT8 MFMA
//      units (v_mfma_i32_16x16x64_i8). The products of one weight class are
//      summed exactly in an int32 accumulator, since (s * 127^2 * k) < 2^31
//      holds per pass over k. Each k chunk is loaded once, and all of its
//      slice pairs are multiplied from registers;
//   3. after the k loop, the class sums are scaled by 2^-7d and added in
//      FP64, smallest weight first. Then c_ij = alpha 2^(e_i+f_j) sum + beta c_ij.
//
// Truncation bound per element:
//   |c_ij - c~_ij| <= |alpha| 2^(e_i+f_j) k (s+2) 2^-7s.
// The Frobenius norm of this bound is reported with every call.
#include "internal.hpp"
#include "gemm_tiles.hpp"

#include <algorithm>
#include <climits>
#include <cmath>

namespace blis_amd {

namespace detail_split {

typedef int i4_t __attribute__((ext_vector_type(4)));

constexpr int KC = 64;          // k per MFMA step (bytes of one slice chunk)
constexpr int ROWS_ALIGN = 64;  // slice buffers are padded to whole tiles
constexpr int EXP_NONFINITE = INT_MAX;

// ---- operand statistics --------------------------------------------------
// Per row of X_eff (rows x K), where X_eff(i,p) = x[i*s_row + p*s_k]:
//   emax = largest binary exponent (INT_MIN for an all-zero row,
//          EXP_NONFINITE if any entry is Inf/NaN);
//   esum, ecnt = sum and count of the exponents of the nonzero entries.
// Exponents follow frexp: v = f 2^e with 0.5 <= |f| < 1.
__global__ void stats_init_kernel(int* emax, long long* esum, int* ecnt, int64_t rows) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows;
         i += (int64_t)gridDim.x * blockDim.x) {
        emax[i] = INT_MIN;
        esum[i] = 0;
        ecnt[i] = 0;
    }
}

struct ExpStat {
    int m = INT_MIN;
    int cnt = 0;
    long long sum = 0;
};

// Exponent of one entry from its bit pattern (subnormals through frexp).
__device__ inline void stat_one(double v, ExpStat& a) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int be = (int)((b >> 52) & 0x7FF);
    int e;
    if (be == 0x7FF) { a.m = EXP_NONFINITE; return; }
    if (be == 0) {
        if ((b << 1) == 0) return;        // +-0
        frexp(v, &e);
    } else {
        e = be - 1022;
    }
    a.m = e > a.m ? e : a.m;
    a.sum += e;
    ++a.cnt;
}

__device__ inline void stat_one(float v, ExpStat& a) {
    const unsigned b = (unsigned)__float_as_int(v);
    const int be = (int)((b >> 23) & 0xFF);
    int e;
    if (be == 0xFF) { a.m = EXP_NONFINITE; return; }
    if (be == 0) {
        if ((b << 1) == 0) return;
        frexpf(v, &e);
    } else {
        e = be - 126;
    }
    a.m = e > a.m ? e : a.m;
    a.sum += e;
    ++a.cnt;
}

__device__ inline void stat_merge(ExpStat& a, const ExpStat& b) {
    a.m = b.m > a.m ? b.m : a.m;
    a.sum += b.sum;
    a.cnt += b.cnt;
}

__device__ inline void stat_commit(const ExpStat& a, int64_t i, int* emax, long long* esum,
                                   int* ecnt) {
    if (a.m != INT_MIN) atomicMax(&emax[i], a.m);
    if (a.cnt) {
        atomicAdd((unsigned long long*)&esum[i], (unsigned long long)a.sum);
        atomicAdd(&ecnt[i], a.cnt);
    }
}

// Rows adjacent in memory (s_row == 1): lanes own 64 consecutive rows, the
// four waves interleave over the block's k range (p = p0 + wave + 4j), and
// the waves' partial statistics are merged through LDS.
template <typename T>
__global__ __launch_bounds__(256) void stats_rows_kernel(const T* __restrict__ x, int64_t rows,
                                                         int64_t K, int64_t s_k, int64_t kchunk,
                                                         int* emax, long long* esum, int* ecnt) {
    __shared__ int sm[4][64], sc[4][64];
    __shared__ long long ss[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = blockIdx.x * (int64_t)64 + lane;
    const int64_t p0 = blockIdx.y * kchunk;
    const int64_t p1 = (p0 + kchunk < K) ? p0 + kchunk : K;
    ExpStat a;
    if (i < rows) {
        const T* xr = x + i;
        int64_t p = p0 + wave;
        for (; p + 28 < p1; p += 32) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = xr[(p + 4 * u) * s_k];
#pragma unroll
            for (int u = 0; u < 8; ++u) stat_one(v[u], a);
        }
        for (; p < p1; p += 4) stat_one(xr[p * s_k], a);
    }
    sm[wave][lane] = a.m;
    sc[wave][lane] = a.cnt;
    ss[wave][lane] = a.sum;
    __syncthreads();
    if (wave == 0 && i < rows) {
        for (int w = 1; w < 4; ++w) {
            ExpStat b;
            b.m = sm[w][lane];
            b.cnt = sc[w][lane];
            b.sum = ss[w][lane];
            stat_merge(a, b);
        }
        stat_commit(a, i, emax, esum, ecnt);
    }
}

// Any other layout: one wave per row, lanes along k, merged by shuffles.
template <typename T>
__global__ __launch_bounds__(256) void stats_cols_kernel(const T* __restrict__ x, int64_t rows,
                                                         int64_t K, int64_t s_row, int64_t s_k,
                                                         int64_t kchunk, int* emax, long long* esum,
                                                         int* ecnt) {
    const int lane = threadIdx.x & 63;
    const int64_t i = blockIdx.x * (int64_t)4 + (threadIdx.x >> 6);
    if (i >= rows) return;
    const int64_t p0 = blockIdx.y * kchunk;
    const int64_t p1 = (p0 + kchunk < K) ? p0 + kchunk : K;
    const T* xr = x + i * s_row;
    ExpStat a;
    int64_t p = p0 + lane;
    for (; p + 448 < p1; p += 512) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = xr[(p + 64 * u) * s_k];
#pragma unroll
        for (int u = 0; u < 8; ++u) stat_one(v[u], a);
    }
    for (; p < p1; p += 64) stat_one(xr[p * s_k], a);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ExpStat b;
        b.m = __shfl_xor(a.m, off);
        b.cnt = __shfl_xor(a.cnt, off);
        b.sum = __shfl_xor(a.sum, off);
        stat_merge(a, b);
    }
    if (lane == 0) stat_commit(a, i, emax, esum, ecnt);
}

// ---- slicing ---------------------------------------------------------------
// Digits of X_eff in the layout the GEMM reads:
//   out[((kc * rows_pad + i) * S + t) * 64 + (p % 64)],  kc = p / 64,
// so a tile of 64 rows at one k chunk is one contiguous block of 64*S*64
// bytes. Rows >= rows and p >= K are zero. A workgroup makes one such block
// at a time: each thread slices 16 consecutive p of one row (threads along
// rows when rows are adjacent in memory, along k otherwise, for coalesced
// reads), the digits are gathered in LDS, and the block is written out
// contiguously.
constexpr int SLICE_LDS = 64 * SPLIT_MAX_SLICES * KC;

template <typename T>
__global__ __launch_bounds__(256) void slice_kernel(const T* __restrict__ x, int64_t rows,
                                                    int64_t rows_pad, int64_t K, int64_t nkc,
                                                    int64_t s_row, int64_t s_k, int row_contig,
                                                    const int* __restrict__ emax, int S,
                                                    signed char* __restrict__ out) {
    // 64 * S * KC bytes, sized at launch so that fewer slices let more
    // workgroups share a CU.
    extern __shared__ __attribute__((aligned(16))) signed char blk[];
    const int tid = threadIdx.x;
    const int64_t rblocks = rows_pad / 64, tiles = rblocks * nkc;
    const int chunks = 64 * S * KC / 16;                 // 16-byte chunks per block
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t rb = tile % rblocks, kc = tile / rblocks;
        double v[16];
        if (row_contig) {
            // Lanes along rows (unit stride): thread (r, sub) takes k
            // positions sub*16 .. sub*16 + 15 of row r and writes its 16
            // digits of each slice as one 16-byte chunk.
            const int r = tid & 63, sub = tid >> 6;
            const int64_t i = rb * 64 + r;
            if (i < rows) {
                const int em = emax[i];
                const int e = (em == INT_MIN) ? 0 : em;
                const int64_t pb = kc * KC + sub * 16;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int64_t p = pb + q;
                    v[q] = (p < K) ? ldexp((double)x[i * s_row + p * s_k], -e) : 0.0;  // exact
                }
            } else {
#pragma unroll
                for (int q = 0; q < 16; ++q) v[q] = 0.0;
            }
            for (int t = 0; t < S; ++t) {
                union { signed char b[16]; i4_t w; } dg;
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const double y = v[q] * 128.0;    // exact
                    const double tr = trunc(y);       // |tr| <= 127
                    dg.b[q] = (signed char)(int)tr;
                    v[q] = y - tr;                    // exact, |v| < 1
                }
                *reinterpret_cast<i4_t*>(blk + ((r * S + t) * KC + sub * 16)) = dg.w;
            }
        } else {
            // Lanes along k (unit stride): lane l takes k position l of rows
            // w*16 .. w*16 + 15 (w = wave), so each load instruction reads 64
            // consecutive elements of one row. Digits go to LDS one byte each.
            const int lane = tid & 63, w = tid >> 6;
            const int64_t p = kc * KC + lane;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int64_t i = rb * 64 + w * 16 + q;
                double val = 0.0;
                if (i < rows && p < K) {
                    const int em = emax[i];
                    const int e = (em == INT_MIN) ? 0 : em;
                    val = ldexp((double)x[i * s_row + p * s_k], -e);              // exact
                }
                v[q] = val;
            }
            for (int t = 0; t < S; ++t) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const double y = v[q] * 128.0;    // exact
                    const double tr = trunc(y);       // |tr| <= 127
                    blk[((w * 16 + q) * S + t) * KC + lane] = (signed char)(int)tr;
                    v[q] = y - tr;                    // exact, |v| < 1
                }
            }
        }
        __syncthreads();
        i4_t* dst = reinterpret_cast<i4_t*>(out + (kc * rows_pad + rb * 64) * S * KC);
        const i4_t* src = reinterpret_cast<const i4_t*>(blk);
        for (int c = tid; c < chunks; c += 256) dst[c] = src[c];
        __syncthreads();
    }
}

// ---- slice GEMM ------------------------------------------------------------
// D[x][y] = sum_p X[x][p] Y[y][p] per weight class. X supplies the MFMA's row
// operand and Y its column operand. A lane's accumulator entries then have y
// (= lane % 16) as their fastest index, which the host maps to C's unit-stride
// dimension.
//
// Workgroup: 64 x-rows by 64 y-rows, 4 waves as 2 x 2, 32 x 32 per wave (2 x 2
// MFMA blocks of 16 x 16). Per 64-wide k chunk a wave holds the fragments of
// all S Y-slices in registers and streams the X-slices, so each slice
// fragment read from LDS feeds up to S products.
//
// Staging: the k chunk of a 64-row tile is one contiguous block of 64*S*64
// bytes in global memory. It is copied straight into LDS with
// global_load_lds_dwordx4: lane l of an instruction fills LDS bytes
// [base + 16 l, base + 16 l + 16), so the lanes choose the source chunks. The
// LDS row (S*64 bytes, no padding) holds its 16-byte chunks in swizzled
// order (swz), so a fragment read, 16 rows at one chunk, touches 16
// different bank groups. There are no staging registers, and one buffer of
// 2*64*S*64 bytes lets two workgroups share a CU.
template <int S>
struct SliceTile {
    static constexpr int NT = 256;
    static constexpr int CPR = 4 * S;                   // 16-byte chunks per row
    static constexpr int OPB = 64 * CPR * 16;           // one operand's LDS tile
    static constexpr int LDS = 2 * OPB;
    static constexpr int ROUNDS = 64 * CPR / NT;        // load instructions per wave per operand
    static_assert(ROUNDS == S, "each wave issues S loads per operand");
    // Two workgroups per CU overlap each other's loads and multiplies. Each
    // also double-buffers its LDS tile when two stages of both workgroups
    // fit (S <= 5); above that a single stage is used (measured at S = 8:
    // one double-buffered workgroup per CU was 16% slower than two single-
    // buffered ones).
    static constexpr int NBUF = (4 * LDS <= 160 * 1024) ? 2 : 1;
    static constexpr int WAVES_PER_EU = S <= 8 ? 2 : 1;
};

// Position of logical chunk c of row r inside the LDS row (an involution).
// Chunks are permuted by XOR within aligned groups of 16 (of 4 in a
// trailing partial group).
template <int S>
__device__ __forceinline__ int swz(int r, int c) {
    constexpr int CPR = 4 * S;
    return ((c | 15) < CPR) ? ((c & ~15) | ((c & 15) ^ (r & 15)))
                            : ((c & ~3) | ((c & 3) ^ (r & 3)));
}

struct SliceGemmArgs {
    const signed char* xs;   // slices of X, rows padded to x_pad
    const signed char* ys;   // slices of Y, rows padded to y_pad
    const int* ex;           // exponents (INT_MIN = zero row)
    const int* ey;
    void* c;                 // C(x, y) at c + x*c_sx + y*c_sy
    int64_t c_sx, c_sy;
    int64_t nx, ny, x_pad, y_pad;
    int64_t kc0, kc1;        // k chunk range of this pass
    double alpha, beta;
};

typedef __attribute__((address_space(3))) void* lds_ptr_t;

constexpr int SPLIT_GROUP = 8;   // raster group height in tiles

template <typename T, int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SliceTile<S>::WAVES_PER_EU,
                                                                     SliceTile<S>::WAVES_PER_EU)))
void slice_gemm_kernel(SliceGemmArgs g) {
    using TL = SliceTile<S>;
    __shared__ __attribute__((aligned(16))) signed char smem[TL::NBUF * TL::LDS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wx = (wave & 1) * 32, wy = (wave >> 1) * 32;
    // 1-D grid over the (x_pad/64) x (y_pad/64) tiles. Workgroups are
    // dispatched round-robin over the 8 XCDs; each XCD gets a contiguous
    // range of logical tiles, rasterised in groups of SPLIT_GROUP tile rows,
    // so the slice tiles a workgroup streams are shared in its XCD's L2.
    int64_t x0, y0;
    {
        const int nb = gridDim.x, gx = (int)(g.x_pad / 64), gy = (int)(g.y_pad / 64);
        int bid = blockIdx.x;
        const int q = nb / 8, rm = nb % 8, xcd = bid % 8, l = bid / 8;
        bid = xcd * q + (xcd < rm ? xcd : rm) + l;
        const int per_group = SPLIT_GROUP * gy;
        const int group = bid / per_group;
        const int first = group * SPLIT_GROUP;
        const int gsz = (gx - first < SPLIT_GROUP) ? gx - first : SPLIT_GROUP;
        const int in = bid - group * per_group;
        x0 = (int64_t)(first + in % gsz) * 64;
        y0 = (int64_t)(in / gsz) * 64;
    }

    i4_t acc[S][2][2];
#pragma unroll
    for (int c = 0; c < S; ++c)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[c][i][j] = i4_t{0, 0, 0, 0};

    // Source chunk of every (round, lane): LDS position P = (q*4 + wave)*64 +
    // lane holds logical chunk swz(r, P % CPR) of row r = P / CPR.
    int src[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int P = (q * 4 + wave) * 64 + lane;
        const int r = P / TL::CPR, cp = P % TL::CPR;
        src[q] = (r * TL::CPR + swz<S>(r, cp)) * 16;
    }
    auto stage = [&](int64_t kc, int buf) {
        const signed char* gx = g.xs + ((kc * g.x_pad + x0) * S) * KC;
        const signed char* gy = g.ys + ((kc * g.y_pad + y0) * S) * KC;
#pragma unroll
        for (int q = 0; q < S; ++q) {
            signed char* dx = smem + buf * TL::LDS + (q * 4 + wave) * 64 * 16;
            __builtin_amdgcn_global_load_lds((const void*)(gx + src[q]), (lds_ptr_t)dx, 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(gy + src[q]), (lds_ptr_t)(dx + TL::OPB),
                                             16, 0, 0);
        }
    };

    const int lr = lane & 15, kk = lane >> 4;
    stage(g.kc0, 0);
    for (int64_t kc = g.kc0; kc < g.kc1; ++kc) {
        const int buf = (TL::NBUF == 2) ? (int)((kc - g.kc0) & 1) : 0;
        __builtin_amdgcn_s_waitcnt(0);   // this wave's loads of chunk kc
        __syncthreads();                 // everyone's loads; the other stage is free
        if (TL::NBUF == 2 && kc + 1 < g.kc1) stage(kc + 1, buf ^ 1);
        const signed char* sx = smem + buf * TL::LDS;
        const signed char* sy = sx + TL::OPB;
        i4_t yf[S][2];
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int r = wy + 16 * j + lr;
                yf[u][j] = *reinterpret_cast<const i4_t*>(sy + (r * TL::CPR + swz<S>(r, 4 * u + kk)) * 16);
            }
#pragma unroll
        for (int t = 0; t < S; ++t) {
            i4_t xf[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = wx + 16 * i + lr;
                xf[i] = *reinterpret_cast<const i4_t*>(sx + (r * TL::CPR + swz<S>(r, 4 * t + kk)) * 16);
            }
#pragma unroll
            for (int u = 0; u + t < S; ++u)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[t + u][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                            xf[i], yf[u][j], acc[t + u][i][j], 0, 0, 0);
        }
        if (TL::NBUF == 1) {
            __syncthreads();             // every wave is done with this chunk
            if (kc + 1 < g.kc1) stage(kc + 1, 0);
        }
    }

    // Epilogue: class c (weight d = c + 2) carries 2^-7(c+2).
    T* C = (T*)g.c;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t y = y0 + wy + 16 * j + lr;
            if (y >= g.ny) continue;
            const int fy = g.ey[y] == INT_MIN ? 0 : g.ey[y];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t x = x0 + wx + 16 * i + 4 * (lane >> 4) + r;
                if (x >= g.nx) continue;
                double v = 0.0;
#pragma unroll
                for (int c = S - 1; c >= 0; --c) v += ldexp((double)acc[c][i][j][r], -7 * (c + 2));
                const int fx = g.ex[x] == INT_MIN ? 0 : g.ex[x];
                const double val = g.alpha * ldexp(v, fx + fy);
                T* dst = C + x * g.c_sx + y * g.c_sy;
                *dst = (T)(g.beta == 0.0 ? val : val + g.beta * (double)*dst);
            }
        }
}

inline unsigned grid1(int64_t threads) {
    const int64_t g = (threads + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

} // namespace detail_split
using namespace detail_split;

// Slice count from the operands' exponent range. A row's entries sit on
// average `gap` bits below the row maximum (the mean of e_max - e over the
// nonzero entries), and the slices must cover those bits plus the significand
// that native arithmetic keeps. The constants were fitted with a
// double-double reference at k = 8192 so that the normwise error stays at or
// below native GEMM's error for u[-1,1] 2^round(N(0, sigma)) data,
// sigma = 0, 4, 8.
int split_choose_slices(Dt dt, double gap_bits) {
    const double need = (dt == Dt::F64 ? 50.0 : 22.0) + gap_bits;
    const int s = (int)std::ceil(need / 7.0);
    return std::max(2, s);
}

// Slice count at which the INT8 emulation stops paying: past it the slice
// products take longer than native FP64 (FP32) MFMA. The value comes from
// gfx950 timings of the two 8192^3 kernels: FP64 at 8 slices (36 products)
// ties native DGEMM (16.1 ms each), 9 slices would take about 20 ms.
int split_max_slices(Dt dt) { return dt == Dt::F64 ? 8 : 6; }

static void launch_stats(Dt dt, View v, int64_t rows, int64_t K, int64_t s_row, int64_t s_k,
                         int* emax, long long* esum, int* ecnt, hipStream_t st) {
    hipLaunchKernelGGL(stats_init_kernel, dim3(grid1(rows)), dim3(256), 0, st, emax, esum, ecnt,
                       rows);
    if (s_row == 1) {
        const int64_t kchunk = 256;
        const dim3 grid((unsigned)((rows + 63) / 64), (unsigned)((K + kchunk - 1) / kchunk));
        if (dt == Dt::F64)
            hipLaunchKernelGGL(stats_rows_kernel<double>, grid, dim3(256), 0, st,
                               (const double*)v.ptr, rows, K, s_k, kchunk, emax, esum, ecnt);
        else
            hipLaunchKernelGGL(stats_rows_kernel<float>, grid, dim3(256), 0, st,
                               (const float*)v.ptr, rows, K, s_k, kchunk, emax, esum, ecnt);
    } else {
        const int64_t kchunk = 2048;
        const dim3 grid((unsigned)((rows + 3) / 4), (unsigned)((K + kchunk - 1) / kchunk));
        if (dt == Dt::F64)
            hipLaunchKernelGGL(stats_cols_kernel<double>, grid, dim3(256), 0, st,
                               (const double*)v.ptr, rows, K, s_row, s_k, kchunk, emax, esum, ecnt);
        else
            hipLaunchKernelGGL(stats_cols_kernel<float>, grid, dim3(256), 0, st,
                               (const float*)v.ptr, rows, K, s_row, s_k, kchunk, emax, esum, ecnt);
    }
    HIP_CHECK(hipGetLastError());
}

static void launch_slice(Dt dt, View v, int64_t rows, int64_t rows_pad, int64_t K, int64_t nkc,
                         int64_t s_row, int64_t s_k, const int* emax, int S, signed char* out,
                         hipStream_t st) {
    const int64_t tiles = rows_pad / 64 * nkc;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 8192));
    const int row_contig = (s_row == 1) ? 1 : 0;
    const size_t lds = (size_t)64 * S * KC;
    if (S < 1 || lds > (size_t)SLICE_LDS) abort_with(BLIS_INVALID_UKR_ID, "split: slice count");
    if (dt == Dt::F64)
        hipLaunchKernelGGL(slice_kernel<double>, dim3(grid), dim3(256), lds, st,
                           (const double*)v.ptr, rows, rows_pad, K, nkc, s_row, s_k, row_contig,
                           emax, S, out);
    else
        hipLaunchKernelGGL(slice_kernel<float>, dim3(grid), dim3(256), lds, st,
                           (const float*)v.ptr, rows, rows_pad, K, nkc, s_row, s_k, row_contig,
                           emax, S, out);
    HIP_CHECK(hipGetLastError());
}

template <typename T>
static void launch_slice_gemm(int S, const SliceGemmArgs& a, hipStream_t st) {
    const dim3 grid((unsigned)((a.x_pad / 64) * (a.y_pad / 64)));
    switch (S) {
#define BLIS_AMD_SLICE_CASE(n) \
    case n: hipLaunchKernelGGL((slice_gemm_kernel<T, n>), grid, dim3(256), 0, st, a); break;
        BLIS_AMD_SLICE_CASE(2) BLIS_AMD_SLICE_CASE(3) BLIS_AMD_SLICE_CASE(4)
        BLIS_AMD_SLICE_CASE(5) BLIS_AMD_SLICE_CASE(6) BLIS_AMD_SLICE_CASE(7)
        BLIS_AMD_SLICE_CASE(8) BLIS_AMD_SLICE_CASE(9) BLIS_AMD_SLICE_CASE(10)
        BLIS_AMD_SLICE_CASE(11) BLIS_AMD_SLICE_CASE(12)
#undef BLIS_AMD_SLICE_CASE
    default: abort_with(BLIS_NOT_YET_IMPLEMENTED, "split gemm: slice count out of range");
    }
    HIP_CHECK(hipGetLastError());
}

// Mean over nonzero rows of (e_max - mean exponent).
static double mean_gap(const int* emax, const long long* esum, const int* ecnt, int64_t rows) {
    double tot = 0.0;
    int64_t used = 0;
    for (int64_t i = 0; i < rows; ++i) {
        if (ecnt[i] == 0) continue;
        tot += (double)emax[i] - (double)esum[i] / (double)ecnt[i];
        ++used;
    }
    return used ? tot / (double)used : 0.0;
}

// Pinned host buffer for the statistics read-back (grow-only; every user
// waits for its copies before the next call can replace it).
static void* stats_host(size_t bytes) {
    static void* p = nullptr;
    static size_t cap = 0;
    if (cap < bytes) {
        if (p) HIP_CHECK(hipHostFree(p));
        p = nullptr;
        HIP_CHECK(hipHostMalloc(&p, bytes));
        cap = bytes;
    }
    return p;
}

long double sum_pow4(const int* e, int64_t cnt) {
    constexpr int LO = -1200, HI = 1200;
    std::vector<int64_t> h(HI - LO + 1, 0);
    long double s = 0.0L;
    for (int64_t i = 0; i < cnt; ++i) {
        const int v = e[i];
        if (v == INT_MIN) continue;
        if (v >= LO && v <= HI) ++h[v - LO];
        else s += ldexpl(1.0L, 2 * v);
    }
    for (int v = LO; v <= HI; ++v)
        if (h[v - LO]) s += (long double)h[v - LO] * ldexpl(1.0L, 2 * v);
    return s;
}

SplitResult gemm_split_run(const GemmDesc& d, int nslices, hipStream_t st) {
    SplitResult res{0, 0.0};
    if (d.m <= 0 || d.n <= 0) return res;
    if (d.k <= 0 || d.alpha == 0.0) {
        scalm_run(d.dt, d.m, d.n, d.beta, d.c, st);
        return res;
    }
    if (nslices > SPLIT_MAX_SLICES)
        abort_with(BLIS_INVALID_DIM_STRIDE_COMBINATION, "split gemm: more than 12 slices");
    if (nslices < 0 && (nslices < -CRT_MAX_MODULI || nslices > -2))
        abort_with(BLIS_INVALID_DIM_STRIDE_COMBINATION, "split gemm: 2 to 32 moduli");
    if (nslices < 0 && d.k > ((int64_t)1 << 17))
        abort_with(BLIS_INVALID_DIM_STRIDE_COMBINATION, "split gemm: modular scheme needs k <= 2^17");

    // 1. Exponent statistics of op(A)'s rows and op(B)'s columns.
    const int64_t rows = d.m + d.n;
    int* emax = (int*)workspace(sizeof(int) * (size_t)rows * 2 + 256, 4);
    int* ecnt = emax + rows;
    long long* esum = (long long*)workspace(sizeof(long long) * (size_t)rows + 256, 5);
    launch_stats(d.dt, d.a, d.m, d.k, d.a.rs, d.a.cs, emax, esum, ecnt, st);
    launch_stats(d.dt, d.b, d.n, d.k, d.b.cs, d.b.rs, emax + d.m, esum + d.m, ecnt + d.m, st);
    // Read back into pinned memory: the maxima and counts (adjacent on the
    // device) in one copy, the exponent sums in another, then one wait.
    char* hb = (char*)stats_host((size_t)rows * (2 * sizeof(int) + sizeof(long long)));
    int* h_emax = (int*)hb;
    int* h_cnt = h_emax + rows;
    long long* h_sum = (long long*)(hb + (size_t)rows * 2 * sizeof(int));
    HIP_CHECK(hipMemcpyAsync(h_emax, emax, sizeof(int) * 2 * (size_t)rows, hipMemcpyDeviceToHost,
                             st));
    HIP_CHECK(hipMemcpyAsync(h_sum, esum, sizeof(long long) * (size_t)rows, hipMemcpyDeviceToHost,
                             st));
    HIP_CHECK(hipStreamSynchronize(st));

    // 2. Slice count; non-finite operands and slice counts past the break-even
    //    point take the native path (which propagates Inf/NaN as BLIS does).
    bool finite = true;
    for (int64_t i = 0; i < rows; ++i) finite &= (h_emax[i] != EXP_NONFINITE);
    //    The automatic choice takes the modular scheme when its moduli are
    //    fewer than the slice scheme's products (crt.hip); a negative count
    //    forces that scheme with -nslices moduli.
    // One slice runs as two, the fewest the slice kernels take.
    int S = nslices == 1 ? 2 : nslices, crt_n = 0, crt_sh = 0;
    if (nslices < 0) {
        S = 0;
        crt_n = -nslices;
        crt_sh = crt_shift_for(crt_n, d.k);
    } else if (S == 0) {
        const double gap = std::max(mean_gap(h_emax, h_sum, h_cnt, d.m),
                                    mean_gap(h_emax + d.m, h_sum + d.m, h_cnt + d.m, d.n));
        S = split_choose_slices(d.dt, gap);
        if (S > split_max_slices(d.dt)) S = 0;
        int sh = 0;
        const int n = crt_choose_moduli(d.dt, gap, d.k, &sh);
        if (n > 0 && (S == 0 || n < S * (S + 1) / 2)) {
            crt_n = n;
            crt_sh = sh;
            S = 0;
        }
    }
    if (finite && crt_n > 0) return crt_run(d, emax, h_emax, crt_n, crt_sh, st);
    if (!finite || S <= 0) {
        gemm_run(d, nullptr, st);
        return res;
    }

    // 3. A-priori bound ||C - C~||_F <= |alpha| k (S+2) 2^-7S
    //    * sqrt(sum_i 4^e_i) * sqrt(sum_j 4^f_j).
    res.err_bound = (double)(std::fabs(d.alpha) * (long double)d.k * (S + 2) *
                             ldexpl(1.0L, -7 * S) * sqrtl(sum_pow4(h_emax, d.m)) *
                             sqrtl(sum_pow4(h_emax + d.m, d.n)));
    res.slices = S;

    // 4. Slices. Y is the operand whose index runs along C's unit stride.
    const int64_t nkc = (d.k + KC - 1) / KC;
    const int64_t m_pad = (d.m + ROWS_ALIGN - 1) / ROWS_ALIGN * ROWS_ALIGN;
    const int64_t n_pad = (d.n + ROWS_ALIGN - 1) / ROWS_ALIGN * ROWS_ALIGN;
    signed char* as = (signed char*)workspace((size_t)nkc * m_pad * S * KC, 2);
    signed char* bs = (signed char*)workspace((size_t)nkc * n_pad * S * KC, 3);
    launch_slice(d.dt, d.a, d.m, m_pad, d.k, nkc, d.a.rs, d.a.cs, emax, S, as, st);
    launch_slice(d.dt, d.b, d.n, n_pad, d.k, nkc, d.b.cs, d.b.rs, emax + d.m, S, bs, st);

    SliceGemmArgs g;
    const bool c_col = (d.c.rs == 1);   // column-stored C: rows of C run along lanes
    if (c_col) {
        g.xs = bs; g.ys = as; g.ex = emax + d.m; g.ey = emax;
        g.nx = d.n; g.ny = d.m; g.x_pad = n_pad; g.y_pad = m_pad;
        g.c_sx = d.c.cs; g.c_sy = d.c.rs;
    } else {
        g.xs = as; g.ys = bs; g.ex = emax; g.ey = emax + d.m;
        g.nx = d.m; g.ny = d.n; g.x_pad = m_pad; g.y_pad = n_pad;
        g.c_sx = d.c.rs; g.c_sy = d.c.cs;
    }
    g.c = d.c.ptr;
    g.alpha = d.alpha;

    // 5. Passes over k short enough that every class sum is exact in int32:
    //    S * 127^2 * 64 * chunks < 2^31.
    const int64_t max_chunks =
        std::max<int64_t>(1, (((int64_t)1 << 31) - 1) / ((int64_t)S * 127 * 127 * KC));
    for (int64_t k0 = 0; k0 < nkc; k0 += max_chunks) {
        g.kc0 = k0;
        g.kc1 = std::min(nkc, k0 + max_chunks);
        g.beta = (k0 == 0) ? d.beta : 1.0;
        if (d.dt == Dt::F64) launch_slice_gemm<double>(S, g, st);
        else launch_slice_gemm<float>(S, g, st);
    }
    return res;
}

} // namespace blis_amd

using namespace blis_amd;

template <typename T>
static int split_typed(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                       const T* alpha, const T* a, inc_t rs_a, inc_t cs_a,
                       const T* b, inc_t rs_b, inc_t cs_b, const T* beta,
                       T* c, inc_t rs_c, inc_t cs_c, int nslices, double* err_bound) {
    init_once();
    if (err_bound) *err_bound = 0.0;
    if (m <= 0 || n <= 0) return 0;
    const bool ta = ((unsigned)transa & BLIS_TRANS_BIT) != 0;
    const bool tb = ((unsigned)transb & BLIS_TRANS_BIT) != 0;
    const Dt dt = dt_of<T>();
    std::lock_guard<std::mutex> lock(api_mutex());
    Staged sa, sb, sc;
    if (k > 0) {
        stage(sa, a, dt, ta ? k : m, ta ? m : k, rs_a, cs_a, false);
        stage(sb, b, dt, tb ? n : k, tb ? k : n, rs_b, cs_b, false);
    }
    stage(sc, c, dt, m, n, rs_c, cs_c, true);
    GemmDesc d;
    d.dt = dt; d.m = m; d.n = n; d.k = k;
    d.a.ptr = sa.dev; d.a.rs = ta ? cs_a : rs_a; d.a.cs = ta ? rs_a : cs_a;
    d.b.ptr = sb.dev; d.b.rs = tb ? cs_b : rs_b; d.b.cs = tb ? rs_b : cs_b;
    d.c.ptr = sc.dev; d.c.rs = rs_c; d.c.cs = cs_c;
    d.alpha = (double)*alpha;
    d.beta = (double)*beta;
    const SplitResult r = gemm_split_run(d, nslices, stream());
    sync_if_needed();
    if (err_bound) *err_bound = r.err_bound;
    return r.slices;
}

extern "C" {

__attribute__((visibility("default")))
int bli_amd_dgemm_split(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                        const double* alpha, const double* a, inc_t rs_a, inc_t cs_a,
                        const double* b, inc_t rs_b, inc_t cs_b, const double* beta,
                        double* c, inc_t rs_c, inc_t cs_c, int nslices, double* err_bound) {
    return split_typed<double>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                               beta, c, rs_c, cs_c, nslices, err_bound);
}

__attribute__((visibility("default")))
int bli_amd_sgemm_split(trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                        const float* alpha, const float* a, inc_t rs_a, inc_t cs_a,
                        const float* b, inc_t rs_b, inc_t cs_b, const float* beta,
                        float* c, inc_t rs_c, inc_t cs_c, int nslices, double* err_bound) {
    return split_typed<float>(transa, transb, m, n, k, alpha, a, rs_a, cs_a, b, rs_b, cs_b,
                              beta, c, rs_c, cs_c, nslices, err_bound);
}

__attribute__((visibility("default")))
int bli_amd_gemm_split_slices(int is_double, double gap_bits) {
    return split_choose_slices(is_double ? Dt::F64 : Dt::F32, gap_bits);
}

__attribute__((visibility("default")))
int bli_amd_gemm_split_max_slices(int is_double) {
    return split_max_slices(is_double ? Dt::F64 : Dt::F32);
}

__attribute__((visibility("default")))
int bli_amd_gemm_split_moduli(int is_double, double gap_bits, dim_t k, int* sh) {
    int s = 0;
    const int n = crt_choose_moduli(is_double ? Dt::F64 : Dt::F32, gap_bits, k, &s);
    if (sh) *sh = s;
    return n;
}

} // extern "C"
