// crt.hip — the modular scheme of split-precision GEMM: the product of the
// operands' integer images is formed exactly from its residues modulo a set
// of pairwise coprime moduli, one INT8 GEMM per modulus, and rebuilt by the
// Chinese remainder theorem in FP64.
//
// The reference has no such path; its GEMM is the FP64 (FP32) microkernel
// (ref_kernels/3/bli_gemm_ref.c:163-317) whose accuracy this path must
// match. The method is Ozaki scheme II: Ozaki, Uchino, Imamura, "Ozaki Scheme
// II: A GEMM-oriented emulation of floating-point matrix multiplication using
// an integer modular technique" (2024), on the INT8 MFMA units:
//
//   1. row i of op(A) is scaled to integers with sh fractional bits below
//      its largest exponent e_i: a'_ip = trunc(a_ip 2^(sh - e_i)), so
//      |a'_ip| < 2^sh and |a_ip - a'_ip 2^(e_i - sh)| < 2^(e_i - sh); column
//      j of op(B) alike (exponent f_j). The products sum_p a'_ip b'_pj are
//      integers of magnitude below k 2^(2 sh) < M / 4, M = prod_l p_l;
//   2. for every modulus p_l the residues of a' and b' (int8: symmetric for
//      odd p_l, two's complement for 256) are multiplied on the INT8 MFMA
//      units (v_mfma_i32_16x16x64_i8). The int32 sums are exact (k * 128^2 <
//      2^31) and reduced mod p_l into one byte per element;
//   3. X = sum_l c_l w_l - q M with w_l = (M/p_l) ((M/p_l)^-1 mod p_l): the
//      weights and M are split into 39-bit parts so that every partial sum
//      sum_l c_l w_lk is exact in FP64, q = round(sum / M), every part of X
//      is exact, and the two leading parts (which may cancel) are added
//      error-free. Then c_ij = alpha 2^(e_i + f_j - 2 sh) X + beta c_ij.
//
// The integer product is exact, so the only error is the truncation in
// step 1: |c_ij - c~_ij| <= |alpha| k 2^(e_i + f_j) 2^(1 - sh), plus the
// final roundings. One INT8 product per modulus: about 16 moduli for FP64
// data of one binade, against 36 slice products for the slice scheme at the
// same accuracy (split.hip).
#include "internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

namespace blis_amd {

namespace detail_crt {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void* lds_ptr_t;

constexpr int KC = 64;          // k per MFMA step (bytes of one residue chunk)
constexpr int TILE = 256;       // rows of X and of Y per workgroup
constexpr int NT = 512;         // eight waves
constexpr int ROWS_ALIGN = 256; // residue planes are padded to whole tiles
constexpr int PARTS_MAX = 7;    // 39-bit parts of a weight (up to 273 bits)
constexpr int POW2_MAX = 128;   // 2^e mod p for e < 128 (sh < 181)
constexpr int PART_BITS = 39;

// Pairwise coprime moduli, largest first: 2^8, and odd numbers whose prime
// factors appear once in the list (255 = 3*5*17, 253 = 11*23, 247 = 13*19,
// the rest prime). Residues fit int8: [-128, 127] for 256, symmetric
// [-(p-1)/2, (p-1)/2] for odd p.
constexpr int MODULI[CRT_MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223,
                                        211, 199, 197, 193, 191, 181, 179, 173, 167, 163, 157,
                                        151, 149, 139, 137, 131, 127, 113, 109, 107, 103};

struct CrtConst {
    int n;                              // moduli
    int parts;                          // parts of each weight
    int p[CRT_MAX_MODULI];
    double inv_p[CRT_MAX_MODULI];       // 1/p in FP64
    double w[CRT_MAX_MODULI][PARTS_MAX];// w_l in parts, most significant first
    double mpart[PARTS_MAX];            // M in the same parts
    double inv_m;                       // 1/M rounded
    unsigned char pow2[CRT_MAX_MODULI][POW2_MAX];   // 2^e mod p_l
};

// ---- residues ----------------------------------------------------------------

// r = v mod p in [0, p) for an integer-valued |v| < 2^53 (fma makes v - q p
// exact; q may be off by one, which the two corrections absorb).
__device__ inline double fmod_p(double v, double p, double ip) {
    const double q = floor(v * ip);
    double r = fma(-q, p, v);
    if (r < 0.0) r += p;
    if (r >= p) r -= p;
    return r;
}

// The symmetric residue of an integer-valued double |v| < 2^53 modulo p
// (in [-(p-1)/2, (p-1)/2], or [-128, 128] for 256, whose byte is the same
// for +-128): v - rint(v/p) p is exact by fma and lies within (0.52 p) of 0
// because v/p is rounded by at most 2^-52 relative; a second step with
// |r/p| < 0.52 is exact. FP64 FMA runs at the int32 rate on this GPU.
__device__ inline double sym_mod(double v, double p, double ip) {
    double r = fma(-rint(v * ip), p, v);
    return fma(-rint(r * ip), p, r);
}

// An integer-valued double t as t = vm 2^ex with |vm| < 2^53 (ex = 0 when
// |t| < 2^53, ex < 128 since |t| < 2^sh, sh < 181).
__device__ inline double int_split(double t, int& ex) {
    int E;
    (void)frexp(t, &E);   // |t| = f 2^E, 0.5 <= f < 1
    ex = (E > 53) ? E - 53 : 0;
    return ex ? ldexp(t, -ex) : t;   // exact: t has at most 53 significant bits
}

// Residue planes of X_eff (rows x K), X_eff(i, p) = x[i*s_row + p*s_k], in
// the layout the GEMM reads: plane l holds
//   out[l*plane + (kc*rows_pad + i)*64 + (p % 64)],  kc = p / 64,
// so the k chunk of a 256-row tile is one contiguous 16 KiB block. Rows >=
// rows and p >= K are zero. A workgroup handles 64 rows x 64 k at a time:
// the integers trunc(x 2^(sh - e_i)) are staged in LDS (read with lanes
// along whichever index is adjacent in memory), then thread (r, sub) takes
// k positions sub*16 .. sub*16 + 15 of row r and writes them as one 16-byte
// chunk per plane.
template <typename T>
__global__ __launch_bounds__(256) void crt_residue_kernel(const T* __restrict__ x, int64_t rows,
                                                          int64_t rows_pad, int64_t K,
                                                          int64_t nkc, int64_t s_row,
                                                          int64_t s_k, int row_contig,
                                                          const int* __restrict__ emax, int sh,
                                                          const CrtConst* __restrict__ cc,
                                                          signed char* __restrict__ out,
                                                          int64_t plane) {
    __shared__ double tt[64][65];
    __shared__ unsigned char pw[CRT_MAX_MODULI][POW2_MAX];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = cc->n;
    for (int i = tid; i < CRT_MAX_MODULI * POW2_MAX / 4; i += blockDim.x)
        reinterpret_cast<uint32_t*>(&pw[0][0])[i] = reinterpret_cast<const uint32_t*>(&cc->pow2[0][0])[i];
    const int64_t rblocks = rows_pad / 64, tiles = rblocks * nkc;
    // Output mapping: a wave writes 16 whole rows (1 KiB) per plane.
    const int r = w * 16 + (lane >> 2), sub = lane & 3;
    for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int64_t rb = tile % rblocks, kc = tile / rblocks;
        if (row_contig) {
            const int64_t i = rb * 64 + lane;
            const int em = (i < rows) ? emax[i] : INT_MIN;
            const int e = (em == INT_MIN) ? 0 : em;
#pragma unroll 4
            for (int q = 0; q < 16; ++q) {
                const int64_t pk = kc * KC + w * 16 + q;
                tt[lane][w * 16 + q] =
                    (i < rows && pk < K) ? trunc(ldexp((double)x[i * s_row + pk * s_k], sh - e)) : 0.0;
            }
        } else {
            const int64_t pk = kc * KC + lane;
#pragma unroll 4
            for (int q = 0; q < 16; ++q) {
                const int64_t i = rb * 64 + w * 16 + q;
                double t = 0.0;
                if (i < rows && pk < K) {
                    const int em = emax[i];
                    const int e = (em == INT_MIN) ? 0 : em;
                    t = trunc(ldexp((double)x[i * s_row + pk * s_k], sh - e));
                }
                tt[w * 16 + q][lane] = t;
            }
        }
        __syncthreads();
        double vm[16];
        int ex[16];
        int any = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            vm[q] = int_split(tt[r][sub * 16 + q], ex[q]);
            any |= ex[q];
        }
        __syncthreads();   // tt is free for the next tile
        const bool big = __any(any != 0);   // some integer of the wave is >= 2^53
        signed char* dst = out + (kc * rows_pad + rb * 64 + r) * KC + sub * 16;
        for (int l = 0; l < n; ++l) {
            const double p = (double)cc->p[l], ip = cc->inv_p[l];
            int res[16];
            if (!big) {
#pragma unroll
                for (int q = 0; q < 16; ++q) res[q] = (int)sym_mod(vm[q], p, ip);
            } else {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    // One step leaves |r0| < 0.52 p; times 2^ex mod p that
                    // stays below 2^16, so one more step is exact and lands
                    // in the symmetric range.
                    const double r0 = fma(-rint(vm[q] * ip), p, vm[q]);
                    const double x = r0 * (double)pw[l][ex[q]];
                    res[q] = (int)fma(-rint(x * ip), p, x);
                }
            }
            uint32_t wd[4];
#pragma unroll
            for (int b4 = 0; b4 < 4; ++b4)
                wd[b4] = (uint32_t)(res[4 * b4] & 0xff) | ((uint32_t)(res[4 * b4 + 1] & 0xff) << 8) |
                         ((uint32_t)(res[4 * b4 + 2] & 0xff) << 16) |
                         ((uint32_t)(res[4 * b4 + 3] & 0xff) << 24);
            *reinterpret_cast<i4_t*>(dst + l * plane) =
                i4_t{(int)wd[0], (int)wd[1], (int)wd[2], (int)wd[3]};
        }
    }
}

// ---- residue GEMM ----------------------------------------------------------------
// For modulus l = blockIdx.y: D[x][y] = sum_p X_l[x][p] Y_l[y][p] mod p_l,
// one byte per element, out[l*out_plane + x*ld_out + y]. X supplies the
// MFMA's row operand and Y its column operand, so a lane's accumulator
// entries have y (= lane % 16) as their fastest index, which the host maps
// to C's unit-stride dimension.
//
// Workgroup: 256 x-rows by 256 y-rows, eight waves as 2 (x) x 4 (y), 128 x 64
// per wave (8 x 4 MFMA blocks of 16 x 16): per 64-deep k chunk a wave reads
// 12 fragments (8 of X, 4 of Y) from LDS for 32 MFMAs. The chunk of each
// operand (256 rows x 64 bytes, contiguous in the plane) is copied straight
// into LDS with global_load_lds_dwordx4, double buffered (64 KiB). The
// 16-byte chunks of a 64-byte LDS row are permuted, chunk c of row r at
// c ^ ((r >> 1) & 3), so the 16 rows of a fragment read fall in distinct
// bank groups.
struct CrtGemmArgs {
    const signed char* xs;   // residue planes of X, rows padded to x_pad
    const signed char* ys;   // residue planes of Y, rows padded to y_pad
    int64_t x_plane, y_plane;
    int64_t x_pad, y_pad, nkc;
    int64_t nx, ny;
    unsigned char* out;      // residue products
    int64_t out_plane, ld_out;
    const CrtConst* cc;
};

__device__ __forceinline__ int swz(int r, int c) { return c ^ ((r >> 1) & 3); }

constexpr int OPB = TILE * KC;         // one operand's LDS tile (16 KiB)
constexpr int STAGE = 2 * OPB;
constexpr int GROUP = 4;               // raster group height in tiles

__global__ __launch_bounds__(NT) void crt_gemm_kernel(CrtGemmArgs g) {
    __shared__ __attribute__((aligned(16))) signed char smem[2 * STAGE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wx = (wave & 1) * 128, wy = (wave >> 1) * 64;
    const int l = blockIdx.y;
    const int p = g.cc->p[l];
    const double pd = (double)p, ip = g.cc->inv_p[l];
    // 1-D grid over the tiles: each XCD gets a contiguous range of logical
    // tiles, rasterised in groups of GROUP tile rows, so the tiles a
    // workgroup streams are shared in its XCD's L2.
    int64_t x0, y0;
    {
        const int nb = gridDim.x, gx = (int)(g.x_pad / TILE), gy = (int)(g.y_pad / TILE);
        int bid = blockIdx.x;
        const int q = nb / 8, rm = nb % 8, xcd = bid % 8, li = bid / 8;
        bid = xcd * q + (xcd < rm ? xcd : rm) + li;
        const int per_group = GROUP * gy;
        const int group = bid / per_group;
        const int first = group * GROUP;
        const int gsz = (gx - first < GROUP) ? gx - first : GROUP;
        const int in = bid - group * per_group;
        x0 = (int64_t)(first + in % gsz) * TILE;
        y0 = (int64_t)(in / gsz) * TILE;
    }
    const signed char* xs = g.xs + l * g.x_plane;
    const signed char* ys = g.ys + l * g.y_plane;

    i4_t acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = i4_t{0, 0, 0, 0};

    // Two load rounds per operand: LDS chunk P = (q*8 + wave)*64 + lane holds
    // logical chunk swz(r, P % 4) of row r = P / 4 (swz is an involution).
    int src[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int P = (q * 8 + wave) * 64 + lane;
        const int r = P >> 2;
        src[q] = (r * 4 + swz(r, P & 3)) * 16;
    }
    auto stage = [&](int64_t kc, int buf) {
        const signed char* gx = xs + (kc * g.x_pad + x0) * KC;
        const signed char* gy = ys + (kc * g.y_pad + y0) * KC;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            signed char* dx = smem + buf * STAGE + (q * 8 + wave) * 64 * 16;
            __builtin_amdgcn_global_load_lds((const void*)(gx + src[q]), (lds_ptr_t)dx, 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(gy + src[q]), (lds_ptr_t)(dx + OPB), 16,
                                             0, 0);
        }
    };

    const int lr = lane & 15, kk = lane >> 4;
    stage(0, 0);
    for (int64_t kc = 0; kc < g.nkc; ++kc) {
        const int buf = (int)(kc & 1);
        __builtin_amdgcn_s_waitcnt(0);   // this wave's loads of chunk kc
        __syncthreads();                 // everyone's loads; the other stage is free
        if (kc + 1 < g.nkc) stage(kc + 1, buf ^ 1);
        const signed char* sx = smem + buf * STAGE;
        const signed char* sy = sx + OPB;
        i4_t yf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = wy + 16 * j + lr;
            yf[j] = *reinterpret_cast<const i4_t*>(sy + (r * 4 + swz(r, kk)) * 16);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = wx + 16 * i + lr;
            const i4_t xf = *reinterpret_cast<const i4_t*>(sx + (r * 4 + swz(r, kk)) * 16);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(xf, yf[j], acc[i][j], 0, 0, 0);
        }
    }

    // Epilogue: each exact int32 sum reduced mod p_l into one byte.
    unsigned char* o = g.out + l * g.out_plane;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t y = y0 + wy + 16 * j + lr;
        if (y >= g.ny) continue;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t x = x0 + wx + 16 * i + 4 * kk + r;
                if (x >= g.nx) continue;
                o[x * g.ld_out + y] = (unsigned char)(int)fmod_p((double)acc[i][j][r], pd, ip);
            }
    }
}

// ---- reconstruction ----------------------------------------------------------
// C(x, y) = alpha 2^(ex_x + ey_y - 2 sh) X + beta C(x, y), X rebuilt from the
// residue products (step 3 of the file comment). beta == 0 does not read C.
// A thread rebuilds four adjacent y of one x (one 4-byte load per plane when
// vec4 is set: ld and plane multiples of 4); blockIdx.y walks x. The moduli's
// constants are wave-uniform and come through scalar loads. PARTS >= 2: the
// weights of a one-part modulus set carry a zero second part.
template <typename T, int PARTS>
__global__ __launch_bounds__(256) void crt_combine_kernel(const unsigned char* __restrict__ prod,
                                                          int64_t plane, int64_t ld,
                                                          int64_t nx, int64_t ny, int vec4,
                                                          const CrtConst* __restrict__ cc,
                                                          const int* __restrict__ ex,
                                                          const int* __restrict__ ey, int sh,
                                                          T* C, int64_t c_sx, int64_t c_sy,
                                                          double alpha, double beta) {
    const int n = cc->n;
    const int64_t y0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (y0 >= ny) return;
    const int ny_here = (int)((ny - y0 < 4) ? ny - y0 : 4);
    for (int64_t x = blockIdx.y; x < nx; x += gridDim.y) {
        double part[4][PARTS];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < PARTS; ++k) part[j][k] = 0.0;
        const unsigned char* src = prod + x * ld + y0;
        for (int l = 0; l < n; ++l) {
            uint32_t v;
            if (vec4 && ny_here == 4) {
                v = *reinterpret_cast<const uint32_t*>(src + l * plane);
            } else {
                v = 0;
                for (int j = 0; j < ny_here; ++j) v |= (uint32_t)src[l * plane + j] << (8 * j);
            }
            double wl[PARTS];
#pragma unroll
            for (int k = 0; k < PARTS; ++k) wl[k] = cc->w[l][k];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double cl = (double)((v >> (8 * j)) & 0xffu);
#pragma unroll
                for (int k = 0; k < PARTS; ++k) part[j][k] = fma(cl, wl[k], part[j][k]);   // exact
            }
        }
        const int fx = ex[x] == INT_MIN ? 0 : ex[x];
        const double inv_m = cc->inv_m;
        double mp[PARTS];
#pragma unroll
        for (int k = 0; k < PARTS; ++k) mp[k] = cc->mpart[k];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= ny_here) break;
            const double q = rint((part[j][0] + part[j][1]) * inv_m);
            // D_k = part_k - q mpart_k, each exact. D_0 and D_1 overlap and
            // may cancel, so they are added error-free (TwoSum); the smaller
            // parts are added to the error term, smallest first.
            double D[PARTS];
#pragma unroll
            for (int k = 0; k < PARTS; ++k) D[k] = fma(-q, mp[k], part[j][k]);
            const double hi = D[0] + D[1];
            const double bb = hi - D[0];
            double lo = (D[0] - (hi - bb)) + (D[1] - bb);
            double tail = 0.0;
#pragma unroll
            for (int k = PARTS - 1; k >= 2; --k) tail += D[k];
            lo += tail;
            const double X = hi + lo;
            const int64_t y = y0 + j;
            const int fy = ey[y] == INT_MIN ? 0 : ey[y];
            const double val = alpha * ldexp(X, fx + fy - 2 * sh);
            T* dst = C + x * c_sx + y * c_sy;
            *dst = (T)(beta == 0.0 ? val : val + beta * (double)*dst);
        }
    }
}

template <typename T>
void launch_combine(const unsigned char* prod, int64_t plane, int64_t ld, int64_t nx,
                    int64_t ny, int parts, const CrtConst* cc, const int* ex, const int* ey,
                    int sh, T* C, int64_t c_sx, int64_t c_sy, double alpha, double beta,
                    hipStream_t st) {
    const int vec4 = (ld % 4 == 0 && plane % 4 == 0) ? 1 : 0;
    const dim3 grid((unsigned)((ny + 1023) / 1024), (unsigned)std::min<int64_t>(nx, 65535));
    switch (std::max(parts, 2)) {
#define BLIS_AMD_CRT_COMBINE(P)                                                                 \
    case P:                                                                                     \
        hipLaunchKernelGGL((crt_combine_kernel<T, P>), grid, dim3(256), 0, st, prod, plane, ld, \
                           nx, ny, vec4, cc, ex, ey, sh, C, c_sx, c_sy, alpha, beta);           \
        break;
        BLIS_AMD_CRT_COMBINE(2)
        BLIS_AMD_CRT_COMBINE(3)
        BLIS_AMD_CRT_COMBINE(4)
        BLIS_AMD_CRT_COMBINE(5)
        BLIS_AMD_CRT_COMBINE(6)
        BLIS_AMD_CRT_COMBINE(7)
#undef BLIS_AMD_CRT_COMBINE
    default:
        abort_with(BLIS_NOT_YET_IMPLEMENTED, "crt: weight parts out of range");
    }
    HIP_CHECK(hipGetLastError());
}

// ---- host: moduli, weights and their parts --------------------------------------

// Unsigned big integer, 32-bit limbs, least significant first.
struct Big {
    std::vector<uint32_t> d{1};
    void mul(uint32_t v) {
        uint64_t carry = 0;
        for (uint32_t& x : d) {
            const uint64_t t = (uint64_t)x * v + carry;
            x = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) d.push_back((uint32_t)carry);
    }
    uint32_t divmod(uint32_t v) {   // *this /= v, returns the remainder
        uint64_t rem = 0;
        for (size_t i = d.size(); i-- > 0;) {
            const uint64_t cur = (rem << 32) | d[i];
            d[i] = (uint32_t)(cur / v);
            rem = cur % v;
        }
        while (d.size() > 1 && d.back() == 0) d.pop_back();
        return (uint32_t)rem;
    }
    int bits() const {
        for (size_t i = d.size(); i-- > 0;)
            if (d[i]) return (int)(i * 32) + (32 - __builtin_clz(d[i]));
        return 0;
    }
    // Bits [lo, lo + width) as an integer (width <= 64).
    uint64_t field(int lo, int width) const {
        uint64_t v = 0;
        for (int b = width - 1; b >= 0; --b) {
            const int bit = lo + b;
            const size_t w = (size_t)bit / 32;
            const uint64_t x = (w < d.size()) ? (d[w] >> (bit % 32)) & 1u : 0u;
            v = (v << 1) | x;
        }
        return v;
    }
};

int inverse_mod(int a, int m) {   // a^-1 mod m (gcd(a, m) = 1)
    int t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        const int q = r / nr;
        int tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return t < 0 ? t + m : t;
}

double log2_modulus(int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::log2((double)MODULI[i]);
    return s;
}

void build_const(int n, CrtConst& c) {
    std::memset(&c, 0, sizeof(c));
    c.n = n;
    Big M;
    for (int l = 0; l < n; ++l) M.mul((uint32_t)MODULI[l]);
    const int L = M.bits();
    c.parts = (L + PART_BITS - 1) / PART_BITS;
    if (c.parts > PARTS_MAX) abort_with(BLIS_NOT_YET_IMPLEMENTED, "crt: modulus too large");
    // Part k covers bits [t_k, t_k + width_k), most significant first.
    int lo[PARTS_MAX], width[PARTS_MAX];
    for (int k = 0; k < c.parts; ++k) {
        const int top = L - PART_BITS * k;
        lo[k] = std::max(0, top - PART_BITS);
        width[k] = top - lo[k];
    }
    auto split = [&](const Big& v, double* out) {
        for (int k = 0; k < c.parts; ++k) out[k] = std::ldexp((double)v.field(lo[k], width[k]), lo[k]);
    };
    split(M, c.mpart);
    const int top_lo = std::max(0, L - 64);
    c.inv_m = std::ldexp(1.0 / (double)M.field(top_lo, L - top_lo), -top_lo);
    for (int l = 0; l < n; ++l) {
        const int p = MODULI[l];
        c.p[l] = p;
        c.inv_p[l] = 1.0 / (double)p;
        Big Ml = M;
        (void)Ml.divmod((uint32_t)p);            // M / p_l
        Big tmp = Ml;
        const int r = (int)tmp.divmod((uint32_t)p);
        Ml.mul((uint32_t)inverse_mod(r, p));     // w_l < M
        split(Ml, c.w[l]);
        int pw = 1 % p;
        for (int e = 0; e < POW2_MAX; ++e) {
            c.pow2[l][e] = (unsigned char)pw;
            pw = (pw * 2) % p;
        }
    }
}

// The device copy of the constants for n moduli (rebuilt when n changes);
// *parts receives the number of weight parts.
const CrtConst* device_const(int n, hipStream_t st, int* parts) {
    static int have = -1;
    static CrtConst host;
    void* dev = workspace(sizeof(CrtConst), 31);
    if (have != n) {
        build_const(n, host);
        HIP_CHECK(hipMemcpyAsync(dev, &host, sizeof(CrtConst), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
        have = n;
    }
    *parts = host.parts;
    return (const CrtConst*)dev;
}

template <typename T>
void launch_residues(View v, int64_t rows, int64_t rows_pad, int64_t K, int64_t nkc,
                     int64_t s_row, int64_t s_k, const int* emax, int sh, const CrtConst* cc,
                     signed char* out, int64_t plane, hipStream_t st) {
    const int64_t tiles = rows_pad / 64 * nkc;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 16384));
    hipLaunchKernelGGL(crt_residue_kernel<T>, dim3(grid), dim3(256), 0, st, (const T*)v.ptr,
                       rows, rows_pad, K, nkc, s_row, s_k, (s_row == 1) ? 1 : 0, emax, sh, cc,
                       out, plane);
    HIP_CHECK(hipGetLastError());
}

} // namespace detail_crt
using namespace detail_crt;

// Fractional bits sh and moduli for a product of depth k whose rows sit on
// average `gap` bits below their maxima: sh = ceil(need) with the slice
// scheme's accuracy target (50 + gap bits for FP64, 22 + gap for FP32), and
// the fewest moduli whose product M exceeds 4 k 2^(2 sh) (so |X| < M / 4).
// Returns 0 moduli when even all of them are too few, or k is too deep for
// exact int32 sums.
int crt_choose_moduli(Dt dt, double gap_bits, int64_t k, int* sh_out) {
    const double need = (dt == Dt::F64 ? 50.0 : 22.0) + gap_bits;
    const int sh = (int)std::ceil(need);
    if (sh_out) *sh_out = sh;
    if (k > ((int64_t)1 << 17)) return 0;   // 128^2 k < 2^31
    const double bits = 2.0 * sh + std::ceil(std::log2((double)std::max<int64_t>(k, 2))) + 2.0;
    for (int n = 2; n <= CRT_MAX_MODULI; ++n)
        if (log2_modulus(n) >= bits) return n;
    return 0;
}

// The largest sh that n moduli support at depth k (forced moduli counts).
int crt_shift_for(int n, int64_t k) {
    const double bits = log2_modulus(n) - std::ceil(std::log2((double)std::max<int64_t>(k, 2))) - 2.0;
    return std::max(1, (int)std::floor(bits / 2.0));
}

SplitResult crt_run(const GemmDesc& d, int* emax, const int* h_emax, int n, int sh,
                    hipStream_t st) {
    SplitResult res{-n, 0.0};
    int parts = 0;
    const CrtConst* cc = device_const(n, st, &parts);
    // A-priori bound ||C - C~||_F <= |alpha| k 2^(1 - sh) (1 + 2^-sh)
    //   * sqrt(sum_i 4^e_i) * sqrt(sum_j 4^f_j).
    res.err_bound = (double)(std::fabs(d.alpha) * (long double)d.k * ldexpl(1.0L, 1 - sh) *
                             (1.0L + ldexpl(1.0L, -sh)) * sqrtl(sum_pow4(h_emax, d.m)) *
                             sqrtl(sum_pow4(h_emax + d.m, d.n)));

    // Residue planes of op(A) (rows i) and op(B) (columns j).
    const int64_t nkc = (d.k + KC - 1) / KC;
    const int64_t m_pad = (d.m + ROWS_ALIGN - 1) / ROWS_ALIGN * ROWS_ALIGN;
    const int64_t n_pad = (d.n + ROWS_ALIGN - 1) / ROWS_ALIGN * ROWS_ALIGN;
    const int64_t a_plane = nkc * m_pad * KC, b_plane = nkc * n_pad * KC;
    signed char* ar = (signed char*)workspace((size_t)a_plane * n, 28);
    signed char* br = (signed char*)workspace((size_t)b_plane * n, 29);
    if (d.dt == Dt::F64) {
        launch_residues<double>(d.a, d.m, m_pad, d.k, nkc, d.a.rs, d.a.cs, emax, sh, cc, ar,
                                a_plane, st);
        launch_residues<double>(d.b, d.n, n_pad, d.k, nkc, d.b.cs, d.b.rs, emax + d.m, sh, cc,
                                br, b_plane, st);
    } else {
        launch_residues<float>(d.a, d.m, m_pad, d.k, nkc, d.a.rs, d.a.cs, emax, sh, cc, ar,
                               a_plane, st);
        launch_residues<float>(d.b, d.n, n_pad, d.k, nkc, d.b.cs, d.b.rs, emax + d.m, sh, cc, br,
                               b_plane, st);
    }

    // Residue products, one plane per modulus. Y is the operand whose index
    // runs along C's unit stride.
    CrtGemmArgs g;
    const bool c_col = (d.c.rs == 1);
    const int* ex;
    const int* ey;
    int64_t c_sx, c_sy;
    if (c_col) {
        g.xs = br; g.ys = ar; g.x_plane = b_plane; g.y_plane = a_plane;
        g.x_pad = n_pad; g.y_pad = m_pad; g.nx = d.n; g.ny = d.m;
        ex = emax + d.m; ey = emax; c_sx = d.c.cs; c_sy = d.c.rs;
    } else {
        g.xs = ar; g.ys = br; g.x_plane = a_plane; g.y_plane = b_plane;
        g.x_pad = m_pad; g.y_pad = n_pad; g.nx = d.m; g.ny = d.n;
        ex = emax; ey = emax + d.m; c_sx = d.c.rs; c_sy = d.c.cs;
    }
    g.nkc = nkc;
    g.ld_out = g.ny;
    g.out_plane = g.nx * g.ny;
    g.out = (unsigned char*)workspace((size_t)g.out_plane * n, 30);
    g.cc = cc;
    const dim3 grid((unsigned)((g.x_pad / TILE) * (g.y_pad / TILE)), (unsigned)n);
    hipLaunchKernelGGL(crt_gemm_kernel, grid, dim3(NT), 0, st, g);
    HIP_CHECK(hipGetLastError());

    // Reconstruction into C.
    if (d.dt == Dt::F64)
        launch_combine<double>(g.out, g.out_plane, g.ld_out, g.nx, g.ny, parts, cc, ex, ey, sh,
                               (double*)d.c.ptr, c_sx, c_sy, d.alpha, d.beta, st);
    else
        launch_combine<float>(g.out, g.out_plane, g.ld_out, g.nx, g.ny, parts, cc, ex, ey, sh,
                              (float*)d.c.ptr, c_sx, c_sy, d.alpha, d.beta, st);
    return res;
}

} // namespace blis_amd
