// gemm_tile.hpp — device-side building blocks of the gfx950 GEMM.
//
// BLIS computes C := beta*C + alpha*A*B with five loops around a register
// microkernel: jc/pc/ic blocking with packm copies of A and B into cache-
// resident micro-panels (frame/3/gemm/bli_gemm_blk_var{1,2,3}.c,
// frame/1m/packm/bli_packm_blk_var1.c), a macro-kernel over MR x NR tiles
// (frame/3/gemm/bli_gemm_ker_var2.c:37-284) and the microkernel itself
// (ref_kernels/3/bli_gemm_ref.c:163-317).
//
// On gfx950 the same blocking is expressed as:
//   jc/ic loops   -> the workgroup grid (one BM x BN tile of C per workgroup),
//                    rasterised in groups and remapped so that neighbouring
//                    tiles share an XCD's L2;
//   pc loop       -> the in-kernel k loop over BK-deep slabs;
//   packm         -> coalesced 16-byte global loads of a BM x BK slab of A and
//                    a BK x BN slab of B into double-buffered LDS, stored
//                    k-major ("packed panels") regardless of the source layout;
//   jr/ir + ukr   -> each wave owns a WM x WN sub-tile built from 16x16x4
//                    MFMA blocks whose accumulators stay in registers for the
//                    whole k loop.
// Inside a wave tile, MFMA block i, row r maps to wave row TM*r + i (and block
// j, column c to wave column TN*c + j). With that interleave a lane's TM A
// fragments (and TN B fragments) are contiguous in LDS, so they are fetched
// as 16-byte vectors, and each accumulator register of a lane covers TM
// consecutive rows of C, which are written back as 16-byte vectors.
//
// Semantics kept from the reference microkernel: beta == 0 overwrites C
// without reading it (bli_gemm_ref.c:257-310); edges m < BM, n < BN and a
// partial last k slab are handled in-kernel by zero-filled loads and masked
// stores.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <cstdint>

namespace blis_amd {
namespace dev {

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float  f4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mma;
// Operand layout (both types, measured on gfx950): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]. Accumulator layout differs:
//   f32 16x16x4: lane l, register r holds D[4*(l/16) + r][l % 16];
//   f64 16x16x4: lane l, register r holds D[4*r + l/16][l % 16].
// acc_row(r, g) gives the D row of register r in lane group g = l / 16.
template <> struct Mma<double> {
    using acc_t = d4_t;
    using vec_t = d2_t;             // 16-byte vector
    static constexpr int V = 2;     // elements per 16-byte vector
    __device__ static inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static constexpr int acc_row(int r, int g) { return 4 * r + g; }
};
template <> struct Mma<float> {
    using acc_t = f4_t;
    using vec_t = f4_t;
    static constexpr int V = 4;
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static constexpr int acc_row(int r, int g) { return 4 * g + r; }
};

// Workgroup tile BM x BN, k slab BK, wave tile WM x WN.
template <typename T, int BM_, int BN_, int BK_, int WM_, int WN_>
struct TileCfg {
    using M = Mma<T>;
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
    static constexpr int WGM = BM / WM, WGN = BN / WN;
    static constexpr int NW = WGM * WGN;          // waves per workgroup
    static constexpr int NT = 64 * NW;            // threads per workgroup
    static constexpr int TM = WM / 16, TN = WN / 16;
    static constexpr int V = M::V;
    // LDS row pitch: rows are padded by one 16-byte vector (a one-element
    // pad for FP64 was measured to raise bank conflicts: it costs the 16-byte
    // alignment of the fragment reads).
    static constexpr int PAD = V;
    static constexpr bool LDS_VEC = (PAD % V) == 0;   // rows start 16-byte aligned
    static constexpr int LDA = BM + PAD;
    static constexpr int LDB = BN + PAD;
    static constexpr int A_ELEMS = BK * LDA;
    static constexpr int B_ELEMS = BK * LDB;
    static constexpr int STAGE = A_ELEMS + B_ELEMS;
    static constexpr int NVA = (BM * BK / V) / NT;  // 16-byte vectors per thread
    static constexpr int NVB = (BN * BK / V) / NT;
    static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be 16-aligned");
    static_assert(BK % 4 == 0 && BK % V == 0, "BK must cover MFMA k and vectors");
    static_assert((BM * BK / V) % NT == 0 && NVA >= 1, "A slab must split evenly");
    static_assert((BN * BK / V) % NT == 0 && NVB >= 1, "B slab must split evenly");
    static_assert(BM % V == 0 && BN % V == 0, "tile must hold whole vectors");
};

// Parameters of one GEMM problem as seen by a tile routine. Layout flags are
// template parameters of the routine: A_K means A is contiguous along k
// (A_eff(i,p) = a[p + i*lda]), otherwise along m (a[i + p*lda]); B_K means B
// is contiguous along k (B_eff(p,j) = b[p + j*ldb]), otherwise along n
// (b[j + p*ldb]). C(i,j) = c[i*c_si + j*c_sj].
template <typename T>
struct TileProblem {
    const T* a;
    const T* b;
    T*       c;
    int64_t  m, n, k;
    int64_t  lda, ldb;
    int64_t  c_si, c_sj;
    T        alpha, beta;
    int      vec_a, vec_b, vec_c;   // 16-byte vector access allowed
};

// Addressing of one operand in an edge tile (or a partial last k slab),
// computed once per tile. The operand is X_eff, rows x K, with
// X_eff(r, p) = x[p + r*ld] when K_CONTIG, else x[r + p*ld]; the tile covers
// rows [row0, row0 + ROWS). Each thread owns NV 16-byte vectors of a slab,
// in the same thread -> element order as the fast path and the LDS stores.
// Elements outside the matrix are loaded from a clamped address and replaced
// by zero, so no load sits behind a per-lane branch. A wave uses 16-byte
// loads whenever none of its lanes straddles the matrix edge (a wave-uniform
// vote), otherwise element loads.
template <typename T, class TL, int ROWS, bool K_CONTIG, int NV>
struct EdgeState {
    using vec_t = typename Mma<T>::vec_t;
    static constexpr int V = TL::V;
    const T* x;
    int64_t ld, K, klast;          // klast = (K-1)*ld (R_CONTIG)
    bool vec;                      // 16-byte access allowed (aligned x, ld % V == 0)
    int64_t rows;
    int64_t off[NV];               // K_CONTIG: clamped row * ld; R_CONTIG: kk * ld
    int sub[NV];                   // K_CONTIG: k offset of the vector; R_CONTIG: kk
    int r0[NV];                    // R_CONTIG: first row of the vector
    unsigned rmask[NV];            // rows inside the matrix (bit e for element e)
    bool vecw[NV];                 // R_CONTIG: the whole wave may use vector loads

    __device__ inline void init(const T* x_, int64_t rows_, int64_t K_, int64_t ld_, int vec_,
                                int64_t row0, int tid) {
        x = x_; ld = ld_; K = K_; vec = vec_ != 0; rows = rows_;
        klast = (K - 1) * ld;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (K_CONTIG) {
                const int kc = v % (TL::BK / V), ii = v / (TL::BK / V);
                const int64_t r = row0 + ii;
                off[q] = (r < rows ? r : rows - 1) * ld;
                sub[q] = kc * V;
                rmask[q] = (r < rows) ? 0xFu : 0u;
            } else {
                const int ic = v % (ROWS / V), kk = v / (ROWS / V);
                const int64_t r = row0 + (int64_t)ic * V;
                off[q] = (int64_t)kk * ld;
                sub[q] = kk;
                unsigned m = 0;
#pragma unroll
                for (int e = 0; e < V; ++e) m |= (r + e < rows ? 1u : 0u) << e;
                rmask[q] = m;
                r0[q] = (int)r;
                const bool full = r + V <= rows, none = r >= rows && rows >= V;
                vecw[q] = __all((vec && (full || none)) ? 1 : 0) != 0;
            }
        }
    }

    template <int NR>
    __device__ inline void load(vec_t (&r)[NR], int64_t k0) const {
        static_assert(NR == NV, "register count");
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            if constexpr (K_CONTIG) {
                const int64_t pk = k0 + sub[q];
                const T* row = x + off[q];
                if (vec && k0 + TL::BK <= K) {
                    const vec_t t = *reinterpret_cast<const vec_t*>(row + pk);
#pragma unroll
                    for (int e = 0; e < V; ++e) r[q][e] = rmask[q] ? t[e] : T(0);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        const bool kin = pk + e < K;
                        const T val = row[kin ? pk + e : K - 1];
                        r[q][e] = (rmask[q] && kin) ? val : T(0);
                    }
                }
            } else {
                const bool kin = k0 + sub[q] < K;
                const T* col = x + (kin ? k0 * ld + off[q] : klast);
                if (vecw[q]) {
                    // rows r0 .. r0+V-1 all inside (vector start r0), or all
                    // outside (any aligned in-bounds vector: start 0)
                    const int rb = (rmask[q] & 1u) ? r0[q] : 0;
                    const vec_t t = *reinterpret_cast<const vec_t*>(col + rb);
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        r[q][e] = (kin && ((rmask[q] >> e) & 1u)) ? t[e] : T(0);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        const int64_t ri = r0[q] + e;
                        const T val = col[ri < rows ? ri : rows - 1];
                        r[q][e] = (kin && ((rmask[q] >> e) & 1u)) ? val : T(0);
                    }
                }
            }
        }
    }
};

// Registers holding one slab of A between its global load and LDS store.
template <typename T, class TL, bool A_K>
struct ALoader {
    using vec_t = typename Mma<T>::vec_t;
    static constexpr int V = TL::V;
    vec_t r[TL::NVA];

    // Edge tile or partial slab, element by element behind bounds checks.
    __device__ inline void load_guarded(const TileProblem<T>& p, int64_t m0, int64_t k0, int tid) {
#pragma unroll
        for (int q = 0; q < TL::NVA; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (!A_K) {
                const int ic = v % (TL::BM / V), kk = v / (TL::BM / V);
                const int64_t i = m0 + ic * V, pk = k0 + kk;
                const T* src = p.a + i + pk * p.lda;
                if (p.vec_a && pk < p.k && i + V <= p.m) {
                    r[q] = *reinterpret_cast<const vec_t*>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        T x = T(0);
                        if (pk < p.k && i + e < p.m) x = src[e];
                        r[q][e] = x;
                    }
                }
            } else {
                const int kc = v % (TL::BK / V), ii = v / (TL::BK / V);
                const int64_t i = m0 + ii, pk = k0 + kc * V;
                const T* src = p.a + pk + i * p.lda;
                if (p.vec_a && i < p.m && pk + V <= p.k) {
                    r[q] = *reinterpret_cast<const vec_t*>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        T x = T(0);
                        if (i < p.m && pk + e < p.k) x = src[e];
                        r[q][e] = x;
                    }
                }
            }
        }
    }

    // Interior tile, full slab, vector access allowed: straight-line 16-byte
    // loads with no per-lane conditions (so the loads issue back to back).
    __device__ inline void load_fast(const TileProblem<T>& p, int64_t m0, int64_t k0, int tid) {
#pragma unroll
        for (int q = 0; q < TL::NVA; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (!A_K) {
                const int ic = v % (TL::BM / V), kk = v / (TL::BM / V);
                r[q] = *reinterpret_cast<const vec_t*>(p.a + (m0 + ic * V) + (k0 + kk) * p.lda);
            } else {
                const int kc = v % (TL::BK / V), ii = v / (TL::BK / V);
                r[q] = *reinterpret_cast<const vec_t*>(p.a + (k0 + kc * V) + (m0 + ii) * p.lda);
            }
        }
    }

    __device__ inline void store(T* As, int tid) const {
#pragma unroll
        for (int q = 0; q < TL::NVA; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (!A_K) {
                const int ic = v % (TL::BM / V), kk = v / (TL::BM / V);
                if constexpr (TL::LDS_VEC) {
                    *reinterpret_cast<vec_t*>(As + kk * TL::LDA + ic * V) = r[q];
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) As[kk * TL::LDA + ic * V + e] = r[q][e];
                }
            } else {
                const int kc = v % (TL::BK / V), ii = v / (TL::BK / V);
#pragma unroll
                for (int e = 0; e < V; ++e) As[(kc * V + e) * TL::LDA + ii] = r[q][e];
            }
        }
    }
};

template <typename T, class TL, bool B_K>
struct BLoader {
    using vec_t = typename Mma<T>::vec_t;
    static constexpr int V = TL::V;
    vec_t r[TL::NVB];

    // Edge tile or partial slab, element by element behind bounds checks.
    __device__ inline void load_guarded(const TileProblem<T>& p, int64_t n0, int64_t k0, int tid) {
#pragma unroll
        for (int q = 0; q < TL::NVB; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (B_K) {
                const int kc = v % (TL::BK / V), jj = v / (TL::BK / V);
                const int64_t j = n0 + jj, pk = k0 + kc * V;
                const T* src = p.b + pk + j * p.ldb;
                if (p.vec_b && j < p.n && pk + V <= p.k) {
                    r[q] = *reinterpret_cast<const vec_t*>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        T x = T(0);
                        if (j < p.n && pk + e < p.k) x = src[e];
                        r[q][e] = x;
                    }
                }
            } else {
                const int jc = v % (TL::BN / V), kk = v / (TL::BN / V);
                const int64_t j = n0 + jc * V, pk = k0 + kk;
                const T* src = p.b + j + pk * p.ldb;
                if (p.vec_b && pk < p.k && j + V <= p.n) {
                    r[q] = *reinterpret_cast<const vec_t*>(src);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        T x = T(0);
                        if (pk < p.k && j + e < p.n) x = src[e];
                        r[q][e] = x;
                    }
                }
            }
        }
    }

    __device__ inline void load_fast(const TileProblem<T>& p, int64_t n0, int64_t k0, int tid) {
#pragma unroll
        for (int q = 0; q < TL::NVB; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (B_K) {
                const int kc = v % (TL::BK / V), jj = v / (TL::BK / V);
                r[q] = *reinterpret_cast<const vec_t*>(p.b + (k0 + kc * V) + (n0 + jj) * p.ldb);
            } else {
                const int jc = v % (TL::BN / V), kk = v / (TL::BN / V);
                r[q] = *reinterpret_cast<const vec_t*>(p.b + (n0 + jc * V) + (k0 + kk) * p.ldb);
            }
        }
    }

    __device__ inline void store(T* Bs, int tid) const {
#pragma unroll
        for (int q = 0; q < TL::NVB; ++q) {
            const int v = tid + q * TL::NT;
            if constexpr (B_K) {
                const int kc = v % (TL::BK / V), jj = v / (TL::BK / V);
#pragma unroll
                for (int e = 0; e < V; ++e) Bs[(kc * V + e) * TL::LDB + jj] = r[q][e];
            } else {
                const int jc = v % (TL::BN / V), kk = v / (TL::BN / V);
                if constexpr (TL::LDS_VEC) {
                    *reinterpret_cast<vec_t*>(Bs + kk * TL::LDB + jc * V) = r[q];
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) Bs[kk * TL::LDB + jc * V + e] = r[q][e];
                }
            }
        }
    }
};

// Rank-BK update of the wave's accumulators from one LDS stage.
template <typename T, class TL>
__device__ inline void mma_slab(const T* As, const T* Bs,
                                typename Mma<T>::acc_t (&acc)[TL::TM][TL::TN],
                                int wm0, int wn0, int lane) {
    using vec_t = typename Mma<T>::vec_t;
    constexpr int V = TL::V;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < TL::BK; ks += 4) {
        const T* ap = As + (ks + lk) * TL::LDA + wm0 + TL::TM * lr;
        const T* bp = Bs + (ks + lk) * TL::LDB + wn0 + TL::TN * lr;
        T af[TL::TM], bf[TL::TN];
        if constexpr (TL::TM % V == 0 && TL::LDS_VEC) {
#pragma unroll
            for (int q = 0; q < TL::TM / V; ++q) {
                vec_t t = *reinterpret_cast<const vec_t*>(ap + q * V);
#pragma unroll
                for (int e = 0; e < V; ++e) af[q * V + e] = t[e];
            }
        } else {
#pragma unroll
            for (int i = 0; i < TL::TM; ++i) af[i] = ap[i];
        }
        if constexpr (TL::TN % V == 0 && TL::LDS_VEC) {
#pragma unroll
            for (int q = 0; q < TL::TN / V; ++q) {
                vec_t t = *reinterpret_cast<const vec_t*>(bp + q * V);
#pragma unroll
                for (int e = 0; e < V; ++e) bf[q * V + e] = t[e];
            }
        } else {
#pragma unroll
            for (int j = 0; j < TL::TN; ++j) bf[j] = bp[j];
        }
#pragma unroll
        for (int i = 0; i < TL::TM; ++i)
#pragma unroll
            for (int j = 0; j < TL::TN; ++j)
                acc[i][j] = Mma<T>::mma(af[i], bf[j], acc[i][j]);
    }
}

// alpha*acc + beta*c as one fixed operation sequence (a product, then a
// fused multiply-add), used by every store path. Each output element then
// has the same arithmetic whatever the tile shape or store path, so every
// tile configuration gives bit-identical results.
template <typename T>
__device__ inline T epilogue(T alpha, T acc, T beta, T old) {
    return fma(beta, old, alpha * acc);
}

// Write the wave's accumulators: C := alpha*acc + beta*C (beta == 0 does not
// read C). For accumulator register r, a lane in group g = lane/16 owns the
// TM consecutive wave rows TM*acc_row(r, g) + [0, TM) of wave column
// TN*(lane%16) + j, written as 16-byte vectors when they fit.
template <typename T, class TL, bool INTERIOR>
__device__ inline void store_tile(const TileProblem<T>& p,
                                  const typename Mma<T>::acc_t (&acc)[TL::TM][TL::TN],
                                  int64_t row0, int64_t col0, int lane) {
    using vec_t = typename Mma<T>::vec_t;
    constexpr int V = TL::V;
    constexpr int TM = TL::TM;
    const int g = lane >> 4, lc = lane & 15;
    const T alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int j = 0; j < TL::TN; ++j) {
        const int64_t col = col0 + TL::TN * lc + j;
        if (!INTERIOR && col >= p.n) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t rb = row0 + (int64_t)TM * Mma<T>::acc_row(r, g);
            if constexpr (TM % V == 0) {
                if (INTERIOR || (p.vec_c && rb + TM <= p.m)) {
                    T* dst = p.c + rb + col * p.c_sj;
#pragma unroll
                    for (int q = 0; q < TM / V; ++q) {
                        vec_t o;
                        if (beta == T(0)) {
#pragma unroll
                            for (int e = 0; e < V; ++e) o[e] = alpha * acc[q * V + e][j][r];
                        } else {
                            const vec_t old = *reinterpret_cast<const vec_t*>(dst + q * V);
#pragma unroll
                            for (int e = 0; e < V; ++e)
                                o[e] = epilogue(alpha, acc[q * V + e][j][r], beta, old[e]);
                        }
                        *reinterpret_cast<vec_t*>(dst + q * V) = o;
                    }
                    continue;
                }
            }
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int64_t row = rb + i;
                if (row < p.m) {
                    T* dst = p.c + row * p.c_si + col * p.c_sj;
                    *dst = (beta == T(0)) ? alpha * acc[i][j][r]
                                          : epilogue(alpha, acc[i][j][r], beta, *dst);
                }
            }
        }
    }
}

// Stored region of a structured C (gemmt, syrk, syr2k): element (i, j) is
// written only when it lies on the stored side of the diagonal j - i == doff
// (lower: j - i <= doff; upper: j - i >= doff), BLIS's convention for
// diagonal offsets (frame/include/bli_param_macro_defs.h:638-673).
// uplo: 0 = every element, 1 = lower, 2 = upper.
struct CMask {
    int     uplo = 0;
    int64_t doff = 0;
};

// store_tile restricted to the stored region of C: a TM-row vector is
// written whole when all its rows are stored, otherwise element by element.
// The per-element arithmetic is the same epilogue() as store_tile.
template <typename T, class TL>
__device__ inline void store_tile_masked(const TileProblem<T>& p,
                                         const typename Mma<T>::acc_t (&acc)[TL::TM][TL::TN],
                                         int64_t row0, int64_t col0, int lane, CMask cm) {
    using vec_t = typename Mma<T>::vec_t;
    constexpr int V = TL::V;
    constexpr int TM = TL::TM;
    const int g = lane >> 4, lc = lane & 15;
    const T alpha = p.alpha, beta = p.beta;
    const bool lower = cm.uplo == 1;
#pragma unroll
    for (int j = 0; j < TL::TN; ++j) {
        const int64_t col = col0 + TL::TN * lc + j;
        if (col >= p.n) continue;
        // lower: rows i >= lim are stored; upper: rows i <= lim.
        const int64_t lim = col - cm.doff;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t rb = row0 + (int64_t)TM * Mma<T>::acc_row(r, g);
            if constexpr (TM % V == 0) {
                const bool all_in = lower ? rb >= lim : rb + TM - 1 <= lim;
                if (all_in && p.vec_c && rb + TM <= p.m) {
                    T* dst = p.c + rb + col * p.c_sj;
#pragma unroll
                    for (int q = 0; q < TM / V; ++q) {
                        vec_t o;
                        if (beta == T(0)) {
#pragma unroll
                            for (int e = 0; e < V; ++e) o[e] = alpha * acc[q * V + e][j][r];
                        } else {
                            const vec_t old = *reinterpret_cast<const vec_t*>(dst + q * V);
#pragma unroll
                            for (int e = 0; e < V; ++e)
                                o[e] = epilogue(alpha, acc[q * V + e][j][r], beta, old[e]);
                        }
                        *reinterpret_cast<vec_t*>(dst + q * V) = o;
                    }
                    continue;
                }
            }
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int64_t row = rb + i;
                const bool in = lower ? row >= lim : row <= lim;
                if (row < p.m && in) {
                    T* dst = p.c + row * p.c_si + col * p.c_sj;
                    *dst = (beta == T(0)) ? alpha * acc[i][j][r]
                                          : epilogue(alpha, acc[i][j][r], beta, *dst);
                }
            }
        }
    }
}

// Epilogue through LDS, for tiles whose wave rows per lane are too few for
// full-width vector access to C: the accumulators are written to LDS as a
// column-major BM x BN tile (column pitch BM + V), then each thread updates
// whole 16-byte chunks of C columns, so a wave reads and writes entire
// cache lines. The arithmetic per element is the same epilogue() as
// store_tile. smem must hold BN * (BM + V) elements; the caller has
// synchronised after the last use of smem by the k loop.
template <typename T, class TL>
__device__ inline void store_tile_lds(const TileProblem<T>& p,
                                      const typename Mma<T>::acc_t (&acc)[TL::TM][TL::TN],
                                      int64_t m0, int64_t n0, int wm0, int wn0, T* smem) {
    using vec_t = typename Mma<T>::vec_t;
    constexpr int V = TL::V;
    constexpr int TM = TL::TM;
    constexpr int LDC = TL::BM + V;
    static_assert(TL::BN * LDC <= 2 * TL::STAGE, "epilogue tile must fit the LDS stages");
    const int tid = threadIdx.x;
    const int lane = tid & 63, g = lane >> 4, lc = lane & 15;
#pragma unroll
    for (int j = 0; j < TL::TN; ++j) {
        const int col = wn0 + TL::TN * lc + j;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = wm0 + TM * Mma<T>::acc_row(r, g);
#pragma unroll
            for (int i = 0; i < TM; ++i) smem[col * LDC + row + i] = acc[i][j][r];
        }
    }
    __syncthreads();
    constexpr int CH = TL::BM / V;                 // 16-byte chunks per column
    const T alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int q0 = 0; q0 < CH * TL::BN; q0 += TL::NT) {
        const int q = q0 + tid;
        if ((CH * TL::BN) % TL::NT != 0 && q >= CH * TL::BN) break;
        const int col = q / CH, ch = q % CH;
        const int64_t gc = n0 + col, gr = m0 + (int64_t)ch * V;
        if (gc >= p.n || gr >= p.m) continue;
        const T* src = smem + col * LDC + ch * V;
        T* dst = p.c + gr * p.c_si + gc * p.c_sj;
        if (p.vec_c && gr + V <= p.m) {
            const vec_t a = *reinterpret_cast<const vec_t*>(src);
            vec_t o;
            if (beta == T(0)) {
#pragma unroll
                for (int e = 0; e < V; ++e) o[e] = alpha * a[e];
            } else {
                const vec_t old = *reinterpret_cast<const vec_t*>(dst);
#pragma unroll
                for (int e = 0; e < V; ++e) o[e] = epilogue(alpha, a[e], beta, old[e]);
            }
            *reinterpret_cast<vec_t*>(dst) = o;
        } else {
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if (gr + e >= p.m) break;
                T* d = dst + e * p.c_si;
                *d = (beta == T(0)) ? alpha * src[e] : epilogue(alpha, src[e], beta, *d);
            }
        }
    }
}

// Compute one BM x BN tile (tile coordinates tm, tn) of one problem. smem
// must hold 2*TL::STAGE elements. The caller synchronises before reusing smem.
// FAST: the tile lies inside the matrix and A, B (and C) allow 16-byte access,
// so every full k slab is loaded without bounds checks.
// PF: k slabs whose global loads are in flight while a slab is multiplied
// (1: one register buffer; 2: two, for tiles that stream their operands
// from HBM, such as the small problems of a batch).
// EDGE: how edge tiles load. false: element loads behind bounds checks
// (fewest registers: the single-GEMM kernels, whose occupancy was tuned with
// it); true: addressing precomputed once per tile (EdgeState) and 16-byte
// loads wherever a whole wave is inside the matrix (fewer instructions per
// slab: the batch kernels, where almost every tile is an edge tile).
// The loaded values, and so the arithmetic, are the same in every variant.
// Pull the C tile's cache lines into L2 ahead of the epilogue (beta != 0):
// one 4-byte global->LDS load per 128-byte line of each tile column, into a
// scratch LDS area of NT words that nothing reads. Addresses are clamped
// into the matrix. C must be column-stored (c_si == 1).
typedef __attribute__((address_space(3))) void* lds_ptr_t;
template <typename T, class TL>
__device__ inline void prefetch_c(const TileProblem<T>& p, int64_t m0, int64_t n0, int* sink) {
    constexpr int LPC = (TL::BM * (int)sizeof(T) + 127) / 128;   // lines per column
    constexpr int ROWS_PER_LINE = 128 / (int)sizeof(T);
    const int tid = threadIdx.x;
    if (p.beta == T(0) || p.c_si != 1) return;
    const int col = tid / LPC, line = tid % LPC;
    if (col >= TL::BN) return;
    const int64_t r = min(m0 + (int64_t)line * ROWS_PER_LINE, p.m - 1);
    const int64_t c = min(n0 + (int64_t)col, p.n - 1);
    int* dst = sink + (tid & ~63);
    __builtin_amdgcn_global_load_lds((const void*)(p.c + r + c * p.c_sj), (lds_ptr_t)dst, 4, 0, 0);
}

template <typename T, class TL, bool A_K, bool B_K, bool FAST, int PF, bool EDGE, bool LDS_EPI,
          bool C_PF = false, bool MASK = false>
__device__ inline void gemm_tile_impl(const TileProblem<T>& p, int64_t m0, int64_t n0, T* smem,
                                      int* sink = nullptr, CMask cm = CMask{}) {
    static_assert(PF == 1 || PF == 2, "prefetch depth");
    using acc_t = typename Mma<T>::acc_t;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % TL::WGM, wn = wave / TL::WGM;
    const int wm0 = wm * TL::WM, wn0 = wn * TL::WN;

    acc_t acc[TL::TM][TL::TN];
#pragma unroll
    for (int i = 0; i < TL::TM; ++i)
#pragma unroll
        for (int j = 0; j < TL::TN; ++j) acc[i][j] = acc_t{};

    const int64_t nk = (p.k + TL::BK - 1) / TL::BK;
    using EA = EdgeState<T, TL, TL::BM, A_K, TL::NVA>;
    using EB = EdgeState<T, TL, TL::BN, B_K, TL::NVB>;
    // With EDGE, edge tiles keep their addressing for every slab; interior
    // (FAST) tiles build it only for a partial last slab, so it holds no
    // registers during their main loop.
    EA ea;
    EB eb;
    if constexpr (EDGE && !FAST) {
        if (nk > 0) {
            ea.init(p.a, p.m, p.k, p.lda, p.vec_a, m0, tid);
            eb.init(p.b, p.n, p.k, p.ldb, p.vec_b, n0, tid);
        }
    }
    auto load = [&](ALoader<T, TL, A_K>& al, BLoader<T, TL, B_K>& bl, int64_t k0) {
        if (FAST && k0 + TL::BK <= p.k) {
            al.load_fast(p, m0, k0, tid);
            bl.load_fast(p, n0, k0, tid);
        } else if constexpr (!EDGE) {
            al.load_guarded(p, m0, k0, tid);
            bl.load_guarded(p, n0, k0, tid);
        } else if constexpr (FAST) {
            EA ta;
            EB tb;
            ta.init(p.a, p.m, p.k, p.lda, p.vec_a, m0, tid);
            tb.init(p.b, p.n, p.k, p.ldb, p.vec_b, n0, tid);
            ta.load(al.r, k0);
            tb.load(bl.r, k0);
        } else {
            ea.load(al.r, k0);
            eb.load(bl.r, k0);
        }
    };

    if constexpr (PF == 1) {
        ALoader<T, TL, A_K> al;
        BLoader<T, TL, B_K> bl;
        // k == 0 (a batch problem with k == 0 or alpha == 0) touches neither
        // A nor B: the loads clamp addresses to the last row/column, which
        // exists only when k >= 1.
        if (nk > 0) {
            load(al, bl, 0);
            al.store(smem, tid);
            bl.store(smem + TL::A_ELEMS, tid);
        }
        __syncthreads();
        if constexpr (C_PF) {
            if (nk <= 1) prefetch_c<T, TL>(p, m0, n0, sink);
        }
        for (int64_t s = 0; s < nk; ++s) {
            T* cur = smem + (s & 1) * TL::STAGE;
            T* nxt = smem + ((s + 1) & 1) * TL::STAGE;
            const bool more = (s + 1) < nk;
            if (more) load(al, bl, (s + 1) * TL::BK);
            if constexpr (C_PF) {
                if (s + 2 == nk) prefetch_c<T, TL>(p, m0, n0, sink);
            }
            mma_slab<T, TL>(cur, cur + TL::A_ELEMS, acc, wm0, wn0, lane);
            if (more) {
                al.store(nxt, tid);
                bl.store(nxt + TL::A_ELEMS, tid);
            }
            __syncthreads();
        }
    } else {
        // Two register buffers. At step s, buffer (s & 1) has already been
        // written to LDS and now receives slab s + 2; buffer ((s + 1) & 1)
        // holds slab s + 1 and is written to the other LDS stage after the
        // multiply. The loop is unrolled by two so that each buffer is a
        // fixed set of registers.
        ALoader<T, TL, A_K> a0, a1;
        BLoader<T, TL, B_K> b0, b1;
        if (nk > 0) {
            load(a0, b0, 0);
            if (nk > 1) load(a1, b1, TL::BK);
            a0.store(smem, tid);
            b0.store(smem + TL::A_ELEMS, tid);
        }
        __syncthreads();
        auto step = [&](auto buf, int64_t s) {
            constexpr int B = decltype(buf)::value;   // == s & 1
            T* cur = smem + B * TL::STAGE;
            T* nxt = smem + (1 - B) * TL::STAGE;
            if (s + 2 < nk) {
                if constexpr (B == 0) load(a0, b0, (s + 2) * TL::BK);
                else load(a1, b1, (s + 2) * TL::BK);
            }
            mma_slab<T, TL>(cur, cur + TL::A_ELEMS, acc, wm0, wn0, lane);
            if (s + 1 < nk) {
                if constexpr (B == 0) { a1.store(nxt, tid); b1.store(nxt + TL::A_ELEMS, tid); }
                else { a0.store(nxt, tid); b0.store(nxt + TL::A_ELEMS, tid); }
            }
            __syncthreads();
        };
        for (int64_t s = 0; s < nk; s += 2) {
            step(std::integral_constant<int, 0>{}, s);
            if (s + 1 < nk) step(std::integral_constant<int, 1>{}, s + 1);
        }
    }

    if constexpr (MASK)
        store_tile_masked<T, TL>(p, acc, m0 + wm0, n0 + wn0, lane, cm);
    else if constexpr (LDS_EPI)
        store_tile_lds<T, TL>(p, acc, m0, n0, wm0, wn0, smem);
    else if (FAST && p.vec_c)
        store_tile<T, TL, true>(p, acc, m0 + wm0, n0 + wn0, lane);
    else
        store_tile<T, TL, false>(p, acc, m0 + wm0, n0 + wn0, lane);
}

// LDS_EPI: write C through store_tile_lds (the batch kernels) instead of
// straight from the accumulators. C_PF: prefetch the C tile into L2 one
// slab before the epilogue (one-slab prefetch depth only), through `sink`
// (NT ints of LDS nobody reads). MASK: write only the stored region `cm` of
// a structured C (store_tile_masked).
template <typename T, class TL, bool A_K, bool B_K, int PF = 1, bool EDGE = false,
          bool LDS_EPI = false, bool C_PF = false, bool MASK = false>
__device__ inline void gemm_tile(const TileProblem<T>& p, int64_t tm, int64_t tn, T* smem,
                                 int* sink = nullptr, CMask cm = CMask{}) {
    static_assert(!C_PF || PF == 1, "C prefetch is placed for one slab in flight");
    static_assert(!MASK || !LDS_EPI, "masked stores go straight from the accumulators");
    const int64_t m0 = tm * TL::BM, n0 = tn * TL::BN;
    const bool fast = p.vec_a && p.vec_b && m0 + TL::BM <= p.m && n0 + TL::BN <= p.n;
    if (fast)
        gemm_tile_impl<T, TL, A_K, B_K, true, PF, EDGE, LDS_EPI, C_PF, MASK>(p, m0, n0, smem, sink,
                                                                            cm);
    else
        gemm_tile_impl<T, TL, A_K, B_K, false, PF, EDGE, LDS_EPI, C_PF, MASK>(p, m0, n0, smem,
                                                                             sink, cm);
}

} // namespace dev
} // namespace blis_amd
