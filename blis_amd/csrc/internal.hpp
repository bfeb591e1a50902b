// internal.hpp — shared host-side declarations of the MI355X BLIS library.
//
// Every public entry point (typed, object and BLAS symbols) reduces its
// arguments to one of the plain descriptors below and hands it to a launch
// planner (gemm.hip, trsm.hip, level12.hip, batch.hip). The descriptors hold
// device pointers only; host-pointer staging happens before they are built.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/blis.h"

#define BLIS_AMD_HIDDEN __attribute__((visibility("hidden")))

namespace blis_amd {

enum class Dt : int { F32 = 0, F64 = 1 };

// Largest m, n and k of the single-call small-problem path (BLIS's
// MT/NT/KT thresholds, frame/base/bli_cntx.h:186-193; sup.hip).
constexpr int64_t SMALL_GEMM_MAX_DIM = 512;

inline size_t dt_size(Dt d) { return d == Dt::F64 ? 8 : 4; }

// Strided 2-D view of device memory: element (i, j) lives at
// ptr + (i*rs + j*cs) * elem_size.
struct View {
    void*   ptr = nullptr;
    int64_t rs = 1;
    int64_t cs = 1;
};

// C := beta*C + alpha*opA(A)*opB(B), after transposition has been folded into
// the strides: A_eff(i,p) = A[i*a.rs + p*a.cs], B_eff(p,j) = B[p*b.rs + j*b.cs].
struct GemmDesc {
    Dt      dt = Dt::F64;
    int64_t m = 0, n = 0, k = 0;
    View    a, b, c;
    double  alpha = 1.0, beta = 0.0;
};

// Tile/schedule configuration chosen by the planner (or the selector).
struct GemmConfig {
    int tile = -1;     // index into the tile table (gemm.hip); -1 = heuristic
    int group_m = 8;   // raster group height in tiles
    int xcd_remap = 1; // map consecutive logical tiles onto one XCD
    int split_k = 1;   // number of k partitions (>1 uses a workspace reduction)
};

// ---- runtime (runtime.cpp) ----
BLIS_AMD_HIDDEN void        init_once();
BLIS_AMD_HIDDEN hipStream_t stream();
BLIS_AMD_HIDDEN int         device_id();
BLIS_AMD_HIDDEN void        sync_if_needed();
// Grow-only device scratch, one buffer per slot: 0-1 GEMM operand packing
// and transposes, 2-5 split-precision slices and statistics, 5-7 level-1/2
// partials, 8-10 autotuner operands, 11-12 batch autotuner, 13-16 TRSM,
// 17 GEMM split-k partitions, 18-19 TRSM 512-row inverses, 20-21 dense
// copies of structured operands (symm/trmm) and trmm's copy of B, 22-24
// complex GEMM (A in 1e format, B in 1r format, a column-major copy of C),
// 25-27 mixed-datatype GEMM (A and B in the computation type, the product).
// 28-31 split precision's modular scheme (residues of A and of B, residue
// products, constants).
constexpr int WS_SLOTS = 32;
BLIS_AMD_HIDDEN void*       workspace(size_t bytes, int slot);
BLIS_AMD_HIDDEN bool        synchronous();   // calls wait for their results
BLIS_AMD_HIDDEN bool        is_device_pointer(const void* p);
BLIS_AMD_HIDDEN void        check_hip(hipError_t e, const char* what, const char* file, int line);
BLIS_AMD_HIDDEN void        abort_with(err_t code, const char* where);
BLIS_AMD_HIDDEN std::mutex& api_mutex();

#define HIP_CHECK(expr) ::blis_amd::check_hip((expr), #expr, __FILE__, __LINE__)

// Host staging of one strided operand. If the pointer is host memory the
// touched span is copied to device memory (and back on release when it is
// an output); device pointers pass through unchanged.
struct Staged {
    void*       dev = nullptr;   // device address matching the caller's origin
    void*       host = nullptr;  // original host origin (nullptr if device)
    void*       buf = nullptr;   // device allocation (nullptr if device)
    int64_t     lo = 0;          // lowest element offset touched (<= 0)
    size_t      bytes = 0;
    bool        writeback = false;
    ~Staged();
    Staged() = default;
    Staged(const Staged&) = delete;
    Staged& operator=(const Staged&) = delete;
};
// Device allocations seen during one call: a pointer inside a known range
// is device memory without another runtime query (a batch of 10,000
// problems in a few arenas then asks the runtime a few times, not 30,000).
// Valid for one call only: memory may be freed and reused between calls.
struct DeviceRanges {
    std::vector<std::pair<uintptr_t, uintptr_t>> r;   // [begin, end), sorted by begin
    bool is_device(const void* p);
};
BLIS_AMD_HIDDEN void stage(Staged& s, const void* p, Dt dt, int64_t m, int64_t n,
                           int64_t rs, int64_t cs, bool is_output,
                           DeviceRanges* ranges = nullptr);
BLIS_AMD_HIDDEN void stage_vec(Staged& s, const void* p, Dt dt, int64_t n,
                               int64_t inc, bool is_output);
// A complex matrix (real precision dt, strides in complex elements).
BLIS_AMD_HIDDEN void stage_complex(Staged& s, const void* p, Dt dt, int64_t m, int64_t n,
                                   int64_t rs, int64_t cs, bool is_output);

// ---- kernels' launchers ----
BLIS_AMD_HIDDEN void gemm_run(const GemmDesc& d, const GemmConfig* cfg, hipStream_t st);
// count products of one shape at fixed element strides between their A, B
// and C (column-stored C; operands with a unit stride), in one launch.
BLIS_AMD_HIDDEN void gemm_strided_run(const GemmDesc& d, int count, int64_t sa, int64_t sb,
                                      int64_t sc, hipStream_t st);
BLIS_AMD_HIDDEN void scalm_run(Dt dt, int64_t m, int64_t n, double beta, View c, hipStream_t st);

// ---- structured level-3 (struc.hip) ----
// gemmt: C := beta*C + alpha*A_eff*B_eff on the uploc triangle of the m x m C
// (diagonal offset doff); syrk/syr2k are built from it.
BLIS_AMD_HIDDEN void gemmt_run(const GemmDesc& d, uplo_t uploc, int64_t doff, hipStream_t st);
// symm: C := beta*C + alpha*A*B (left) or alpha*B*A (right), A symmetric and
// stored in its uploa triangle.
BLIS_AMD_HIDDEN void symm_run(Dt dt, side_t side, uplo_t uploa, int64_t m, int64_t n,
                              double alpha, View a, View b, double beta, View c, hipStream_t st);
// trmm3: C := beta*C + alpha*op(A)*B_eff (left) or alpha*B_eff*op(A) (right),
// A triangular (uploa, diaga, diagonal offset doff); B_eff has transb folded in.
BLIS_AMD_HIDDEN void trmm3_run(Dt dt, side_t side, uplo_t uploa, trans_t transa, diag_t diaga,
                               int64_t doff, int64_t m, int64_t n, double alpha, View a, View b,
                               double beta, View c, hipStream_t st);
// trmm: B := alpha*op(A)*B (left) or alpha*B*op(A) (right), in place.
BLIS_AMD_HIDDEN void trmm_run(Dt dt, side_t side, uplo_t uploa, trans_t transa, diag_t diaga,
                              int64_t doff, int64_t m, int64_t n, double alpha, View a, View b,
                              hipStream_t st);
BLIS_AMD_HIDDEN void trsm_run(Dt dt, side_t side, uplo_t uplo, trans_t trans, diag_t diag,
                              int64_t m, int64_t n, double alpha, View a, View b,
                              hipStream_t st);
BLIS_AMD_HIDDEN void axpyv_run(Dt dt, int64_t n, double alpha, const void* x, int64_t incx,
                               void* y, int64_t incy, hipStream_t st);
BLIS_AMD_HIDDEN void dotv_run(Dt dt, int64_t n, const void* x, int64_t incx,
                              const void* y, int64_t incy, void* rho_dev, hipStream_t st);
BLIS_AMD_HIDDEN void gemv_run(Dt dt, trans_t trans, int64_t m, int64_t n, double alpha,
                              View a, const void* x, int64_t incx, double beta,
                              void* y, int64_t incy, hipStream_t st);
BLIS_AMD_HIDDEN void scalv_run(Dt dt, int64_t n, double beta, void* y, int64_t incy,
                               hipStream_t st);

struct BatchProblem {
    int64_t m, n, k;
    View    a, b, c;      // transposition folded into strides, as in GemmDesc
    double  alpha, beta;
};
BLIS_AMD_HIDDEN void gemm_batch_run(Dt dt, const std::vector<BatchProblem>& probs,
                                    hipStream_t st);

// ---- complex and mixed-datatype GEMM (complex.hip) ----
// C := beta*C + alpha*op(A)*op(B) for complex operands of real precision dt.
// Views are in complex elements with transposition folded into the strides;
// conja/conjb conjugate the operand.
struct ZGemmDesc {
    Dt      dt = Dt::F64;
    int64_t m = 0, n = 0, k = 0;
    View    a, b, c;
    bool    conja = false, conjb = false;
    double  alpha_re = 1.0, alpha_im = 0.0, beta_re = 0.0, beta_im = 0.0;
};
BLIS_AMD_HIDDEN void zgemm_run(const ZGemmDesc& z, hipStream_t st);
BLIS_AMD_HIDDEN void zscalm_run(Dt dt, int64_t m, int64_t n, double br, double bi, View c,
                                hipStream_t st);
// Mixed-datatype element conversion: dst (column-major, ld ldd, complex
// when dst_cx) := src (strided, complex when src_cx); a complex source going
// to a real destination gives its real part, a real source going to a
// complex destination gets a zero imaginary part.
BLIS_AMD_HIDDEN void md_convert_run(Dt sdt, bool src_cx, const void* src, int64_t m, int64_t n,
                                    int64_t rs, int64_t cs, Dt ddt, bool dst_cx, void* dst,
                                    int64_t ldd, hipStream_t st);
// C := beta*C + P for a product P (column-major, ld ldp) in another datatype.
// Complex C and P: complex beta. Complex C, real P: the real parts are
// updated with real(beta), imaginary parts untouched. Real C, complex P: the
// real part of P is added. beta == 0 does not read C.
BLIS_AMD_HIDDEN void md_accum_run(Dt pdt, bool p_cx, const void* p, int64_t ldp, int64_t m,
                                  int64_t n, double br, double bi, Dt cdt, bool c_cx, View c,
                                  hipStream_t st);

// ---- configuration selection (select.hip) ----
enum { SELECT_HEURISTIC = 0, SELECT_LEARNED = 1, SELECT_EXHAUSTIVE = 2 };
BLIS_AMD_HIDDEN GemmConfig config_of(int cfg);
BLIS_AMD_HIDDEN int selector_mode();
BLIS_AMD_HIDDEN int selector_pick(Dt dt, int64_t m, int64_t n, int64_t k, bool a_k, bool b_k,
                                  bool va, bool vb, bool vc);
BLIS_AMD_HIDDEN void gemm_features(Dt dt, int64_t m, int64_t n, int64_t k, bool a_k, bool b_k,
                                   bool va, bool vb, bool vc, float* f);
BLIS_AMD_HIDDEN int gemm_autotune(Dt dt, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                                  int64_t lda, int64_t ldb, int64_t ldc, float* ms_out);

// ---- split precision (split.hip) ----
constexpr int SPLIT_MAX_SLICES = 12;
constexpr int CRT_MAX_MODULI = 32;
struct SplitResult {
    int    slices;      // > 0: INT8 slices per operand (slice scheme); < 0: minus the
                        // number of moduli (modular scheme, crt.hip); 0: native path
    double err_bound;   // a-priori Frobenius bound of the truncation error
};
BLIS_AMD_HIDDEN SplitResult gemm_split_run(const GemmDesc& d, int nslices, hipStream_t st);
BLIS_AMD_HIDDEN int split_choose_slices(Dt dt, double gap_bits);
BLIS_AMD_HIDDEN int split_max_slices(Dt dt);
// Modular scheme (crt.hip): moduli and fractional bits for a gap and depth
// (0 moduli: not applicable), the fractional bits n moduli support, and the
// run itself on operand exponents already on the device (emax: m rows of
// op(A), then n columns of op(B); h_emax the host copy).
BLIS_AMD_HIDDEN int crt_choose_moduli(Dt dt, double gap_bits, int64_t k, int* sh_out);
BLIS_AMD_HIDDEN int crt_shift_for(int n, int64_t k);
// sum_i 4^e_i over the exponents e_i != INT_MIN, in long double, from a
// histogram of the exponents (one ldexpl per distinct exponent).
BLIS_AMD_HIDDEN long double sum_pow4(const int* e, int64_t cnt);
BLIS_AMD_HIDDEN SplitResult crt_run(const GemmDesc& d, int* emax, const int* h_emax,
                                    int n, int sh, hipStream_t st);

} // namespace blis_amd
