/*
 * blis_amd.h — extensions of the MI355X BLIS library that have no
 * counterpart in the reference (BLIS is a synchronous CPU library with no
 * notion of streams, device plans or tile selection).
 *
 * Include after blis.h.
 */
#ifndef BLIS_AMD_H
#define BLIS_AMD_H

#include "blis.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- execution control -------------------------------------------------
   Run subsequent calls on a caller-provided hipStream_t (NULL restores the
   library's own stream, a blocking stream that orders with the legacy
   default stream). With sync == 0 calls return as soon as
   the work is enqueued (only meaningful for device-resident operands); with
   sync != 0 (the default) every call waits for its results, as BLIS does. */
void  bli_amd_set_stream( void* hip_stream, int sync );
void* bli_amd_get_stream( void );
void  bli_amd_synchronize( void );

/* ---- GEMM configuration ------------------------------------------------
   Force the tile configuration used by subsequent GEMMs whose caller gives
   none (tile = index into the gfx950 tile table, -1 = built-in heuristic),
   the raster group height, the XCD remapping switch and the k split
   (split_k > 1: up to 64 partitions of whole 64-deep k blocks, each computed
   into library workspace, then added in partition order and combined with
   beta*C; 1 = one pass, or the built-in rule when tile is -1, which splits
   only products with m, n <= k/4 and k >= 2048). Results depend on the k
   partition only, never on the tile index or on how it was chosen. */
void bli_amd_gemm_set_config( int tile, int group_m, int xcd_remap, int split_k );
int  bli_amd_gemm_tile_count( int is_double );
int  bli_amd_gemm_default_tile( int is_double, long long m, long long n, long long k );

/* ---- GEMM without device allocation -------------------------------------
   No counterpart in the reference. For callers that cannot allocate device
   memory: code under HIP graph capture, serving loops, harnesses that
   forbid hipMalloc. C := beta*C + alpha*op(A)*op(B) for dt BLIS_FLOAT or
   BLIS_DOUBLE, with the same kernels and default configuration as
   bli_?gemm (the results are the same bits). A, B and C must be device
   memory; alpha and beta are host pointers. The call is asynchronous on
   `stream` (a hipStream_t; NULL is the null stream) and allocates nothing:
   the scratch it needs (split-k partitions, packed copies of general-stride
   operands) comes from one fixed 64 MiB array per GPU that the library
   holds until the process ends. Every call on a GPU shares that array, so
   calls that overlap on one GPU race. Returns BLIS_SUCCESS;
   BLIS_INSUFFICIENT_STACK_BUF_SIZE when the problem needs more scratch than
   the array holds (nothing is run); BLIS_FAILURE for host operands;
   BLIS_INVALID_DATATYPE, BLIS_INVALID_TRANS, BLIS_NEGATIVE_DIMENSION or
   BLIS_NULL_POINTER for invalid arguments. */
int bli_amd_gemm_noalloc( num_t dt, trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                          const void* alpha, const void* a, inc_t rs_a, inc_t cs_a,
                          const void* b, inc_t rs_b, inc_t cs_b, const void* beta,
                          void* c, inc_t rs_c, inc_t cs_c, void* stream );

/* ---- batched GEMM plans ------------------------------------------------
   A plan holds a batch (same arguments as bli_?gemm_batch, operands already
   in device memory) as device descriptors and LPT-ordered work lists. Each
   execute runs the batch as persistent kernels, one per tile variant (or,
   when the autotuner finds it faster, as one dispatch whose workgroups take
   tiles from per-XCD device counters); use_graph != 0 replays it as a HIP
   graph captured on first use. The plan keeps the pointers it was built
   with, and one plan's executions must not overlap. */
void* bli_amd_dgemm_batch_plan_create( const trans_t* transa, const trans_t* transb,
                                       const dim_t* m, const dim_t* n, const dim_t* k,
                                       const double* alpha,
                                       const double* const* a, const inc_t* rs_a, const inc_t* cs_a,
                                       const double* const* b, const inc_t* rs_b, const inc_t* cs_b,
                                       const double* beta,
                                       double* const* c, const inc_t* rs_c, const inc_t* cs_c,
                                       dim_t group_count, const dim_t* group_size );
void* bli_amd_sgemm_batch_plan_create( const trans_t* transa, const trans_t* transb,
                                       const dim_t* m, const dim_t* n, const dim_t* k,
                                       const float* alpha,
                                       const float* const* a, const inc_t* rs_a, const inc_t* cs_a,
                                       const float* const* b, const inc_t* rs_b, const inc_t* cs_b,
                                       const float* beta,
                                       float* const* c, const inc_t* rs_c, const inc_t* cs_c,
                                       dim_t group_count, const dim_t* group_size );
void  bli_amd_gemm_batch_plan_execute( void* plan, int use_graph );
int   bli_amd_gemm_batch_plan_work_items( void* plan );
void  bli_amd_gemm_batch_plan_destroy( void* plan );
/* Batch schedule autotuner: times every kernel mode of the plan (k slabs in
   flight, edge-tile loading) on the plan's own problems, keeps the fastest
   and returns it. C is saved and restored, so the caller's data is left as
   it was. All modes compute bit-identical results. */
int   bli_amd_gemm_batch_plan_autotune( void* plan, int reps );
void  bli_amd_gemm_batch_plan_set_mode( void* plan, int mode );
int   bli_amd_gemm_batch_plan_get_mode( void* plan );
/* on != 0: one dispatch for all tile variants; 0 (the default): one
   persistent kernel per variant (tile size x operand layouts), one after
   another. The autotuner chooses between the two. */
void  bli_amd_gemm_batch_plan_set_single_dispatch( void* plan, int on );
int   bli_amd_gemm_batch_plan_get_single_dispatch( void* plan );

/* ---- split-precision GEMM ----------------------------------------------
   Same arguments and semantics as bli_?gemm (C := beta*C + alpha*op(A)*op(B);
   beta == 0 does not read C; alpha == 0 or k == 0 scales C by beta). The
   product is computed from exact INT8 slices of the operands (Ozaki scheme:
   rows of op(A) and columns of op(B) are scaled by powers of two and split
   into 7-bit digits, slice products are accumulated exactly in int32 and
   summed in FP64).
   A second scheme (Ozaki scheme II) scales the rows and columns to integers
   and forms their product exactly from its residues modulo up to 32
   pairwise coprime moduli, one INT8 GEMM per modulus, rebuilt in FP64 by the
   Chinese remainder theorem (k <= 2^17).
   nslices > 0 forces the slice scheme with that many slices (at most 12;
   1 runs as 2, the fewest the slice kernels take);
   nslices in [-32, -2] forces the modular scheme with -nslices moduli;
   nslices == 0 chooses from the operands' exponent range: the modular
   scheme when its moduli are fewer than the slice scheme's products, else
   the slice scheme. If the operands hold Inf/NaN, or if
   the choice is past the point where the emulation is slower than native
   MFMA, the native bli_?gemm path runs instead.
   Returns the slice count (> 0), minus the number of moduli (< 0), or 0
   when the native path ran. If err_bound is non-NULL, it receives an
   a-priori bound on ||C_split - alpha*op(A)*op(B) - beta*C||_F from the
   truncation (0 when the native path ran). The call waits for the device
   to read back the operand statistics. */
int bli_amd_dgemm_split( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                         const double* alpha, const double* a, inc_t rs_a, inc_t cs_a,
                         const double* b, inc_t rs_b, inc_t cs_b, const double* beta,
                         double* c, inc_t rs_c, inc_t cs_c, int nslices, double* err_bound );
int bli_amd_sgemm_split( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                         const float* alpha, const float* a, inc_t rs_a, inc_t cs_a,
                         const float* b, inc_t rs_b, inc_t cs_b, const float* beta,
                         float* c, inc_t rs_c, inc_t cs_c, int nslices, double* err_bound );
/* Slice count the policy picks for a mean exponent gap (bits between a row's
   largest exponent and its entries' exponents). */
int bli_amd_gemm_split_slices( int is_double, double gap_bits );
/* Largest slice count the automatic policy uses before it falls back to the
   native path (the break-even point against native MFMA). */
int bli_amd_gemm_split_max_slices( int is_double );
/* Moduli the automatic policy uses for a mean exponent gap at depth k (0 when
   the modular scheme does not apply), and the fractional bits kept per
   operand entry (*sh, if non-NULL). */
int bli_amd_gemm_split_moduli( int is_double, double gap_bits, dim_t k, int* sh );

/* ---- configuration space, autotuner and tile selector -------------------
   A GEMM configuration is a tile of the gfx950 tile table combined with a
   tile schedule (raster group height, XCD remapping). All configurations
   compute bit-identical results; they differ in speed only.
   Configuration c: tile = c / 3, schedule = c % 3. */
int  bli_amd_gemm_config_count( int is_double );
int  bli_amd_gemm_config_describe( int is_double, int cfg, int* tile, int* group_m,
                                   int* xcd_remap, int* bm, int* bn, int* bk );
/* Time every configuration on one column-major shape (library-owned random
   operands with the given leading dimensions); ms receives
   bli_amd_gemm_config_count() milliseconds. Returns 0, or -1 for invalid
   arguments. */
int  bli_amd_gemm_autotune( int is_double, trans_t transa, trans_t transb,
                            dim_t m, dim_t n, dim_t k, inc_t lda, inc_t ldb, inc_t ldc,
                            float* ms );
/* The learned selector's 16 input features for a column-major shape. */
void bli_amd_gemm_features( int is_double, trans_t transa, trans_t transb,
                            dim_t m, dim_t n, dim_t k, inc_t lda, inc_t ldb, inc_t ldc,
                            float* out );
/* How GEMMs without an explicit configuration pick one:
   0 = built-in size heuristic (default), 1 = learned (the MLP net, or the
   library's shipped weights when net is NULL), 2 = exhaustive autotuning on
   first use of a shape. Choices are cached per shape. The environment
   variable BLIS_AMD_SELECTOR={heuristic,mlp,exhaustive} sets the initial
   mode (BLIS_AMD_SELECTOR_WEIGHTS names a weight file). */
void bli_amd_gemm_set_selector( int mode, void* net );
int  bli_amd_gemm_get_selector( void );
int  bli_amd_gemm_selected_config( int is_double, trans_t transa, trans_t transb,
                                   dim_t m, dim_t n, dim_t k, inc_t lda, inc_t ldb, inc_t ldc );

/* ---- the selector network (FP32 MLP 16 -> 64 -> 64 -> configurations) ----
   Trained and evaluated by HIP kernels. Targets are log2(t_c / t_best) per
   configuration; the prediction is the argmin. Parameters are laid out as
   W1[64][16] b1[64] W2[64][64] b2[64] W3[C][64] b3[C]; mu_sd holds the
   feature normalisation mu[16] then sd[16]. Feature/target arrays are host
   arrays, row-major. */
int    bli_amd_mlp_param_count( void );
void*  bli_amd_mlp_create( unsigned seed );
void   bli_amd_mlp_destroy( void* net );
void   bli_amd_mlp_set_params( void* net, const float* params, const float* mu_sd );
void   bli_amd_mlp_get_params( void* net, float* params, float* mu_sd );
double bli_amd_mlp_loss_and_grad( void* net, const float* feats, const float* targets, int n,
                                  float* grad );
double bli_amd_mlp_train( void* net, const float* feats, const float* targets, int n,
                          int epochs, int batch, float lr, unsigned seed );
void   bli_amd_mlp_predict( void* net, const float* feats, int n, float* scores, int* best );
int    bli_amd_mlp_save( void* net, const char* path );
void*  bli_amd_mlp_load( const char* path );

#ifdef __cplusplus
}
#endif

#endif /* BLIS_AMD_H */
