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
   the raster group height, the XCD remapping switch and the k split. The
   arithmetic of a given tile index is fixed, so results depend only on the
   tile index, never on how it was chosen. */
void bli_amd_gemm_set_config( int tile, int group_m, int xcd_remap, int split_k );
int  bli_amd_gemm_tile_count( int is_double );
int  bli_amd_gemm_default_tile( int is_double, long long m, long long n, long long k );

/* ---- batched GEMM plans ------------------------------------------------
   A plan holds a batch (same arguments as bli_?gemm_batch, operands already
   in device memory) as device descriptors and an LPT-ordered work list. Each
   execute runs the whole batch as one persistent kernel; use_graph != 0
   replays it as a HIP graph captured on first use. The plan keeps the
   pointers it was built with. */
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

/* ---- split-precision GEMM ----------------------------------------------
   Same arguments and semantics as bli_?gemm (C := beta*C + alpha*op(A)*op(B);
   beta == 0 does not read C; alpha == 0 or k == 0 scales C by beta). The
   product is computed from exact INT8 slices of the operands (Ozaki scheme:
   rows of op(A) and columns of op(B) are scaled by powers of two and split
   into 7-bit digits, slice products are accumulated exactly in int32 and
   summed in FP64).
   nslices > 0 forces the slice count (at most 12). nslices <= 0 chooses it
   from the operands' exponent range. If the operands hold Inf/NaN, or if the
   chosen count is past the point where the emulation is slower than native
   MFMA, the native bli_?gemm path runs instead.
   Returns the slice count used (0 = native path). If err_bound is non-NULL,
   it receives an a-priori bound on ||C_split - alpha*op(A)*op(B) - beta*C||_F
   from the truncation (0 when the native path ran). The call waits for the
   device to read back the operand statistics. */
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

#ifdef __cplusplus
}
#endif

#endif /* BLIS_AMD_H */
