/*
 * blis.h — public C ABI of the MI355X (gfx950) BLIS level-3 hot path.
 *
 * This header is the drop-in boundary. It declares the subset of BLIS's
 * public interface that this library implements, with the same names,
 * argument meaning, integer widths, enum bit encodings and struct layouts as
 * BLIS, so that code written against BLIS's typed, object and BLAS interfaces
 * compiles and links against this library unchanged.
 *
 * Reference interfaces (paths relative to the BLIS source tree):
 *   integer widths        frame/include/bli_type_defs.h:79-115, 224-230
 *   info bit fields       frame/include/bli_type_defs.h:254-380
 *   enums                 frame/include/bli_type_defs.h:395-520
 *   obj_t                 frame/include/bli_type_defs.h:1233-1263
 *   typed level-3         frame/3/bli_l3_tapi_ex.c:43-99 (gemm), :532-586 (trsm)
 *   object level-3        frame/3/bli_l3_oapi_ex.c:48-148 (gemm), :692-801 (trsm)
 *   level-1v typed        frame/1/bli_l1v_tapi.c:155-191 (axpyv), :195-232 (dotv)
 *   level-2 typed         frame/2/bli_l2_tapi.c:43-123 (gemv)
 *   structured level-3    frame/3/bli_l3_tapi_ex.h:60-225, bli_l3_oapi_ex.c:151-690
 *   complex / mixed GEMM  frame/3/bli_l3_ind.c (1m), docs/MixedDatatypes.md
 *   registry / info       frame/base/bli_gks.c, bli_cntx.h:73-204, bli_info.c
 *   BLAS compatibility    frame/compat/bla_gemm.c:122-260, bla_trsm.c:121-216,
 *                         bla_gemv.c:41-141, bla_axpy.c:41-85, bla_dot.c:41-88,
 *                         extra/bla_gemm_batch.c:135-246
 *
 * Memory residency: matrix and vector arguments may be device pointers
 * (hipMalloc) or ordinary host pointers. Device pointers are used in place;
 * host pointers are staged through device memory for the duration of the
 * call. Scalars (alpha, beta) are host pointers, as in BLIS. Every call is
 * synchronous: results are visible when it returns (BLIS semantics).
 */
#ifndef BLIS_H
#define BLIS_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- version / configuration macros ---------------------------------- */

#define BLIS_VERSION_STRING "3.0-dev-mi355x"
#define BLIS_INT_TYPE_SIZE 64
#define BLIS_BLAS_INT_TYPE_SIZE 32
#define BLIS_ENABLE_BLAS 1
#define BLIS_ENABLE_TRSM_PREINVERSION 1

#ifndef TRUE
#define TRUE true
#endif
#ifndef FALSE
#define FALSE false
#endif

/* ---- integer types (bli_type_defs.h:79-115) --------------------------- */

typedef int64_t  gint_t;
typedef uint64_t guint_t;
#ifndef _DEFINED_DIM_T
#define _DEFINED_DIM_T
typedef gint_t dim_t;
#endif
typedef gint_t   inc_t;
typedef gint_t   doff_t;
typedef guint_t  siz_t;
typedef uint32_t objbits_t;

/* BLAS integers (bli_type_defs.h:224-230; bli_f2c.h:40-51). */
typedef int32_t  f77_int;
typedef char     f77_char;
typedef float    f77_float;
typedef double   f77_double;
typedef f77_int  bla_integer;
typedef f77_int  ftnlen;

/* ---- complex types and the scalar atom -------------------------------- */

typedef struct { float  real; float  imag; } scomplex;
typedef struct { double real; double imag; } dcomplex;
typedef dcomplex atom_t;
typedef void* void_fp;

/* ---- info bit fields (bli_type_defs.h:254-380) ------------------------ */

#define BLIS_DATATYPE_SHIFT            0
#define BLIS_DOMAIN_SHIFT              0
#define BLIS_PRECISION_SHIFT           1
#define BLIS_CONJTRANS_SHIFT           3
#define BLIS_TRANS_SHIFT               3
#define BLIS_CONJ_SHIFT                4
#define BLIS_UPLO_SHIFT                5
#define BLIS_UPPER_SHIFT               5
#define BLIS_DIAG_SHIFT                6
#define BLIS_LOWER_SHIFT               7
#define BLIS_UNIT_DIAG_SHIFT           8
#define BLIS_INVERT_DIAG_SHIFT         9
#define BLIS_PACK_SCHEMA_SHIFT        10
#define BLIS_PACK_PANEL_SHIFT         10
#define BLIS_PACK_FORMAT_SHIFT        11
#define BLIS_PACK_SHIFT               15
#define BLIS_PACK_REV_IF_UPPER_SHIFT  16
#define BLIS_PACK_REV_IF_LOWER_SHIFT  17
#define BLIS_PACK_BUFFER_SHIFT        18
#define BLIS_STRUC_SHIFT              20
#define BLIS_COMP_PREC_SHIFT          22
#define BLIS_SCALAR_DT_SHIFT          24

#define BLIS_DATATYPE_BITS    ( 0x7u << BLIS_DATATYPE_SHIFT )
#define BLIS_DOMAIN_BIT       ( 0x1u << BLIS_DOMAIN_SHIFT )
#define BLIS_PRECISION_BIT    ( 0x3u << BLIS_PRECISION_SHIFT )
#define BLIS_CONJTRANS_BITS   ( 0x3u << BLIS_CONJTRANS_SHIFT )
#define BLIS_TRANS_BIT        ( 0x1u << BLIS_TRANS_SHIFT )
#define BLIS_CONJ_BIT         ( 0x1u << BLIS_CONJ_SHIFT )
#define BLIS_UPLO_BITS        ( 0x7u << BLIS_UPLO_SHIFT )
#define BLIS_UPPER_BIT        ( 0x1u << BLIS_UPPER_SHIFT )
#define BLIS_DIAG_BIT         ( 0x1u << BLIS_DIAG_SHIFT )
#define BLIS_LOWER_BIT        ( 0x1u << BLIS_LOWER_SHIFT )
#define BLIS_UNIT_DIAG_BIT    ( 0x1u << BLIS_UNIT_DIAG_SHIFT )
#define BLIS_INVERT_DIAG_BIT  ( 0x1u << BLIS_INVERT_DIAG_SHIFT )
#define BLIS_PACK_SCHEMA_BITS ( 0x3Fu << BLIS_PACK_SCHEMA_SHIFT )
#define BLIS_STRUC_BITS       ( 0x3u << BLIS_STRUC_SHIFT )
#define BLIS_COMP_PREC_BIT    ( 0x3u << BLIS_COMP_PREC_SHIFT )
#define BLIS_SCALAR_DT_BITS   ( 0x7u << BLIS_SCALAR_DT_SHIFT )

#define BLIS_BITVAL_REAL           0x0u
#define BLIS_BITVAL_COMPLEX        BLIS_DOMAIN_BIT
#define BLIS_BITVAL_SINGLE_PREC    0x0u
#define BLIS_BITVAL_DOUBLE_PREC    ( 0x1u << BLIS_PRECISION_SHIFT )
#define BLIS_BITVAL_FLOAT_TYPE     0x0u
#define BLIS_BITVAL_SCOMPLEX_TYPE  BLIS_DOMAIN_BIT
#define BLIS_BITVAL_DOUBLE_TYPE    BLIS_BITVAL_DOUBLE_PREC
#define BLIS_BITVAL_DCOMPLEX_TYPE  ( BLIS_DOMAIN_BIT | BLIS_BITVAL_DOUBLE_PREC )
#define BLIS_BITVAL_INT_TYPE       0x04u
#define BLIS_BITVAL_CONST_TYPE     0x05u
#define BLIS_BITVAL_TRANS          BLIS_TRANS_BIT
#define BLIS_BITVAL_CONJ           BLIS_CONJ_BIT
#define BLIS_BITVAL_CONJ_TRANS     ( BLIS_CONJ_BIT | BLIS_TRANS_BIT )
#define BLIS_BITVAL_ZEROS          0x0u
#define BLIS_BITVAL_UPPER          ( BLIS_UPPER_BIT | BLIS_DIAG_BIT )
#define BLIS_BITVAL_LOWER          ( BLIS_LOWER_BIT | BLIS_DIAG_BIT )
#define BLIS_BITVAL_DENSE          BLIS_UPLO_BITS
#define BLIS_BITVAL_UNIT_DIAG      BLIS_UNIT_DIAG_BIT
#define BLIS_BITVAL_INVERT_DIAG    BLIS_INVERT_DIAG_BIT
#define BLIS_BITVAL_GENERAL        0x0u
#define BLIS_BITVAL_HERMITIAN      ( 0x1u << BLIS_STRUC_SHIFT )
#define BLIS_BITVAL_SYMMETRIC      ( 0x2u << BLIS_STRUC_SHIFT )
#define BLIS_BITVAL_TRIANGULAR     ( 0x3u << BLIS_STRUC_SHIFT )

/* ---- enums (bli_type_defs.h:395-520) ---------------------------------- */

typedef enum trans_e {
    BLIS_NO_TRANSPOSE      = 0x0,
    BLIS_TRANSPOSE         = BLIS_BITVAL_TRANS,
    BLIS_CONJ_NO_TRANSPOSE = BLIS_BITVAL_CONJ,
    BLIS_CONJ_TRANSPOSE    = BLIS_BITVAL_CONJ_TRANS
} trans_t;

typedef enum conj_e {
    BLIS_NO_CONJUGATE = 0x0,
    BLIS_CONJUGATE    = BLIS_BITVAL_CONJ
} conj_t;

typedef enum uplo_e {
    BLIS_ZEROS = BLIS_BITVAL_ZEROS,
    BLIS_LOWER = BLIS_BITVAL_LOWER,
    BLIS_UPPER = BLIS_BITVAL_UPPER,
    BLIS_DENSE = BLIS_BITVAL_DENSE
} uplo_t;

typedef enum side_e {
    BLIS_LEFT = 0x0,
    BLIS_RIGHT
} side_t;

typedef enum diag_e {
    BLIS_NONUNIT_DIAG = 0x0,
    BLIS_UNIT_DIAG    = BLIS_BITVAL_UNIT_DIAG
} diag_t;

typedef enum struc_e {
    BLIS_GENERAL    = BLIS_BITVAL_GENERAL,
    BLIS_HERMITIAN  = BLIS_BITVAL_HERMITIAN,
    BLIS_SYMMETRIC  = BLIS_BITVAL_SYMMETRIC,
    BLIS_TRIANGULAR = BLIS_BITVAL_TRIANGULAR
} struc_t;

typedef enum num_e {
    BLIS_FLOAT    = BLIS_BITVAL_FLOAT_TYPE,
    BLIS_DOUBLE   = BLIS_BITVAL_DOUBLE_TYPE,
    BLIS_SCOMPLEX = BLIS_BITVAL_SCOMPLEX_TYPE,
    BLIS_DCOMPLEX = BLIS_BITVAL_DCOMPLEX_TYPE,
    BLIS_INT      = BLIS_BITVAL_INT_TYPE,
    BLIS_CONSTANT = BLIS_BITVAL_CONST_TYPE,
    BLIS_DT_LO    = BLIS_FLOAT,
    BLIS_DT_HI    = BLIS_DCOMPLEX
} num_t;

typedef enum prec_e {
    BLIS_SINGLE_PREC = BLIS_BITVAL_SINGLE_PREC,
    BLIS_DOUBLE_PREC = BLIS_BITVAL_DOUBLE_PREC
} prec_t;

/* Error codes (frame/include/bli_type_defs.h:1500-1608; same values). */
typedef enum err_e {
    BLIS_SUCCESS                              = (  -1),
    BLIS_FAILURE                              = (  -2),
    BLIS_ERROR_CODE_MIN                       = (  -9),
    BLIS_INVALID_ERROR_CHECKING_LEVEL         = ( -10),
    BLIS_UNDEFINED_ERROR_CODE                 = ( -11),
    BLIS_NULL_POINTER                         = ( -12),
    BLIS_NOT_YET_IMPLEMENTED                  = ( -13),
    BLIS_OUT_OF_BOUNDS                        = ( -14),
    BLIS_LOCK_FAILURE                         = ( -15),
    BLIS_INVALID_SIDE                         = ( -20),
    BLIS_INVALID_UPLO                         = ( -21),
    BLIS_INVALID_TRANS                        = ( -22),
    BLIS_INVALID_CONJ                         = ( -23),
    BLIS_INVALID_DIAG                         = ( -24),
    BLIS_INVALID_MACHVAL                      = ( -25),
    BLIS_EXPECTED_NONUNIT_DIAG                = ( -26),
    BLIS_INVALID_DATATYPE                     = ( -30),
    BLIS_EXPECTED_FLOATING_POINT_DATATYPE     = ( -31),
    BLIS_EXPECTED_NONINTEGER_DATATYPE         = ( -32),
    BLIS_EXPECTED_NONCONSTANT_DATATYPE        = ( -33),
    BLIS_EXPECTED_REAL_DATATYPE               = ( -34),
    BLIS_EXPECTED_INTEGER_DATATYPE            = ( -35),
    BLIS_INCONSISTENT_DATATYPES               = ( -36),
    BLIS_EXPECTED_REAL_PROJ_OF                = ( -37),
    BLIS_EXPECTED_REAL_VALUED_OBJECT          = ( -38),
    BLIS_INCONSISTENT_PRECISIONS              = ( -39),
    BLIS_NONCONFORMAL_DIMENSIONS              = ( -40),
    BLIS_EXPECTED_SCALAR_OBJECT               = ( -41),
    BLIS_EXPECTED_VECTOR_OBJECT               = ( -42),
    BLIS_UNEQUAL_VECTOR_LENGTHS               = ( -43),
    BLIS_EXPECTED_SQUARE_OBJECT               = ( -44),
    BLIS_UNEXPECTED_OBJECT_LENGTH             = ( -45),
    BLIS_UNEXPECTED_OBJECT_WIDTH              = ( -46),
    BLIS_UNEXPECTED_VECTOR_DIM                = ( -47),
    BLIS_UNEXPECTED_DIAG_OFFSET               = ( -48),
    BLIS_NEGATIVE_DIMENSION                   = ( -49),
    BLIS_INVALID_ROW_STRIDE                   = ( -50),
    BLIS_INVALID_COL_STRIDE                   = ( -51),
    BLIS_INVALID_DIM_STRIDE_COMBINATION       = ( -52),
    BLIS_EXPECTED_GENERAL_OBJECT              = ( -60),
    BLIS_EXPECTED_HERMITIAN_OBJECT            = ( -61),
    BLIS_EXPECTED_SYMMETRIC_OBJECT            = ( -62),
    BLIS_EXPECTED_TRIANGULAR_OBJECT           = ( -63),
    BLIS_EXPECTED_UPPER_OR_LOWER_OBJECT       = ( -70),
    BLIS_INVALID_3x1_SUBPART                  = ( -80),
    BLIS_INVALID_1x3_SUBPART                  = ( -81),
    BLIS_INVALID_3x3_SUBPART                  = ( -82),
    BLIS_UNEXPECTED_NULL_CONTROL_TREE         = ( -90),
    BLIS_PACK_SCHEMA_NOT_SUPPORTED_FOR_UNPACK = (-100),
    BLIS_PACK_SCHEMA_NOT_SUPPORTED_FOR_PART   = (-101),
    BLIS_EXPECTED_NONNULL_OBJECT_BUFFER       = (-110),
    BLIS_MALLOC_RETURNED_NULL                 = (-120),
    BLIS_INVALID_PACKBUF                      = (-130),
    BLIS_EXHAUSTED_CONTIG_MEMORY_POOL         = (-131),
    BLIS_INSUFFICIENT_STACK_BUF_SIZE          = (-132),
    BLIS_ALIGNMENT_NOT_POWER_OF_TWO           = (-133),
    BLIS_ALIGNMENT_NOT_MULT_OF_PTR_SIZE       = (-134),
    BLIS_EXPECTED_OBJECT_ALIAS                = (-140),
    BLIS_INVALID_ARCH_ID                      = (-150),
    BLIS_UNINITIALIZED_GKS_CNTX               = (-151),
    BLIS_INVALID_UKR_ID                       = (-152),
    BLIS_MC_DEF_NONMULTIPLE_OF_MR             = (-160),
    BLIS_MC_MAX_NONMULTIPLE_OF_MR             = (-161),
    BLIS_NC_DEF_NONMULTIPLE_OF_NR             = (-162),
    BLIS_NC_MAX_NONMULTIPLE_OF_NR             = (-163),
    BLIS_KC_DEF_NONMULTIPLE_OF_KR             = (-164),
    BLIS_KC_MAX_NONMULTIPLE_OF_KR             = (-165),
    BLIS_MR_NOT_EVEN_FOR_REAL_TYPE            = (-166),
    BLIS_PACKMR_NOT_EVEN_FOR_REAL_TYPE        = (-167),
    BLIS_NR_NOT_EVEN_FOR_REAL_TYPE            = (-168),
    BLIS_PACKNR_NOT_EVEN_FOR_REAL_TYPE        = (-169),
    BLIS_ERROR_CODE_MAX                       = (-170)
} err_t;

typedef enum errlev_e {
    BLIS_NO_ERROR_CHECKING   = 0,
    BLIS_FULL_ERROR_CHECKING
} errlev_t;

/* ---- obj_t (bli_type_defs.h:1233-1263), layout-identical ------------- */

typedef struct obj_s {
    struct obj_s* root;
    dim_t         off[2];
    dim_t         dim[2];
    doff_t        diag_off;
    objbits_t     info;
    objbits_t     info2;
    siz_t         elem_size;
    void*         buffer;
    inc_t         rs;
    inc_t         cs;
    inc_t         is;
    atom_t        scalar;
    dim_t         m_padded;
    dim_t         n_padded;
    inc_t         ps;
    inc_t         pd;
    dim_t         m_panel;
    dim_t         n_panel;
} obj_t;

/* ---- kernel/blocksize registry types (bli_type_defs.h:583-935, 1131-1150) --
   Same enumerator values as BLIS. The context itself is this library's own
   structure: its fields are read through the bli_cntx_* functions below. */

#define BLIS_NUM_FP_TYPES 4

#define BLIS_NTYPE_KER_SHIFT 28
#define BLIS_NTYPE_KER_BITS  ( 0xFu << BLIS_NTYPE_KER_SHIFT )
#define BLIS_1TYPE_KER       ( 0u << BLIS_NTYPE_KER_SHIFT )
#define BLIS_2TYPE_KER       ( 1u << BLIS_NTYPE_KER_SHIFT )
#define bli_ker_idx( ker )   ( ( kerid_t )( ( ker ) & ~BLIS_NTYPE_KER_BITS ) )
#define bli_ker_ntype( ker ) ( ( ( ( ker ) & BLIS_NTYPE_KER_BITS ) >> BLIS_NTYPE_KER_SHIFT ) + 1 )
#define BLIS_VA_END ( ( kerid_t )-1 )

typedef uint32_t kerid_t;

typedef enum ind_e { BLIS_1M = 0, BLIS_NAT, BLIS_NUM_IND_METHODS } ind_t;

typedef enum ukr_e {
    BLIS_ADDV_KER = BLIS_1TYPE_KER, BLIS_AMAXV_KER, BLIS_AXPBYV_KER, BLIS_AXPYV_KER,
    BLIS_COPYV_KER, BLIS_DOTV_KER, BLIS_DOTXV_KER, BLIS_INVERTV_KER, BLIS_INVSCALV_KER,
    BLIS_SCALV_KER, BLIS_SCAL2V_KER, BLIS_SETV_KER, BLIS_SUBV_KER, BLIS_SWAPV_KER,
    BLIS_XPBYV_KER, BLIS_AXPY2V_KER, BLIS_DOTAXPYV_KER,
    BLIS_AXPYF_KER, BLIS_DOTXF_KER, BLIS_DOTXAXPYF_KER,
    BLIS_GEMMTRSM_L_UKR, BLIS_GEMMTRSM_U_UKR, BLIS_TRSM_L_UKR, BLIS_TRSM_U_UKR,
    BLIS_GEMMTRSM1M_L_UKR, BLIS_GEMMTRSM1M_U_UKR,
    BLIS_GEMMSUP_RRR_UKR, BLIS_GEMMSUP_RRC_UKR, BLIS_GEMMSUP_RCR_UKR, BLIS_GEMMSUP_RCC_UKR,
    BLIS_GEMMSUP_CRR_UKR, BLIS_GEMMSUP_CRC_UKR, BLIS_GEMMSUP_CCR_UKR, BLIS_GEMMSUP_CCC_UKR,
    BLIS_GEMMSUP_XXX_UKR,
    BLIS_NUM_UKRS_, BLIS_NUM_UKRS = bli_ker_idx( BLIS_NUM_UKRS_ ),
    BLIS_PACKM_KER = BLIS_2TYPE_KER, BLIS_PACKM_1ER_KER, BLIS_PACKM_RO_KER,
    BLIS_PACKM_DIAG_KER, BLIS_PACKM_DIAG_1ER_KER, BLIS_PACKM_DIAG_RO_KER,
    BLIS_UNPACKM_KER,
    BLIS_GEMM_UKR,
    BLIS_GEMM1M_UKR,
    BLIS_GEMM_CCR_UKR, BLIS_GEMM_RCC_UKR, BLIS_GEMM_CRR_UKR,
    BLIS_NUM_UKR2S_, BLIS_NUM_UKR2S = bli_ker_idx( BLIS_NUM_UKR2S_ ),
    BLIS_UKRS_END_ = BLIS_VA_END
} ukr_t;

typedef enum ukr_pref_e {
    BLIS_GEMM_UKR_ROW_PREF, BLIS_GEMMTRSM_L_UKR_ROW_PREF, BLIS_GEMMTRSM_U_UKR_ROW_PREF,
    BLIS_TRSM_L_UKR_ROW_PREF, BLIS_TRSM_U_UKR_ROW_PREF,
    BLIS_GEMMSUP_RRR_UKR_ROW_PREF, BLIS_GEMMSUP_RRC_UKR_ROW_PREF,
    BLIS_GEMMSUP_RCR_UKR_ROW_PREF, BLIS_GEMMSUP_RCC_UKR_ROW_PREF,
    BLIS_GEMMSUP_CRR_UKR_ROW_PREF, BLIS_GEMMSUP_CRC_UKR_ROW_PREF,
    BLIS_GEMMSUP_CCR_UKR_ROW_PREF, BLIS_GEMMSUP_CCC_UKR_ROW_PREF,
    BLIS_GEMMSUP_XXX_UKR_ROW_PREF,
    BLIS_NUM_UKR_PREFS,
    BLIS_UKR_PREFS_END_ = BLIS_VA_END
} ukr_pref_t;

typedef enum kimpl_e {
    BLIS_REFERENCE_UKERNEL = 0, BLIS_VIRTUAL_UKERNEL, BLIS_OPTIMIZED_UKERNEL,
    BLIS_NOTAPPLIC_UKERNEL, BLIS_NUM_UKR_IMPL_TYPES
} kimpl_t;

typedef enum bszid_e {
    BLIS_KR, BLIS_MR, BLIS_NR, BLIS_MC, BLIS_KC, BLIS_NC,
    BLIS_BBM, BLIS_BBN,
    BLIS_M2, BLIS_N2,
    BLIS_AF, BLIS_DF, BLIS_XF,
    BLIS_MT, BLIS_NT, BLIS_KT,
    BLIS_KR_SUP, BLIS_MR_SUP, BLIS_NR_SUP, BLIS_MC_SUP, BLIS_KC_SUP, BLIS_NC_SUP,
    BLIS_NO_PART, BLIS_NUM_BLKSZS = BLIS_NO_PART,
    BLIS_BLKSZS_END_ = BLIS_VA_END
} bszid_t;

/* The one architecture this library targets. */
typedef enum arch_e { BLIS_ARCH_GFX950 = 0, BLIS_NUM_ARCHS } arch_t;

/* Blocksize: default (v) and maximum (e) value per datatype, indexed
   float, scomplex, double, dcomplex. */
typedef struct blksz_s {
    dim_t v[ BLIS_NUM_FP_TYPES ];
    dim_t e[ BLIS_NUM_FP_TYPES ];
} blksz_t;

typedef struct mbool_s { bool v[ BLIS_NUM_FP_TYPES ]; } mbool_t;

typedef struct auxinfo_s auxinfo_t;

/* A context: blocksizes, their multiples, kernels and kernel preferences of
   one architecture and induced method (bli_gks_query_cntx). */
typedef struct cntx_s {
    blksz_t blkszs[ BLIS_NUM_BLKSZS ];
    bszid_t bmults[ BLIS_NUM_BLKSZS ];
    void_fp ukrs[ BLIS_NUM_UKRS ][ BLIS_NUM_FP_TYPES ];
    void_fp ukr2s[ BLIS_NUM_UKR2S ][ BLIS_NUM_FP_TYPES ];
    mbool_t ukr_prefs[ BLIS_NUM_UKR_PREFS ];
    ind_t   method;
} cntx_t;

/* Runtime objects are accepted for signature compatibility and ignored:
   parallelism is chosen by the GPU launch planner. */
typedef struct rntm_s rntm_t;


#define BLIS_OBJECT_INITIALIZER \
{ NULL, { 0, 0 }, { 0, 0 }, 0, \
  0x0u | BLIS_BITVAL_DENSE | BLIS_BITVAL_GENERAL, 0x0u, sizeof( float ), \
  NULL, 0, 0, 1, { 0.0, 0.0 }, 0, 0, 0, 0, 0, 0 }

#define BLIS_OBJECT_INITIALIZER_1X1 \
{ NULL, { 0, 0 }, { 1, 1 }, 0, \
  0x0u | BLIS_BITVAL_DENSE | BLIS_BITVAL_GENERAL, 0x0u, sizeof( float ), \
  NULL, 0, 0, 1, { 0.0, 0.0 }, 0, 0, 0, 0, 0, 0 }

/* Global scalar constants (frame/base/bli_const.c). */
extern obj_t BLIS_TWO;
extern obj_t BLIS_ONE;
extern obj_t BLIS_ZERO;
extern obj_t BLIS_MINUS_ONE;
extern obj_t BLIS_MINUS_TWO;

/* ---- inline object accessors (frame/include/bli_obj_macro_defs.h) ----- */

static inline num_t   bli_obj_dt( const obj_t* o )        { return ( num_t )( o->info & BLIS_DATATYPE_BITS ); }
static inline trans_t bli_obj_conjtrans_status( const obj_t* o ) { return ( trans_t )( o->info & BLIS_CONJTRANS_BITS ); }
static inline trans_t bli_obj_onlytrans_status( const obj_t* o ) { return ( trans_t )( o->info & BLIS_TRANS_BIT ); }
static inline bool    bli_obj_has_trans( const obj_t* o ) { return ( o->info & BLIS_TRANS_BIT ) != 0; }
static inline uplo_t  bli_obj_uplo( const obj_t* o )      { return ( uplo_t )( o->info & BLIS_UPLO_BITS ); }
static inline diag_t  bli_obj_diag( const obj_t* o )      { return ( diag_t )( o->info & BLIS_UNIT_DIAG_BIT ); }
static inline struc_t bli_obj_struc( const obj_t* o )     { return ( struc_t )( o->info & BLIS_STRUC_BITS ); }
static inline dim_t   bli_obj_length( const obj_t* o )    { return o->dim[0]; }
static inline dim_t   bli_obj_width( const obj_t* o )     { return o->dim[1]; }
static inline dim_t   bli_obj_length_after_trans( const obj_t* o ) { return bli_obj_has_trans( o ) ? o->dim[1] : o->dim[0]; }
static inline dim_t   bli_obj_width_after_trans( const obj_t* o )  { return bli_obj_has_trans( o ) ? o->dim[0] : o->dim[1]; }
static inline inc_t   bli_obj_row_stride( const obj_t* o ) { return o->rs; }
static inline inc_t   bli_obj_col_stride( const obj_t* o ) { return o->cs; }
static inline void*   bli_obj_buffer( const obj_t* o )    { return o->buffer; }
static inline dim_t   bli_obj_row_off( const obj_t* o )   { return o->off[0]; }
static inline dim_t   bli_obj_col_off( const obj_t* o )   { return o->off[1]; }
static inline doff_t  bli_obj_diag_offset( const obj_t* o ) { return o->diag_off; }

static inline void bli_obj_set_conjtrans( trans_t t, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_CONJTRANS_BITS ) | ( objbits_t )t ); }
static inline void bli_obj_set_onlytrans( trans_t t, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_TRANS_BIT ) | ( ( objbits_t )t & BLIS_TRANS_BIT ) ); }
static inline void bli_obj_set_conj( conj_t c, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_CONJ_BIT ) | ( objbits_t )c ); }
static inline void bli_obj_set_uplo( uplo_t u, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_UPLO_BITS ) | ( objbits_t )u ); }
static inline void bli_obj_set_diag( diag_t d, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_UNIT_DIAG_BIT ) | ( objbits_t )d ); }
static inline void bli_obj_set_struc( struc_t s, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_STRUC_BITS ) | ( objbits_t )s ); }
static inline void bli_obj_set_diag_offset( doff_t d, obj_t* o ) { o->diag_off = d; }
static inline void bli_obj_toggle_trans( obj_t* o ) { o->info ^= BLIS_TRANS_BIT; }
/* Computation precision of a (mixed-datatype) GEMM is read from C
   (frame/include/bli_obj_macro_defs.h:149-153, 401-406). */
static inline prec_t bli_obj_comp_prec( const obj_t* o )
{ return ( prec_t )( ( o->info & BLIS_COMP_PREC_BIT ) >> BLIS_COMP_PREC_SHIFT ); }
static inline void bli_obj_set_comp_prec( prec_t p, obj_t* o )
{ o->info = ( objbits_t )( ( o->info & ~BLIS_COMP_PREC_BIT ) | ( ( objbits_t )p << BLIS_COMP_PREC_SHIFT ) ); }
static inline bool bli_obj_is_complex( const obj_t* o ) { return ( o->info & BLIS_DOMAIN_BIT ) != 0; }
static inline bool bli_dt_is_complex( num_t dt ) { return ( ( unsigned )dt & BLIS_DOMAIN_BIT ) != 0 && dt != BLIS_CONSTANT; }

/* ---- initialization / information (frame/base/bli_init.c:40-97) ------- */

void        bli_init( void );
void        bli_finalize( void );
const char* bli_info_get_version_str( void );
void        bli_error_checking_level_set( errlev_t new_level );
errlev_t    bli_error_checking_level( void );
bool        bli_error_checking_is_enabled( void );
const char* bli_error_string_for_code( gint_t code );

/* ---- object management (frame/base/bli_obj.c:38-266) ----------------- */

void bli_obj_create( num_t dt, dim_t m, dim_t n, inc_t rs, inc_t cs, obj_t* obj );
void bli_obj_create_with_attached_buffer( num_t dt, dim_t m, dim_t n, void* p,
                                          inc_t rs, inc_t cs, obj_t* obj );
void bli_obj_create_without_buffer( num_t dt, dim_t m, dim_t n, obj_t* obj );
void bli_obj_attach_buffer( void* p, inc_t rs, inc_t cs, inc_t is, obj_t* obj );
void bli_obj_create_1x1( num_t dt, obj_t* obj );
void bli_obj_create_1x1_with_attached_buffer( num_t dt, void* p, obj_t* obj );
void bli_obj_scalar_init_detached( num_t dt, obj_t* beta );
void bli_obj_free( obj_t* obj );
void bli_setsc( double zeta_r, double zeta_i, obj_t* chi );
void bli_getsc( const obj_t* chi, double* zeta_r, double* zeta_i );

/* ---- kernel/blocksize registry (frame/base/bli_gks.c:58-609,
   frame/base/bli_cntx.h:73-204, frame/base/bli_info.c) ----------------------
   The registered context describes gfx950: MR x NR is the register block one
   wave owns (the wave tile of the default large workgroup tile), MC x NC the
   workgroup tile, KC its k slab, KR the MFMA k step, and MT/NT/KT the
   thresholds of the single-call small-problem path. The gemm microkernel
   entries take BLIS's packed micro-panels (A: MR x k, column-major, ld MR;
   B: k x NR, row-major, ld NR) and m <= MR, n <= NR. */
const cntx_t*  bli_gks_query_cntx( void );
const cntx_t*  bli_gks_query_nat_cntx( void );
const cntx_t*  bli_gks_query_ind_cntx( ind_t ind );
kimpl_t        bli_gks_l3_ukr_impl_type( ukr_t ukr, ind_t method, num_t dt );
const char*    bli_gks_l3_ukr_impl_string( ukr_t ukr, ind_t method, num_t dt );
arch_t         bli_arch_query_id( void );
const char*    bli_arch_string( arch_t id );
const blksz_t* bli_cntx_get_blksz( bszid_t bs_id, const cntx_t* cntx );
dim_t          bli_cntx_get_blksz_def_dt( num_t dt, bszid_t bs_id, const cntx_t* cntx );
dim_t          bli_cntx_get_blksz_max_dt( num_t dt, bszid_t bs_id, const cntx_t* cntx );
dim_t          bli_cntx_get_bmult_dt( num_t dt, bszid_t bs_id, const cntx_t* cntx );
void_fp        bli_cntx_get_ukr_dt( num_t dt, ukr_t ukr_id, const cntx_t* cntx );
bool           bli_cntx_get_ukr_prefs_dt( num_t dt, ukr_pref_t ukr_pref_id, const cntx_t* cntx );
bool           bli_cntx_ukr_prefers_rows_dt( num_t dt, ukr_t ukr_id, const cntx_t* cntx );
bool           bli_cntx_ukr_prefers_cols_dt( num_t dt, ukr_t ukr_id, const cntx_t* cntx );
bool           bli_cntx_l3_sup_thresh_is_met( num_t dt, dim_t m, dim_t n, dim_t k,
                                              const cntx_t* cntx );

void bli_sgemm_ukernel( dim_t m, dim_t n, dim_t k, const float* alpha, const float* a,
                        const float* b, const float* beta, float* c, inc_t rs_c, inc_t cs_c,
                        auxinfo_t* data, const cntx_t* cntx );
void bli_dgemm_ukernel( dim_t m, dim_t n, dim_t k, const double* alpha, const double* a,
                        const double* b, const double* beta, double* c, inc_t rs_c, inc_t cs_c,
                        auxinfo_t* data, const cntx_t* cntx );
void bli_cgemm_ukernel( dim_t m, dim_t n, dim_t k, const scomplex* alpha, const scomplex* a,
                        const scomplex* b, const scomplex* beta, scomplex* c, inc_t rs_c,
                        inc_t cs_c, auxinfo_t* data, const cntx_t* cntx );
void bli_zgemm_ukernel( dim_t m, dim_t n, dim_t k, const dcomplex* alpha, const dcomplex* a,
                        const dcomplex* b, const dcomplex* beta, dcomplex* c, inc_t rs_c,
                        inc_t cs_c, auxinfo_t* data, const cntx_t* cntx );

gint_t      bli_info_get_int_type_size( void );
const char* bli_info_get_int_type_size_str( void );
gint_t      bli_info_get_num_fp_types( void );
gint_t      bli_info_get_max_type_size( void );
gint_t      bli_info_get_page_size( void );
gint_t      bli_info_get_simd_size( void );
gint_t      bli_info_get_simd_num_registers( void );
gint_t      bli_info_get_simd_align_size( void );
gint_t      bli_info_get_stack_buf_max_size( void );
gint_t      bli_info_get_stack_buf_align_size( void );
gint_t      bli_info_get_heap_addr_align_size( void );
gint_t      bli_info_get_enable_blas( void );
gint_t      bli_info_get_enable_cblas( void );
gint_t      bli_info_get_blas_int_type_size( void );
gint_t      bli_info_get_enable_threading( void );
gint_t      bli_info_get_enable_openmp( void );
gint_t      bli_info_get_enable_pthreads( void );
gint_t      bli_info_get_enable_sandbox( void );
const char* bli_info_get_gemm_ukr_impl_string( ind_t method, num_t dt );
const char* bli_info_get_gemmtrsm_l_ukr_impl_string( ind_t method, num_t dt );
const char* bli_info_get_gemmtrsm_u_ukr_impl_string( ind_t method, num_t dt );
const char* bli_info_get_trsm_l_ukr_impl_string( ind_t method, num_t dt );
const char* bli_info_get_trsm_u_ukr_impl_string( ind_t method, num_t dt );
const char* bli_info_get_gemm_impl_string( num_t dt );
const char* bli_info_get_gemmt_impl_string( num_t dt );
const char* bli_info_get_symm_impl_string( num_t dt );
const char* bli_info_get_syrk_impl_string( num_t dt );
const char* bli_info_get_syr2k_impl_string( num_t dt );
const char* bli_info_get_trmm_impl_string( num_t dt );
const char* bli_info_get_trmm3_impl_string( num_t dt );
const char* bli_info_get_trsm_impl_string( num_t dt );

/* ---- level-3: typed API (frame/3/bli_l3_tapi.h:44-58) ----------------- */

void bli_sgemm( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                const float* alpha, const float* a, inc_t rs_a, inc_t cs_a,
                const float* b, inc_t rs_b, inc_t cs_b,
                const float* beta, float* c, inc_t rs_c, inc_t cs_c );
void bli_dgemm( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                const double* alpha, const double* a, inc_t rs_a, inc_t cs_a,
                const double* b, inc_t rs_b, inc_t cs_b,
                const double* beta, double* c, inc_t rs_c, inc_t cs_c );
void bli_sgemm_ex( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                   const float* alpha, const float* a, inc_t rs_a, inc_t cs_a,
                   const float* b, inc_t rs_b, inc_t cs_b,
                   const float* beta, float* c, inc_t rs_c, inc_t cs_c,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_dgemm_ex( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                   const double* alpha, const double* a, inc_t rs_a, inc_t cs_a,
                   const double* b, inc_t rs_b, inc_t cs_b,
                   const double* beta, double* c, inc_t rs_c, inc_t cs_c,
                   const cntx_t* cntx, const rntm_t* rntm );

void bli_strsm( side_t side, uplo_t uploa, trans_t transa, diag_t diaga,
                dim_t m, dim_t n, const float* alpha,
                const float* a, inc_t rs_a, inc_t cs_a,
                float* b, inc_t rs_b, inc_t cs_b );
void bli_dtrsm( side_t side, uplo_t uploa, trans_t transa, diag_t diaga,
                dim_t m, dim_t n, const double* alpha,
                const double* a, inc_t rs_a, inc_t cs_a,
                double* b, inc_t rs_b, inc_t cs_b );
void bli_strsm_ex( side_t side, uplo_t uploa, trans_t transa, diag_t diaga,
                   dim_t m, dim_t n, const float* alpha,
                   const float* a, inc_t rs_a, inc_t cs_a,
                   float* b, inc_t rs_b, inc_t cs_b,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_dtrsm_ex( side_t side, uplo_t uploa, trans_t transa, diag_t diaga,
                   dim_t m, dim_t n, const double* alpha,
                   const double* a, inc_t rs_a, inc_t cs_a,
                   double* b, inc_t rs_b, inc_t cs_b,
                   const cntx_t* cntx, const rntm_t* rntm );

/* Complex GEMM, typed (frame/3/bli_l3_tapi_ex.c:43-99, scomplex/dcomplex
   instances). trans_t may conjugate: BLIS_CONJ_NO_TRANSPOSE and
   BLIS_CONJ_TRANSPOSE. Computed by the 1m method on the real GEMM tiles. */
void bli_cgemm( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                const scomplex* alpha, const scomplex* a, inc_t rs_a, inc_t cs_a,
                const scomplex* b, inc_t rs_b, inc_t cs_b,
                const scomplex* beta, scomplex* c, inc_t rs_c, inc_t cs_c );
void bli_zgemm( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                const dcomplex* alpha, const dcomplex* a, inc_t rs_a, inc_t cs_a,
                const dcomplex* b, inc_t rs_b, inc_t cs_b,
                const dcomplex* beta, dcomplex* c, inc_t rs_c, inc_t cs_c );
void bli_cgemm_ex( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                   const scomplex* alpha, const scomplex* a, inc_t rs_a, inc_t cs_a,
                   const scomplex* b, inc_t rs_b, inc_t cs_b,
                   const scomplex* beta, scomplex* c, inc_t rs_c, inc_t cs_c,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_zgemm_ex( trans_t transa, trans_t transb, dim_t m, dim_t n, dim_t k,
                   const dcomplex* alpha, const dcomplex* a, inc_t rs_a, inc_t cs_a,
                   const dcomplex* b, inc_t rs_b, inc_t cs_b,
                   const dcomplex* beta, dcomplex* c, inc_t rs_c, inc_t cs_c,
                   const cntx_t* cntx, const rntm_t* rntm );

/* ---- structured level-3: typed API (frame/3/bli_l3_tapi_ex.h:60-225) --
   gemmt: C := beta*C + alpha*op(A)*op(B) on the uploc triangle of the
   m x m C; syrk: gemmt(A, A^T); syr2k: alpha*(op(A)op(B)^T + op(B)op(A)^T);
   symm: C := beta*C + alpha*A*op(B) (left) or alpha*op(B)*A (right), A
   symmetric in its uploa triangle (conja has no effect for real types);
   trmm: B := alpha*op(A)*B (left) or alpha*B*op(A) (right), A triangular;
   trmm3: C := beta*C + alpha*op(A)*op(B) (left) or alpha*op(B)*op(A). */
#define BLIS_AMD_STRUC_PROTOS( ch, T ) \
void bli_##ch##gemmt( uplo_t uploc, trans_t transa, trans_t transb, dim_t m, dim_t k, \
                      const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                      const T* b, inc_t rs_b, inc_t cs_b, \
                      const T* beta, T* c, inc_t rs_c, inc_t cs_c ); \
void bli_##ch##gemmt_ex( uplo_t uploc, trans_t transa, trans_t transb, dim_t m, dim_t k, \
                         const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                         const T* b, inc_t rs_b, inc_t cs_b, \
                         const T* beta, T* c, inc_t rs_c, inc_t cs_c, \
                         const cntx_t* cntx, const rntm_t* rntm ); \
void bli_##ch##syrk( uplo_t uploc, trans_t transa, dim_t m, dim_t k, \
                     const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                     const T* beta, T* c, inc_t rs_c, inc_t cs_c ); \
void bli_##ch##syrk_ex( uplo_t uploc, trans_t transa, dim_t m, dim_t k, \
                        const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                        const T* beta, T* c, inc_t rs_c, inc_t cs_c, \
                        const cntx_t* cntx, const rntm_t* rntm ); \
void bli_##ch##syr2k( uplo_t uploc, trans_t transa, trans_t transb, dim_t m, dim_t k, \
                      const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                      const T* b, inc_t rs_b, inc_t cs_b, \
                      const T* beta, T* c, inc_t rs_c, inc_t cs_c ); \
void bli_##ch##syr2k_ex( uplo_t uploc, trans_t transa, trans_t transb, dim_t m, dim_t k, \
                         const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                         const T* b, inc_t rs_b, inc_t cs_b, \
                         const T* beta, T* c, inc_t rs_c, inc_t cs_c, \
                         const cntx_t* cntx, const rntm_t* rntm ); \
void bli_##ch##symm( side_t side, uplo_t uploa, conj_t conja, trans_t transb, \
                     dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                     const T* b, inc_t rs_b, inc_t cs_b, \
                     const T* beta, T* c, inc_t rs_c, inc_t cs_c ); \
void bli_##ch##symm_ex( side_t side, uplo_t uploa, conj_t conja, trans_t transb, \
                        dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                        const T* b, inc_t rs_b, inc_t cs_b, \
                        const T* beta, T* c, inc_t rs_c, inc_t cs_c, \
                        const cntx_t* cntx, const rntm_t* rntm ); \
void bli_##ch##trmm( side_t side, uplo_t uploa, trans_t transa, diag_t diaga, \
                     dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                     T* b, inc_t rs_b, inc_t cs_b ); \
void bli_##ch##trmm_ex( side_t side, uplo_t uploa, trans_t transa, diag_t diaga, \
                        dim_t m, dim_t n, const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                        T* b, inc_t rs_b, inc_t cs_b, \
                        const cntx_t* cntx, const rntm_t* rntm ); \
void bli_##ch##trmm3( side_t side, uplo_t uploa, trans_t transa, diag_t diaga, \
                      trans_t transb, dim_t m, dim_t n, \
                      const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                      const T* b, inc_t rs_b, inc_t cs_b, \
                      const T* beta, T* c, inc_t rs_c, inc_t cs_c ); \
void bli_##ch##trmm3_ex( side_t side, uplo_t uploa, trans_t transa, diag_t diaga, \
                         trans_t transb, dim_t m, dim_t n, \
                         const T* alpha, const T* a, inc_t rs_a, inc_t cs_a, \
                         const T* b, inc_t rs_b, inc_t cs_b, \
                         const T* beta, T* c, inc_t rs_c, inc_t cs_c, \
                         const cntx_t* cntx, const rntm_t* rntm );
BLIS_AMD_STRUC_PROTOS( s, float )
BLIS_AMD_STRUC_PROTOS( d, double )

/* Typed batched GEMM. No counterpart exists in the reference, whose batch
   interface is only the BLAS-style ?gemm_batch_ (frame/compat/extra/
   bla_gemm_batch.c:135-246). Same group semantics: group g holds
   group_size[g] problems that share trans/m/n/k/strides/alpha/beta; the
   pointer arrays list the problems of all groups in order. Results equal a
   loop of bli_?gemm over the problems. */
void bli_sgemm_batch( const trans_t* transa, const trans_t* transb,
                      const dim_t* m, const dim_t* n, const dim_t* k,
                      const float* alpha,
                      const float* const* a, const inc_t* rs_a, const inc_t* cs_a,
                      const float* const* b, const inc_t* rs_b, const inc_t* cs_b,
                      const float* beta,
                      float* const* c, const inc_t* rs_c, const inc_t* cs_c,
                      dim_t group_count, const dim_t* group_size );
void bli_dgemm_batch( const trans_t* transa, const trans_t* transb,
                      const dim_t* m, const dim_t* n, const dim_t* k,
                      const double* alpha,
                      const double* const* a, const inc_t* rs_a, const inc_t* cs_a,
                      const double* const* b, const inc_t* rs_b, const inc_t* cs_b,
                      const double* beta,
                      double* const* c, const inc_t* rs_c, const inc_t* cs_c,
                      dim_t group_count, const dim_t* group_size );

/* ---- level-3: object API (frame/3/bli_l3_oapi.h) --------------------- */

void bli_gemm( const obj_t* alpha, const obj_t* a, const obj_t* b,
               const obj_t* beta, const obj_t* c );
void bli_gemm_ex( const obj_t* alpha, const obj_t* a, const obj_t* b,
                  const obj_t* beta, const obj_t* c,
                  const cntx_t* cntx, const rntm_t* rntm );
void bli_trsm( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b );
void bli_trsm_ex( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                  const cntx_t* cntx, const rntm_t* rntm );

/* Structured level-3 object API (frame/3/bli_l3_oapi_ex.c:151-690). The
   stored triangle of gemmt's C, and the structure (uplo, diag, diagonal
   offset) of symm's and trmm's A, come from the objects' info fields. */
void bli_gemmt( const obj_t* alpha, const obj_t* a, const obj_t* b,
                const obj_t* beta, const obj_t* c );
void bli_gemmt_ex( const obj_t* alpha, const obj_t* a, const obj_t* b,
                   const obj_t* beta, const obj_t* c,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_syrk( const obj_t* alpha, const obj_t* a, const obj_t* beta, const obj_t* c );
void bli_syrk_ex( const obj_t* alpha, const obj_t* a, const obj_t* beta, const obj_t* c,
                  const cntx_t* cntx, const rntm_t* rntm );
void bli_syr2k( const obj_t* alpha, const obj_t* a, const obj_t* b,
                const obj_t* beta, const obj_t* c );
void bli_syr2k_ex( const obj_t* alpha, const obj_t* a, const obj_t* b,
                   const obj_t* beta, const obj_t* c,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_symm( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
               const obj_t* beta, const obj_t* c );
void bli_symm_ex( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                  const obj_t* beta, const obj_t* c,
                  const cntx_t* cntx, const rntm_t* rntm );
void bli_trmm( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b );
void bli_trmm_ex( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                  const cntx_t* cntx, const rntm_t* rntm );
void bli_trmm3( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                const obj_t* beta, const obj_t* c );
void bli_trmm3_ex( side_t side, const obj_t* alpha, const obj_t* a, const obj_t* b,
                   const obj_t* beta, const obj_t* c,
                   const cntx_t* cntx, const rntm_t* rntm );

/* ---- level-1v / level-2 typed API ------------------------------------ */

void bli_saxpyv( conj_t conjx, dim_t n, const float* alpha,
                 const float* x, inc_t incx, float* y, inc_t incy );
void bli_daxpyv( conj_t conjx, dim_t n, const double* alpha,
                 const double* x, inc_t incx, double* y, inc_t incy );
void bli_saxpyv_ex( conj_t conjx, dim_t n, const float* alpha,
                    const float* x, inc_t incx, float* y, inc_t incy,
                    const cntx_t* cntx, const rntm_t* rntm );
void bli_daxpyv_ex( conj_t conjx, dim_t n, const double* alpha,
                    const double* x, inc_t incx, double* y, inc_t incy,
                    const cntx_t* cntx, const rntm_t* rntm );

void bli_sdotv( conj_t conjx, conj_t conjy, dim_t n,
                const float* x, inc_t incx, const float* y, inc_t incy, float* rho );
void bli_ddotv( conj_t conjx, conj_t conjy, dim_t n,
                const double* x, inc_t incx, const double* y, inc_t incy, double* rho );
void bli_sdotv_ex( conj_t conjx, conj_t conjy, dim_t n,
                   const float* x, inc_t incx, const float* y, inc_t incy, float* rho,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_ddotv_ex( conj_t conjx, conj_t conjy, dim_t n,
                   const double* x, inc_t incx, const double* y, inc_t incy, double* rho,
                   const cntx_t* cntx, const rntm_t* rntm );

void bli_sgemv( trans_t transa, conj_t conjx, dim_t m, dim_t n,
                const float* alpha, const float* a, inc_t rs_a, inc_t cs_a,
                const float* x, inc_t incx, const float* beta, float* y, inc_t incy );
void bli_dgemv( trans_t transa, conj_t conjx, dim_t m, dim_t n,
                const double* alpha, const double* a, inc_t rs_a, inc_t cs_a,
                const double* x, inc_t incx, const double* beta, double* y, inc_t incy );
void bli_sgemv_ex( trans_t transa, conj_t conjx, dim_t m, dim_t n,
                   const float* alpha, const float* a, inc_t rs_a, inc_t cs_a,
                   const float* x, inc_t incx, const float* beta, float* y, inc_t incy,
                   const cntx_t* cntx, const rntm_t* rntm );
void bli_dgemv_ex( trans_t transa, conj_t conjx, dim_t m, dim_t n,
                   const double* alpha, const double* a, inc_t rs_a, inc_t cs_a,
                   const double* x, inc_t incx, const double* beta, double* y, inc_t incy,
                   const cntx_t* cntx, const rntm_t* rntm );

/* ---- level-1v / level-2 object API ----------------------------------- */

void bli_axpyv( const obj_t* alpha, const obj_t* x, const obj_t* y );
void bli_dotv( const obj_t* x, const obj_t* y, const obj_t* rho );
void bli_gemv( const obj_t* alpha, const obj_t* a, const obj_t* x,
               const obj_t* beta, const obj_t* y );

/* ---- BLAS compatibility (f77 ABI, column-major) ----------------------- */

/* As in frame/compat/bli_blas.h:46-80: the BLAS prototypes are declared
   unless the netlib BLAS test drivers are being compiled (they declare the
   routines themselves) or the includer defines BLIS_DISABLE_BLAS_DEFS. */
#define BLIS_ENABLE_BLAS_DEFS
#ifdef BLIS_VIA_BLASTEST
#undef BLIS_ENABLE_BLAS_DEFS
#endif
#ifdef BLIS_DISABLE_BLAS_DEFS
#undef BLIS_ENABLE_BLAS_DEFS
#endif
#if defined(BLIS_ENABLE_BLAS_DEFS) || defined(BLIS_IS_BUILDING_LIBRARY)

void sgemm_( const f77_char* transa, const f77_char* transb,
             const f77_int* m, const f77_int* n, const f77_int* k,
             const float* alpha, const float* a, const f77_int* lda,
             const float* b, const f77_int* ldb,
             const float* beta, float* c, const f77_int* ldc );
void dgemm_( const f77_char* transa, const f77_char* transb,
             const f77_int* m, const f77_int* n, const f77_int* k,
             const double* alpha, const double* a, const f77_int* lda,
             const double* b, const f77_int* ldb,
             const double* beta, double* c, const f77_int* ldc );
void strsm_( const f77_char* side, const f77_char* uploa,
             const f77_char* transa, const f77_char* diaga,
             const f77_int* m, const f77_int* n, const float* alpha,
             const float* a, const f77_int* lda, float* b, const f77_int* ldb );
void dtrsm_( const f77_char* side, const f77_char* uploa,
             const f77_char* transa, const f77_char* diaga,
             const f77_int* m, const f77_int* n, const double* alpha,
             const double* a, const f77_int* lda, double* b, const f77_int* ldb );
void sgemm_batch_( const f77_char* transa_array, const f77_char* transb_array,
                   const f77_int* m_array, const f77_int* n_array, const f77_int* k_array,
                   const float* alpha_array,
                   const float** a_array, const f77_int* lda_array,
                   const float** b_array, const f77_int* ldb_array,
                   const float* beta_array,
                   float** c_array, const f77_int* ldc_array,
                   const f77_int* group_count, const f77_int* group_size );
void dgemm_batch_( const f77_char* transa_array, const f77_char* transb_array,
                   const f77_int* m_array, const f77_int* n_array, const f77_int* k_array,
                   const double* alpha_array,
                   const double** a_array, const f77_int* lda_array,
                   const double** b_array, const f77_int* ldb_array,
                   const double* beta_array,
                   double** c_array, const f77_int* ldc_array,
                   const f77_int* group_count, const f77_int* group_size );
void cgemm_( const f77_char* transa, const f77_char* transb,
             const f77_int* m, const f77_int* n, const f77_int* k,
             const scomplex* alpha, const scomplex* a, const f77_int* lda,
             const scomplex* b, const f77_int* ldb,
             const scomplex* beta, scomplex* c, const f77_int* ldc );
void zgemm_( const f77_char* transa, const f77_char* transb,
             const f77_int* m, const f77_int* n, const f77_int* k,
             const dcomplex* alpha, const dcomplex* a, const f77_int* lda,
             const dcomplex* b, const f77_int* ldb,
             const dcomplex* beta, dcomplex* c, const f77_int* ldc );
void cgemm_batch_( const f77_char* transa_array, const f77_char* transb_array,
                   const f77_int* m_array, const f77_int* n_array, const f77_int* k_array,
                   const scomplex* alpha_array,
                   const scomplex** a_array, const f77_int* lda_array,
                   const scomplex** b_array, const f77_int* ldb_array,
                   const scomplex* beta_array,
                   scomplex** c_array, const f77_int* ldc_array,
                   const f77_int* group_count, const f77_int* group_size );
void zgemm_batch_( const f77_char* transa_array, const f77_char* transb_array,
                   const f77_int* m_array, const f77_int* n_array, const f77_int* k_array,
                   const dcomplex* alpha_array,
                   const dcomplex** a_array, const f77_int* lda_array,
                   const dcomplex** b_array, const f77_int* ldb_array,
                   const dcomplex* beta_array,
                   dcomplex** c_array, const f77_int* ldc_array,
                   const f77_int* group_count, const f77_int* group_size );
void sgemv_( const f77_char* transa, const f77_int* m, const f77_int* n,
             const float* alpha, const float* a, const f77_int* lda,
             const float* x, const f77_int* incx,
             const float* beta, float* y, const f77_int* incy );
void dgemv_( const f77_char* transa, const f77_int* m, const f77_int* n,
             const double* alpha, const double* a, const f77_int* lda,
             const double* x, const f77_int* incx,
             const double* beta, double* y, const f77_int* incy );
void saxpy_( const f77_int* n, const float* alpha, const float* x, const f77_int* incx,
             float* y, const f77_int* incy );
void daxpy_( const f77_int* n, const double* alpha, const double* x, const f77_int* incx,
             double* y, const f77_int* incy );
float  sdot_( const f77_int* n, const float* x, const f77_int* incx,
              const float* y, const f77_int* incy );
double ddot_( const f77_int* n, const double* x, const f77_int* incx,
              const double* y, const f77_int* incy );

/* Structured level-3 BLAS (frame/compat/bla_symm.c, bla_trmm.c, bla_syrk.c,
   bla_syr2k.c, extra/bla_gemmt.c). */
#define BLIS_AMD_STRUC_BLAS( ch, T ) \
void ch##symm_( const f77_char* side, const f77_char* uploa, \
                const f77_int* m, const f77_int* n, const T* alpha, \
                const T* a, const f77_int* lda, const T* b, const f77_int* ldb, \
                const T* beta, T* c, const f77_int* ldc ); \
void ch##trmm_( const f77_char* side, const f77_char* uploa, \
                const f77_char* transa, const f77_char* diaga, \
                const f77_int* m, const f77_int* n, const T* alpha, \
                const T* a, const f77_int* lda, T* b, const f77_int* ldb ); \
void ch##syrk_( const f77_char* uploc, const f77_char* transa, \
                const f77_int* m, const f77_int* k, const T* alpha, \
                const T* a, const f77_int* lda, const T* beta, T* c, const f77_int* ldc ); \
void ch##syr2k_( const f77_char* uploc, const f77_char* trans, \
                 const f77_int* m, const f77_int* k, const T* alpha, \
                 const T* a, const f77_int* lda, const T* b, const f77_int* ldb, \
                 const T* beta, T* c, const f77_int* ldc ); \
void ch##gemmt_( const f77_char* uploc, const f77_char* transa, const f77_char* transb, \
                 const f77_int* m, const f77_int* k, const T* alpha, \
                 const T* a, const f77_int* lda, const T* b, const f77_int* ldb, \
                 const T* beta, T* c, const f77_int* ldc );
BLIS_AMD_STRUC_BLAS( s, float )
BLIS_AMD_STRUC_BLAS( d, double )

/* xerbla_ prints the netlib message and returns (frame/compat/f2c/
   bla_xerbla.c). It is a weak symbol so that a program may supply its own,
   as the netlib test drivers do. lsame_ is the case-insensitive compare. */
int     xerbla_( const char* srname, const f77_int* info, ftnlen srname_len );
f77_int lsame_( const char* ca, const char* cb, ftnlen ca_len, ftnlen cb_len );

#endif /* BLAS prototypes */

#ifdef __cplusplus
}
#endif

#endif /* BLIS_H */
