// api_common.hpp — argument checking, scalar and obj_t helpers and BLAS
// character decoding shared by the library's front ends (api.cpp, the
// structured and complex level-3 interfaces, the CBLAS layer).
//
// Reference behaviour kept here: error checking aborts with a message in
// BLIS's FULL checking mode (frame/base/bli_check.c:40-58, bli_error.c:
// 144-151); BLAS argument errors go through xerbla_ with the 6-character
// routine name (frame/compat/check/bla_gemm_check.h:56-79).
#pragma once

#include "internal.hpp"

#include <cctype>
#include <cstring>

namespace blis_amd {
namespace api {

// Layout of the buffer behind BLIS's global scalar constants
// (frame/include/bli_type_defs.h constdata_t): one value per datatype.
struct ConstData {
    float    s;
    double   d;
    scomplex c;
    dcomplex z;
    gint_t   i;
};

// Error-checking level (bli_error_checking_level_set).
errlev_t& errlev();

inline bool check_on() { return errlev() == BLIS_FULL_ERROR_CHECKING; }

inline void require(bool ok, err_t code, const char* where) {
    if (!ok && check_on()) abort_with(code, where);
}

inline bool valid_trans(trans_t t) {
    return t == BLIS_NO_TRANSPOSE || t == BLIS_TRANSPOSE ||
           t == BLIS_CONJ_NO_TRANSPOSE || t == BLIS_CONJ_TRANSPOSE;
}
inline bool does_trans(trans_t t) { return ( (unsigned)t & BLIS_TRANS_BIT ) != 0; }

template <typename T> constexpr Dt dt_for();
template <> constexpr Dt dt_for<double>() { return Dt::F64; }
template <> constexpr Dt dt_for<float>() { return Dt::F32; }

// Read a scalar held by a host or device pointer.
template <typename T>
T read_scalar(const T* p) {
    if (is_device_pointer(p)) {
        T v;
        HIP_CHECK(hipMemcpy(&v, p, sizeof(T), hipMemcpyDeviceToHost));
        return v;
    }
    return *p;
}

inline bool obj_is_real_fp(const obj_t* o) {
    const num_t dt = bli_obj_dt(o);
    return dt == BLIS_FLOAT || dt == BLIS_DOUBLE;
}

// Value of a 1x1 scalar object in double precision.
inline double obj_scalar(const obj_t* o) {
    require(o != nullptr, BLIS_NULL_POINTER, "scalar object");
    const num_t dt = bli_obj_dt(o);
    const char* base = (const char*)o->buffer;
    require(base != nullptr, BLIS_EXPECTED_NONNULL_OBJECT_BUFFER, "scalar object buffer");
    if (dt == BLIS_CONSTANT) return ((const ConstData*)base)->d;
    const char* p = base + (o->off[0] * o->rs + o->off[1] * o->cs) * (inc_t)o->elem_size;
    if (dt == BLIS_DOUBLE) return read_scalar((const double*)p);
    if (dt == BLIS_FLOAT) return (double)read_scalar((const float*)p);
    abort_with(BLIS_EXPECTED_REAL_DATATYPE, "scalar object datatype");
    return 0.0;
}

inline void* obj_origin(const obj_t* o) {
    return (char*)o->buffer + (o->off[0] * o->rs + o->off[1] * o->cs) * (inc_t)o->elem_size;
}

inline dim_t obj_vec_dim(const obj_t* o) { return o->dim[0] == 1 ? o->dim[1] : o->dim[0]; }
inline inc_t obj_vec_inc(const obj_t* o) { return o->dim[0] == 1 ? o->cs : o->rs; }

inline void check_matrix_obj(const obj_t* o, const char* where) {
    require(o != nullptr, BLIS_NULL_POINTER, where);
    require(obj_is_real_fp(o), BLIS_EXPECTED_REAL_DATATYPE, where);
    require(o->dim[0] >= 0 && o->dim[1] >= 0, BLIS_NEGATIVE_DIMENSION, where);
    require(o->buffer != nullptr || o->dim[0] == 0 || o->dim[1] == 0,
            BLIS_EXPECTED_NONNULL_OBJECT_BUFFER, where);
}

inline char up(char ch) { return (char)std::toupper((unsigned char)ch); }

inline bool map_trans(char ch, trans_t* t) {
    switch (up(ch)) {
    case 'N': *t = BLIS_NO_TRANSPOSE; return true;
    case 'T': *t = BLIS_TRANSPOSE; return true;
    case 'C': *t = BLIS_CONJ_TRANSPOSE; return true;
    default: return false;
    }
}

inline void blas_error(const char* dt, const char* op, f77_int info) {
    char name[16];
    std::snprintf(name, sizeof(name), "%s%-5s", dt, op);
    for (char* q = name; *q; ++q) *q = up(*q);
    xerbla_(name, &info, (ftnlen)6);
}

// Object GEMM whose operands are complex, of differing datatypes, or whose
// computation precision differs from C's storage precision
// (api_complex.cpp).
void gemm_obj_md(const obj_t* alpha, const obj_t* a, const obj_t* b, const obj_t* beta,
                 const obj_t* c);

} // namespace api
} // namespace blis_amd
