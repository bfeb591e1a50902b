/*
 * blas1_extra.c — the level-1 BLAS routines that the reference's netlib
 * driver blastest/src/{d,s}blat1.c exercises besides ?axpy_ and ?dot_ (TEST
 * INFRASTRUCTURE ONLY; they are outside the library's hot path). With them
 * the driver runs whole, both against the oracle (?axpy_/?dot_ from
 * blas_f77.c) and against libblis_amd.so (?axpy_/?dot_ from the library);
 * this file defines neither of those two.
 *
 * Semantics restated from the BLAS definitions as the reference implements
 * them (frame/compat/f2c/bla_rotg.c, bla_rotmg.c, bla_rot.c, bla_rotm.c;
 * frame/compat/bla_{copy,swap,nrm2,asum,scal,amax}.c, bla_dot.c for the
 * ?sdot variants). Increments follow the BLAS convention: a negative
 * increment walks the vector from its last element
 * (frame/include/bli_blas_macro_defs.h:50-74); routines that take one
 * vector return at once for n < 1 or incx < 1.
 */
#include <math.h>
#include <stdint.h>

typedef int32_t f77_int;

/* First element touched for n elements at increment inc (BLAS convention). */
static int64_t first(f77_int n, f77_int inc) { return inc < 0 ? (int64_t)(1 - n) * inc : 0; }

#define DEFINE_L1(T, ch, AMAX)                                                          \
int ch##rotg_(T* a, T* b, T* c, T* s)                                                   \
{                                                                                       \
    /* r = sign(roe) * hypot(a, b) with roe the larger of a, b in magnitude,            \
       z reconstructs the rotation from one number. */                                  \
    const T roe = (fabs((double)*a) > fabs((double)*b)) ? *a : *b;                      \
    const T scale = (T)(fabs((double)*a) + fabs((double)*b));                           \
    T r, z;                                                                             \
    if (scale == (T)0) {                                                                \
        *c = (T)1; *s = (T)0; r = (T)0; z = (T)0;                                       \
    } else {                                                                            \
        const T p = *a / scale, q = *b / scale;                                         \
        r = scale * (T)sqrt((double)(p * p + q * q));                                   \
        if (roe < (T)0) r = -r;                                                         \
        *c = *a / r;                                                                    \
        *s = *b / r;                                                                    \
        z = (T)1;                                                                       \
        if (fabs((double)*a) > fabs((double)*b)) z = *s;                                \
        if (fabs((double)*b) >= fabs((double)*a) && *c != (T)0) z = (T)1 / *c;          \
    }                                                                                   \
    *a = r;                                                                             \
    *b = z;                                                                             \
    return 0;                                                                           \
}                                                                                       \
                                                                                        \
int ch##rot_(const f77_int* n, T* x, const f77_int* incx, T* y, const f77_int* incy,    \
             const T* c, const T* s)                                                    \
{                                                                                       \
    if (*n <= 0) return 0;                                                              \
    int64_t ix = first(*n, *incx), iy = first(*n, *incy);                               \
    for (f77_int i = 0; i < *n; ++i, ix += *incx, iy += *incy) {                        \
        const T t = *c * x[ix] + *s * y[iy];                                            \
        y[iy] = *c * y[iy] - *s * x[ix];                                                \
        x[ix] = t;                                                                      \
    }                                                                                   \
    return 0;                                                                           \
}                                                                                       \
                                                                                        \
int ch##copy_(const f77_int* n, const T* x, const f77_int* incx, T* y,                  \
              const f77_int* incy)                                                      \
{                                                                                       \
    if (*n <= 0) return 0;                                                              \
    int64_t ix = first(*n, *incx), iy = first(*n, *incy);                               \
    for (f77_int i = 0; i < *n; ++i, ix += *incx, iy += *incy) y[iy] = x[ix];           \
    return 0;                                                                           \
}                                                                                       \
                                                                                        \
int ch##swap_(const f77_int* n, T* x, const f77_int* incx, T* y, const f77_int* incy)   \
{                                                                                       \
    if (*n <= 0) return 0;                                                              \
    int64_t ix = first(*n, *incx), iy = first(*n, *incy);                               \
    for (f77_int i = 0; i < *n; ++i, ix += *incx, iy += *incy) {                        \
        const T t = x[ix]; x[ix] = y[iy]; y[iy] = t;                                    \
    }                                                                                   \
    return 0;                                                                           \
}                                                                                       \
                                                                                        \
T ch##nrm2_(const f77_int* n, const T* x, const f77_int* incx)                          \
{                                                                                       \
    /* Scaled sum of squares: ssq * scale^2 = sum x_i^2 with scale the largest          \
       magnitude seen so far, which avoids overflow and underflow. */                   \
    if (*n < 1 || *incx < 1) return (T)0;                                               \
    double scale = 0.0, ssq = 1.0;                                                      \
    for (f77_int i = 0; i < *n; ++i) {                                                  \
        const double v = fabs((double)x[(int64_t)i * *incx]);                           \
        if (v == 0.0) continue;                                                         \
        if (scale < v) { ssq = 1.0 + ssq * (scale / v) * (scale / v); scale = v; }      \
        else ssq += (v / scale) * (v / scale);                                          \
    }                                                                                   \
    return (T)(scale * sqrt(ssq));                                                      \
}                                                                                       \
                                                                                        \
T ch##asum_(const f77_int* n, const T* x, const f77_int* incx)                          \
{                                                                                       \
    if (*n < 1 || *incx < 1) return (T)0;                                               \
    T s = (T)0;                                                                         \
    for (f77_int i = 0; i < *n; ++i) s += (T)fabs((double)x[(int64_t)i * *incx]);       \
    return s;                                                                           \
}                                                                                       \
                                                                                        \
int ch##scal_(const f77_int* n, const T* a, T* x, const f77_int* incx)                  \
{                                                                                       \
    if (*n < 1 || *incx < 1) return 0;                                                  \
    for (f77_int i = 0; i < *n; ++i) x[(int64_t)i * *incx] *= *a;                       \
    return 0;                                                                           \
}                                                                                       \
                                                                                        \
f77_int AMAX(const f77_int* n, const T* x, const f77_int* incx)                         \
{                                                                                       \
    /* 1-based index of the first element of largest magnitude. */                      \
    if (*n < 1 || *incx < 1) return 0;                                                  \
    f77_int best = 1;                                                                   \
    double bv = fabs((double)x[0]);                                                     \
    for (f77_int i = 1; i < *n; ++i) {                                                  \
        const double v = fabs((double)x[(int64_t)i * *incx]);                           \
        if (v > bv) { bv = v; best = i + 1; }                                           \
    }                                                                                   \
    return best;                                                                        \
}                                                                                       \
                                                                                        \
int ch##rotmg_(T* d1, T* d2, T* x1, const T* y1, T* param)                              \
{                                                                                       \
    /* Modified Givens: H zeroes the second component of                               \
       (sqrt(d1) x1, sqrt(d2) y1). param[0] = flag selects H's form:                    \
       -1: full (h11 h12 / h21 h22); 0: (1 h12 / h21 1); 1: (h11 1 / -1 h22);           \
       -2: identity. d1, d2 are kept within [1/gam^2, gam^2] by powers of gam. */      \
    const T gam = (T)4096, gamsq = (T)16777216, rgamsq = (T)5.9604645e-8;               \
    T flag, h11 = 0, h12 = 0, h21 = 0, h22 = 0;                                         \
    int zero_all = 0;                                                                   \
    if (*d1 < (T)0) {                                                                   \
        zero_all = 1;                                                                   \
    } else {                                                                            \
        const T p2 = *d2 * *y1;                                                         \
        if (p2 == (T)0) { param[0] = (T)-2; return 0; }                                 \
        const T p1 = *d1 * *x1, q2 = p2 * *y1, q1 = p1 * *x1;                           \
        if (fabs((double)q1) > fabs((double)q2)) {                                      \
            h21 = -*y1 / *x1;                                                           \
            h12 = p2 / p1;                                                              \
            const T u = (T)1 - h12 * h21;                                               \
            if (u <= (T)0) zero_all = 1;                                                \
            else { flag = (T)0; *d1 /= u; *d2 /= u; *x1 *= u; }                         \
        } else if (q2 < (T)0) {                                                         \
            zero_all = 1;                                                               \
        } else {                                                                        \
            flag = (T)1;                                                                \
            h11 = p1 / p2;                                                              \
            h22 = *x1 / *y1;                                                            \
            const T u = (T)1 + h11 * h22;                                               \
            const T t = *d2 / u;                                                        \
            *d2 = *d1 / u;                                                              \
            *d1 = t;                                                                    \
            *x1 = *y1 * u;                                                              \
        }                                                                               \
    }                                                                                   \
    if (zero_all) {                                                                     \
        flag = (T)-1;                                                                   \
        h11 = h12 = h21 = h22 = (T)0;                                                   \
        *d1 = *d2 = *x1 = (T)0;                                                         \
    } else {                                                                            \
        /* Make every entry of H explicit before rescaling a row of it. */              \
        _Bool full = 0;                                                                 \
        if (*d1 != (T)0) {                                                              \
            while (*d1 <= rgamsq || *d1 >= gamsq) {                                     \
                if (!full) {                                                            \
                    if (flag == (T)0) { h11 = (T)1; h22 = (T)1; }                       \
                    else if (flag == (T)1) { h21 = (T)-1; h12 = (T)1; }                 \
                    flag = (T)-1; full = 1;                                             \
                }                                                                       \
                if (*d1 <= rgamsq) {                                                    \
                    *d1 *= gam * gam; *x1 /= gam; h11 /= gam; h12 /= gam;               \
                } else {                                                                \
                    *d1 /= gam * gam; *x1 *= gam; h11 *= gam; h12 *= gam;               \
                }                                                                       \
            }                                                                           \
        }                                                                               \
        if (*d2 != (T)0) {                                                              \
            while (fabs((double)*d2) <= rgamsq || fabs((double)*d2) >= gamsq) {         \
                if (!full) {                                                            \
                    if (flag == (T)0) { h11 = (T)1; h22 = (T)1; }                       \
                    else if (flag == (T)1) { h21 = (T)-1; h12 = (T)1; }                 \
                    flag = (T)-1; full = 1;                                             \
                }                                                                       \
                if (fabs((double)*d2) <= rgamsq) {                                      \
                    *d2 *= gam * gam; h21 /= gam; h22 /= gam;                           \
                } else {                                                                \
                    *d2 /= gam * gam; h21 *= gam; h22 *= gam;                           \
                }                                                                       \
            }                                                                           \
        }                                                                               \
    }                                                                                   \
    if (flag < (T)0) { param[1] = h11; param[2] = h21; param[3] = h12; param[4] = h22; } \
    else if (flag == (T)0) { param[2] = h21; param[3] = h12; }                          \
    else { param[1] = h11; param[4] = h22; }                                            \
    param[0] = flag;                                                                    \
    return 0;                                                                           \
}                                                                                       \
                                                                                        \
int ch##rotm_(const f77_int* n, T* x, const f77_int* incx, T* y, const f77_int* incy,   \
              const T* param)                                                           \
{                                                                                       \
    const T flag = param[0];                                                            \
    if (*n <= 0 || flag == (T)-2) return 0;                                             \
    T h11, h12, h21, h22;                                                               \
    if (flag < (T)0) { h11 = param[1]; h21 = param[2]; h12 = param[3]; h22 = param[4]; } \
    else if (flag == (T)0) { h11 = (T)1; h21 = param[2]; h12 = param[3]; h22 = (T)1; }  \
    else { h11 = param[1]; h21 = (T)-1; h12 = (T)1; h22 = param[4]; }                   \
    int64_t ix = first(*n, *incx), iy = first(*n, *incy);                               \
    for (f77_int i = 0; i < *n; ++i, ix += *incx, iy += *incy) {                        \
        const T w = x[ix], z = y[iy];                                                   \
        if (flag == (T)0) { x[ix] = w + z * h12; y[iy] = w * h21 + z; }                 \
        else if (flag > (T)0) { x[ix] = w * h11 + z; y[iy] = -w + h22 * z; }            \
        else { x[ix] = w * h11 + z * h12; y[iy] = w * h21 + z * h22; }                  \
    }                                                                                   \
    return 0;                                                                           \
}

DEFINE_L1(double, d, idamax_)
DEFINE_L1(float, s, isamax_)

/* Dot products of single-precision vectors accumulated in double precision;
   sdsdot adds sb and returns single precision. */
static double dsdot_acc(f77_int n, const float* x, f77_int incx, const float* y, f77_int incy)
{
    double acc = 0.0;
    if (n <= 0) return 0.0;
    int64_t ix = first(n, incx), iy = first(n, incy);
    for (f77_int i = 0; i < n; ++i, ix += incx, iy += incy) acc += (double)x[ix] * (double)y[iy];
    return acc;
}

double dsdot_(const f77_int* n, const float* x, const f77_int* incx, const float* y,
              const f77_int* incy)
{
    return dsdot_acc(*n, x, *incx, y, *incy);
}

float sdsdot_(const f77_int* n, const float* sb, const float* x, const f77_int* incx,
              const float* y, const f77_int* incy)
{
    return (float)((double)*sb + dsdot_acc(*n, x, *incx, y, *incy));
}
