"""ctypes wrapper around the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker / CPU baseline, never as
the thing measured. Arrays are numpy arrays; strides are in elements.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

TRANS = {"n": 0x0, "t": 0x8, "r": 0x10, "c": 0x18}   # r: conjugate only; c: conj-transpose
LOWER, UPPER = 0xC0, 0x60
LEFT, RIGHT = 0, 1
NONUNIT, UNIT = 0x0, 0x100

_lib = None
i64 = C.c_int64
P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        for ch, T in (("d", C.c_double), ("s", C.c_float)):
            getattr(L, f"oracle_{ch}gemm").argtypes = [C.c_int, C.c_int, i64, i64, i64, T, P, i64,
                                                       i64, P, i64, i64, T, P, i64, i64]
            getattr(L, f"oracle_{ch}gemm_mt").argtypes = [C.c_int, C.c_int, i64, i64, i64, T, P,
                                                          i64, i64, P, i64, i64, T, P, i64, i64,
                                                          C.c_int]
            getattr(L, f"oracle_{ch}trsm").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, i64,
                                                       i64, T, P, i64, i64, P, i64, i64]
            getattr(L, f"oracle_{ch}axpyv").argtypes = [i64, T, P, i64, P, i64]
            getattr(L, f"oracle_{ch}dotv").argtypes = [i64, P, i64, P, i64]
            getattr(L, f"oracle_{ch}dotv").restype = T
            getattr(L, f"oracle_{ch}gemv").argtypes = [C.c_int, i64, i64, T, P, i64, i64, P, i64,
                                                       T, P, i64]
            getattr(L, f"oracle_{ch}gemmt").argtypes = [C.c_int, C.c_int, C.c_int, i64, i64, T,
                                                        P, i64, i64, P, i64, i64, T, P, i64, i64]
            getattr(L, f"oracle_{ch}syrk").argtypes = [C.c_int, C.c_int, i64, i64, T, P, i64,
                                                       i64, T, P, i64, i64]
            getattr(L, f"oracle_{ch}syr2k").argtypes = [C.c_int, C.c_int, C.c_int, i64, i64, T,
                                                        P, i64, i64, P, i64, i64, T, P, i64, i64]
            getattr(L, f"oracle_{ch}symm").argtypes = [C.c_int, C.c_int, C.c_int, i64, i64, T,
                                                       P, i64, i64, P, i64, i64, T, P, i64, i64]
            getattr(L, f"oracle_{ch}trmm3").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                                        C.c_int, i64, i64, T, P, i64, i64, P,
                                                        i64, i64, T, P, i64, i64]
            getattr(L, f"oracle_{ch}trmm").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, i64,
                                                       i64, T, P, i64, i64, P, i64, i64]
        for ch, R in (("z", C.c_double), ("c", C.c_float)):
            getattr(L, f"oracle_{ch}gemm").argtypes = [C.c_int, C.c_int, i64, i64, i64, R, R, P,
                                                       i64, i64, P, i64, i64, R, R, P, i64, i64]
        L.oracle_dgemm_dd_rows.argtypes = [C.c_int, C.c_int, i64, i64, i64, C.c_double, P, i64,
                                           i64, P, i64, i64, C.c_double, P, i64, i64, i64, i64,
                                           C.c_int]
        _lib = L
    return _lib


def _ch(a):
    return "d" if a.dtype == np.float64 else "s"


def _es(a):
    return a.itemsize


def gemm(transa, transb, alpha, a, b, beta, c, nthreads=1):
    """C := beta*C + alpha*op(A)*op(B) in place on numpy arrays (any strides)."""
    ch = _ch(c)
    m, n = c.shape
    k = a.shape[0] if transa == "t" else a.shape[1]
    args = [TRANS[transa], TRANS[transb], m, n, k, alpha,
            a.ctypes.data, a.strides[0] // _es(a), a.strides[1] // _es(a),
            b.ctypes.data, b.strides[0] // _es(b), b.strides[1] // _es(b), beta,
            c.ctypes.data, c.strides[0] // _es(c), c.strides[1] // _es(c)]
    if nthreads > 1:
        getattr(lib(), f"oracle_{ch}gemm_mt")(*args, nthreads)
    else:
        getattr(lib(), f"oracle_{ch}gemm")(*args)
    return c


def trsm(side, uplo, transa, diag, alpha, a, b):
    ch = _ch(b)
    m, n = b.shape
    getattr(lib(), f"oracle_{ch}trsm")(side, uplo, TRANS[transa], diag, m, n, alpha,
                                       a.ctypes.data, a.strides[0] // _es(a),
                                       a.strides[1] // _es(a), b.ctypes.data,
                                       b.strides[0] // _es(b), b.strides[1] // _es(b))
    return b


def axpyv(alpha, x, y, incx=1, incy=1, n=None):
    ch = _ch(y)
    n = len(x) // abs(incx) if n is None else n
    getattr(lib(), f"oracle_{ch}axpyv")(n, alpha, x.ctypes.data, incx, y.ctypes.data, incy)
    return y


def dotv(x, y, incx=1, incy=1, n=None):
    ch = _ch(x)
    n = len(x) // abs(incx) if n is None else n
    return getattr(lib(), f"oracle_{ch}dotv")(n, x.ctypes.data, incx, y.ctypes.data, incy)


def gemv(transa, alpha, a, x, beta, y, incx=1, incy=1):
    ch = _ch(y)
    m, n = a.shape
    getattr(lib(), f"oracle_{ch}gemv")(TRANS[transa], m, n, alpha, a.ctypes.data,
                                       a.strides[0] // _es(a), a.strides[1] // _es(a),
                                       x.ctypes.data, incx, beta, y.ctypes.data, incy)
    return y


def gemm_dd_rows(transa, transb, alpha, a, b, beta, c, r0, r1, nthreads=1):
    """Double-double GEMM on rows [r0, r1) of C (float64 arrays)."""
    m, n = c.shape
    k = a.shape[0] if transa == "t" else a.shape[1]
    lib().oracle_dgemm_dd_rows(TRANS[transa], TRANS[transb], m, n, k, alpha,
                               a.ctypes.data, a.strides[0] // 8, a.strides[1] // 8,
                               b.ctypes.data, b.strides[0] // 8, b.strides[1] // 8, beta,
                               c.ctypes.data, c.strides[0] // 8, c.strides[1] // 8, r0, r1,
                               nthreads)
    return c


def _st(x):
    return [x.ctypes.data, x.strides[0] // _es(x), x.strides[1] // _es(x)]


def gemmt(uploc, transa, transb, alpha, a, b, beta, c):
    """C := beta*C + alpha*op(A)*op(B) on the uploc triangle of the m x m C."""
    m = c.shape[0]
    k = a.shape[0] if transa == "t" else a.shape[1]
    getattr(lib(), f"oracle_{_ch(c)}gemmt")(uploc, TRANS[transa], TRANS[transb], m, k, alpha,
                                            *_st(a), *_st(b), beta, *_st(c))
    return c


def syrk(uploc, transa, alpha, a, beta, c):
    m = c.shape[0]
    k = a.shape[0] if transa == "t" else a.shape[1]
    getattr(lib(), f"oracle_{_ch(c)}syrk")(uploc, TRANS[transa], m, k, alpha, *_st(a), beta,
                                           *_st(c))
    return c


def syr2k(uploc, transa, transb, alpha, a, b, beta, c):
    m = c.shape[0]
    k = a.shape[0] if transa == "t" else a.shape[1]
    getattr(lib(), f"oracle_{_ch(c)}syr2k")(uploc, TRANS[transa], TRANS[transb], m, k, alpha,
                                            *_st(a), *_st(b), beta, *_st(c))
    return c


def symm(side, uploa, transb, alpha, a, b, beta, c):
    m, n = c.shape
    getattr(lib(), f"oracle_{_ch(c)}symm")(side, uploa, TRANS[transb], m, n, alpha, *_st(a),
                                           *_st(b), beta, *_st(c))
    return c


def trmm3(side, uploa, transa, diaga, transb, alpha, a, b, beta, c):
    m, n = c.shape
    getattr(lib(), f"oracle_{_ch(c)}trmm3")(side, uploa, TRANS[transa], diaga, TRANS[transb],
                                            m, n, alpha, *_st(a), *_st(b), beta, *_st(c))
    return c


def trmm(side, uploa, transa, diaga, alpha, a, b):
    m, n = b.shape
    getattr(lib(), f"oracle_{_ch(b)}trmm")(side, uploa, TRANS[transa], diaga, m, n, alpha,
                                           *_st(a), *_st(b))
    return b


def zgemm(transa, transb, alpha, a, b, beta, c):
    """Complex C := beta*C + alpha*op(A)*op(B) (1m method) on numpy complex
    arrays; op is one of n, t, r (conjugate), c (conjugate transpose)."""
    ch = "z" if c.dtype == np.complex128 else "c"
    m, n = c.shape
    k = a.shape[0] if transa in ("t", "c") else a.shape[1]
    es = c.itemsize

    def st(x):
        return [x.ctypes.data, x.strides[0] // es, x.strides[1] // es]

    getattr(lib(), f"oracle_{ch}gemm")(TRANS[transa], TRANS[transb], m, n, k,
                                       complex(alpha).real, complex(alpha).imag, *st(a), *st(b),
                                       complex(beta).real, complex(beta).imag, *st(c))
    return c


def gemm_md(transa, transb, alpha, a, b, beta, c, comp_double):
    """Mixed-datatype C := beta*C + alpha*op(A)*op(B), following BLIS's rules
    (docs/MixedDatatypes.md): the product is formed in the computation
    precision (double if comp_double) and domain -- complex when A and B are
    complex, or when C and one of A, B are complex, else real; a complex
    operand in a real computation contributes its real part, a real operand
    in a complex one a zero imaginary part. A real C keeps the real part of a
    complex product; a complex C updated by a real product keeps its
    imaginary parts (updated: real parts := real(beta)*re + product). The
    products themselves are the oracle's GEMMs; the input arrays are numpy
    arrays of any of float32/float64/complex64/complex128."""
    acx, bcx, ccx = (np.iscomplexobj(x) for x in (a, b, c))
    comp_cx = (acx and bcx) or (ccx and (acx or bcx))
    rdt = np.float64 if comp_double else np.float32
    cdt = np.complex128 if comp_double else np.complex64
    m, n = c.shape
    k = a.shape[0] if transa in ("t", "c") else a.shape[1]

    def conv(x):
        if comp_cx:
            return np.asfortranarray(x.astype(cdt))
        return np.asfortranarray((x.real if np.iscomplexobj(x) else x).astype(rdt))

    A, B = conv(a), conv(b)
    p = np.zeros((m, n), dtype=cdt if comp_cx else rdt, order="F")
    if comp_cx:
        zgemm(transa, transb, alpha, A, B, 0.0, p)
    else:
        ta = "t" if transa in ("t", "c") else "n"
        tb = "t" if transb in ("t", "c") else "n"
        gemm(ta, tb, complex(alpha).real, A, B, 0.0, p)
    beta = complex(beta)
    out = c.copy()
    if ccx and comp_cx:
        out[...] = (p if beta == 0 else beta * c.astype(cdt) + p).astype(c.dtype)
    elif ccx:
        re = p if beta.real == 0 else beta.real * c.real.astype(rdt) + p
        out.real[...] = re
    else:
        pr = p.real if comp_cx else p
        out[...] = (pr if beta.real == 0 else beta.real * c.astype(rdt) + pr).astype(c.dtype)
    return out
