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

TRANS = {"n": 0x0, "t": 0x8}
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
