"""axpyv / dotv / gemv parity on the GPU through the C ABI against the oracle.

Bars: axpyv is elementwise (y + alpha*x rounds once per element, or twice
without contraction) so it must agree to 2u relative per element; dotv and
gemv reduce in a different order than the oracle, so they must agree within
the standard dot-product bound n*u*sum|x_i y_i| (per output for gemv).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import Mat, say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}


def _T(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def _ch(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,incx,incy", [(1, 1, 1), (1000, 1, 1), (1003, 1, 1), (777, 2, 3),
                                         (500, -1, 2), (65536 + 7, 1, 1)])
def test_axpyv(blis, oracle, dtype, n, incx, incy):
    say(f"axpyv {np.dtype(dtype).name} n={n} incx={incx} incy={incy}")
    u = unit_roundoff(dtype)
    x = Mat(1, n, rs=1, cs=abs(incx), dtype=dtype, seed=n)
    y = Mat(1, n, rs=1, cs=abs(incy), dtype=dtype, seed=n + 1)
    y0 = y.buf.copy()
    x.to_device(); y.to_device()
    # A negative BLIS increment walks backwards from the given origin.
    xo = x.ptr + (n - 1) * abs(incx) * x.dtype.itemsize if incx < 0 else x.ptr
    getattr(blis, f"bli_{_ch(dtype)}axpyv")(0, n, C.byref(_T(dtype)(0.75)), xo, incx, y.ptr, incy)
    got = y.dbuf.cpu().numpy()
    ref = y0.copy()
    # Same origin convention on the host: the view starts at element 0 of
    # the BLIS vector (the last stored element when incx < 0).
    xh = x.buf[(n - 1) * abs(incx):] if incx < 0 else x.buf
    oracle.axpyv(0.75, xh, ref, incx=incx, incy=incy, n=n)
    np.testing.assert_allclose(got, ref, rtol=2 * u, atol=2 * u)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,incx,incy", [(0, 1, 1), (1, 1, 1), (1000, 1, 1), (123457, 1, 1),
                                         (999, 3, 2)])
def test_dotv(blis, oracle, dtype, n, incx, incy):
    say(f"dotv {np.dtype(dtype).name} n={n}")
    u = unit_roundoff(dtype)
    x = Mat(1, max(n, 1), rs=1, cs=incx, dtype=dtype, seed=3 * n + 1)
    y = Mat(1, max(n, 1), rs=1, cs=incy, dtype=dtype, seed=3 * n + 2)
    x.to_device(); y.to_device()
    rho = _T(dtype)(123.0)
    getattr(blis, f"bli_{_ch(dtype)}dotv")(0, 0, n, x.ptr, incx, y.ptr, incy, C.byref(rho))
    ref = oracle.dotv(x.buf, y.buf, incx, incy, n=n)
    bound = n * u * float(np.sum(np.abs(x.buf[::incx][:n].astype(np.float64) *
                                        y.buf[::incy][:n].astype(np.float64)))) + 1e-300
    assert abs(rho.value - ref) <= 2 * bound, (rho.value, ref)
    if n == 0:
        assert rho.value == 0.0


def test_dotv_constant_closed_form(blis):
    n = 1 << 20
    say("ddotv constant vectors")
    x = Mat(1, n, rs=1, cs=1, fill=0.5).to_device()
    y = Mat(1, n, rs=1, cs=1, fill=3.0).to_device()
    rho = C.c_double(0)
    blis.bli_ddotv(0, 0, n, x.ptr, 1, y.ptr, 1, C.byref(rho))
    assert rho.value == 1.5 * n
    # BLAS symbol, negative increments
    ni, inc = C.c_int32(n), C.c_int32(-1)
    assert blis.ddot_(C.byref(ni), x.ptr, C.byref(inc), y.ptr, C.byref(inc)) == 1.5 * n


def _gemv_case(blis, oracle, dtype, ta, m, n, layout, alpha, beta, incx=1, incy=1):
    say(f"gemv {np.dtype(dtype).name} {ta} m={m} n={n} {layout} incx={incx} incy={incy}")
    u = unit_roundoff(dtype)
    if layout == "col":
        a = Mat(m, n, dtype=dtype, seed=m + n)
    elif layout == "row":
        a = Mat(m, n, rs=n, cs=1, dtype=dtype, seed=m + n)
    else:
        a = Mat(m, n, rs=2, cs=2 * m + 1, dtype=dtype, seed=m + n)
    nx, my = (m, n) if ta == "t" else (n, m)
    x = Mat(1, nx, rs=1, cs=incx, dtype=dtype, seed=nx + 5)
    y = Mat(1, my, rs=1, cs=incy, dtype=dtype, seed=my + 6)
    if beta == 0.0:
        y.buf[:] = np.nan
    y0 = y.buf.copy()
    for v in (a, x, y):
        v.to_device()
    T = _T(dtype)
    getattr(blis, f"bli_{_ch(dtype)}gemv")(TR[ta], 0, m, n, C.byref(T(alpha)), a.ptr, a.rs, a.cs,
                                           x.ptr, incx, C.byref(T(beta)), y.ptr, incy)
    got = y.dbuf.cpu().numpy()[::incy][:my]
    ref = y0.copy()
    oracle.gemv(ta, alpha, a.view().copy(order="F"), x.buf, beta, ref, incx=incx, incy=incy)
    ref = ref[::incy][:my]
    A = a.view().astype(np.float64)
    A = A.T if ta == "t" else A
    xs = x.buf[::incx][:nx].astype(np.float64)
    bound = nx * u * (abs(alpha) * (np.abs(A) @ np.abs(xs))) + 2 * u * np.abs(ref) + 1e-300
    assert np.isfinite(got).all()
    assert (np.abs(got - ref) <= 2 * bound).all(), np.max(np.abs(got - ref) / bound)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("layout", ["col", "row", "general"])
def test_gemv(blis, oracle, dtype, ta, layout):
    # (3000, 700): vector-aligned columns whose last row block holds whole
    # and partial row vectors (the no-transpose kernel's two vectors per
    # thread, 1024 FP64 / 2048 FP32 rows per block)
    for (m, n) in [(1, 1), (37, 53), (300, 200), (1025, 2049), (3000, 700)]:
        _gemv_case(blis, oracle, dtype, ta, m, n, layout, 0.7, 1.3)
    _gemv_case(blis, oracle, dtype, ta, 129, 77, layout, -1.0, 0.0)
    _gemv_case(blis, oracle, dtype, ta, 129, 77, layout, 2.0, 0.5, incx=2, incy=3)


def test_dgemv_f77_quirks(blis):
    """dgemv_ returns without scaling y when x is empty (bla_gemv.c:92-111),
    and alpha == 0 only scales y (bli_l2_tapi.c:77-90)."""
    say("dgemv_ quirks")
    m = 50
    a = Mat(m, 1, seed=1).to_device()
    y = Mat(1, m, rs=1, cs=1, seed=2)
    y0 = y.buf.copy()
    y.to_device()
    mi, ni, ld, inc = C.c_int32(m), C.c_int32(0), C.c_int32(m), C.c_int32(1)
    x = Mat(1, 1, fill=1.0).to_device()
    blis.dgemv_(b"N", C.byref(mi), C.byref(ni), C.byref(C.c_double(1.0)), a.ptr, C.byref(ld),
                x.ptr, C.byref(inc), C.byref(C.c_double(0.0)), y.ptr, C.byref(inc))
    np.testing.assert_array_equal(y.dbuf.cpu().numpy(), y0)
    ni = C.c_int32(1)
    blis.dgemv_(b"N", C.byref(mi), C.byref(ni), C.byref(C.c_double(0.0)), a.ptr, C.byref(ld),
                x.ptr, C.byref(inc), C.byref(C.c_double(2.0)), y.ptr, C.byref(inc))
    np.testing.assert_array_equal(y.dbuf.cpu().numpy(), 2.0 * y0)


def test_large_bandwidth_shapes(blis):
    """Config-2 shapes via size-independent properties: axpy of constants,
    dot of constants, gemv against row sums."""
    import torch
    n = 1 << 26
    say("large axpy/dot constants")
    x = torch.full((n,), 0.5, dtype=torch.float64, device="cuda")
    y = torch.full((n,), 2.0, dtype=torch.float64, device="cuda")
    blis.bli_daxpyv(0, n, C.byref(C.c_double(4.0)), x.data_ptr(), 1, y.data_ptr(), 1)
    assert bool((y == 4.0).all())
    rho = C.c_double(0)
    blis.bli_ddotv(0, 0, n, x.data_ptr(), 1, y.data_ptr(), 1, C.byref(rho))
    assert rho.value == 2.0 * n
    del x, y
    m = 8192
    say("large gemv row sums")
    a = torch.rand((m, m), dtype=torch.float64, device="cuda").t()   # column-major view
    ones = torch.ones(m, dtype=torch.float64, device="cuda")
    out = torch.zeros(m, dtype=torch.float64, device="cuda")
    blis.bli_dgemv(0, 0, m, m, C.byref(C.c_double(1.0)), a.data_ptr(), 1, m, ones.data_ptr(), 1,
                   C.byref(C.c_double(0.0)), out.data_ptr(), 1)
    ref = a.sum(dim=1)
    assert float((out - ref).abs().max()) <= m * 2.0 ** -53 * m
