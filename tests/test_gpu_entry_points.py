"""Entry points with logic of their own, through the C ABI, against the
oracle (or bitwise against the typed entry point they forward to):

  * the object APIs bli_trsm / bli_trsm_ex (every side, uplo, trans and
    diag, read from the obj_t info bits: frame/3/bli_l3_oapi_ex.c:692-801),
    bli_axpyv, bli_dotv and bli_gemv (vector objects of either orientation);
  * the BLAS symbols daxpy_/saxpy_/ddot_/sdot_ with negative increments (the
    vector is walked from its last element, frame/include/
    bli_blas_macro_defs.h:50-74) and sgemm_batch_ with several groups;
  * every _ex variant (cntx and rntm passed as NULL), which must give the
    same bits as the entry point without _ex.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import blis_amd
from tests.conftest import Mat, normwise_gemm_err, say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}
LOWER, UPPER = 0xC0, 0x60
UNIT = 0x100
TRIANGULAR = 0x3 << 20
DT = {np.dtype(np.float64): blis_amd.BLIS_DOUBLE, np.dtype(np.float32): blis_amd.BLIS_FLOAT}


def _T(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def _obj(blis, dtype, mat: Mat):
    o = blis_amd.obj_t()
    blis.bli_obj_create_with_attached_buffer(DT[np.dtype(dtype)], mat.m, mat.n, mat.ptr,
                                             mat.rs, mat.cs, C.byref(o))
    return o


def _scalar(blis, dtype, value):
    v = _T(dtype)(value)
    o = blis_amd.obj_t()
    blis.bli_obj_create_1x1_with_attached_buffer(DT[np.dtype(dtype)], C.byref(v), C.byref(o))
    return o, v


def _tri(mn, uplo, unit, dtype, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (mn, mn)) / mn
    a[np.diag_indices(mn)] = 1.0 + rng.uniform(0, 1, mn)
    if uplo == LOWER:
        a[np.triu_indices(mn, 1)] = np.nan
    else:
        a[np.tril_indices(mn, -1)] = np.nan
    if unit:
        a[np.diag_indices(mn)] = np.nan
    m = Mat(mn, mn, dtype=dtype)
    m.set(a.astype(dtype))
    return m


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("diag", [0, UNIT])
@pytest.mark.parametrize("ex", [False, True])
def test_trsm_object_api(blis, oracle, dtype, side, uplo, ta, diag, ex):
    m, n = 37, 29
    mn = m if side == 0 else n
    say(f"bli_trsm{'_ex' if ex else ''} {np.dtype(dtype).name} side={side} uplo={uplo:#x} "
        f"{ta} diag={diag:#x}")
    a = _tri(mn, uplo, diag == UNIT, dtype, seed=mn + side)
    b = Mat(m, n, dtype=dtype, seed=5)
    b0 = b.view().copy()
    a.to_device()
    b.to_device()
    ao, bo = _obj(blis, dtype, a), _obj(blis, dtype, b)
    ao.info = (ao.info & ~(0x18 | 0xE0 | UNIT | (0x7 << 20))) | TR[ta] | uplo | diag | TRIANGULAR
    alo, keep = _scalar(blis, dtype, 1.5)
    if ex:
        blis.bli_trsm_ex.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
        blis.bli_trsm_ex(side, C.byref(alo), C.byref(ao), C.byref(bo), None, None)
    else:
        blis.bli_trsm(side, C.byref(alo), C.byref(ao), C.byref(bo))
    got = b.device_view()
    ref = b0.copy(order="F")
    oracle.trsm(side, uplo, ta, diag, 1.5, a.view().copy(order="F"), ref)
    err = np.linalg.norm((got - ref).astype(np.float64)) / np.linalg.norm(ref.astype(np.float64))
    assert err <= 8 * max(m, n) * unit_roundoff(dtype), err


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_level12_object_api(blis, oracle, dtype):
    n = 1000
    say(f"bli_axpyv / bli_dotv / bli_gemv objects {np.dtype(dtype).name}")
    T = _T(dtype)
    # axpyv: x a column vector (n x 1, stride 1), y a row vector (1 x n, cs 3).
    x = Mat(n, 1, dtype=dtype, seed=1)
    y = Mat(1, n, rs=1, cs=3, dtype=dtype, seed=2)
    y0 = y.view().ravel().copy()
    x.to_device()
    y.to_device()
    alo, k1 = _scalar(blis, dtype, 0.5)
    blis.bli_axpyv.argtypes = [C.c_void_p] * 3
    blis.bli_axpyv(C.byref(alo), C.byref(_obj(blis, dtype, x)), C.byref(_obj(blis, dtype, y)))
    ref = y0.copy()
    oracle.axpyv(0.5, x.view().ravel().copy(), ref)
    np.testing.assert_array_equal(y.device_view().ravel(), ref)
    # dotv into a host scalar object.
    rho = T(0)
    rho_o = blis_amd.obj_t()
    blis.bli_obj_create_1x1_with_attached_buffer(DT[np.dtype(dtype)], C.byref(rho),
                                                 C.byref(rho_o))
    blis.bli_dotv.argtypes = [C.c_void_p] * 3
    blis.bli_dotv(C.byref(_obj(blis, dtype, x)), C.byref(_obj(blis, dtype, y)), C.byref(rho_o))
    want = oracle.dotv(x.view().ravel().copy(), y.device_view().ravel().copy())
    xs, ys = x.view().ravel().astype(np.float64), y.device_view().ravel().astype(np.float64)
    assert abs(rho.value - want) <= 4 * n * unit_roundoff(dtype) * np.sum(np.abs(xs * ys))
    # gemv with op(A) = A^T: y (length 70) := beta*y + alpha*A^T x.
    a = Mat(n, 70, dtype=dtype, seed=3)
    yv = Mat(70, 1, dtype=dtype, seed=4)
    yv0 = yv.view().ravel().copy()
    a.to_device()
    yv.to_device()
    ao = _obj(blis, dtype, a)
    ao.info = (ao.info & ~0x18) | TR["t"]
    beo, k2 = _scalar(blis, dtype, -0.25)
    blis.bli_gemv.argtypes = [C.c_void_p] * 5
    blis.bli_gemv(C.byref(alo), C.byref(ao), C.byref(_obj(blis, dtype, x)), C.byref(beo),
                  C.byref(_obj(blis, dtype, yv)))
    ref = yv0.copy()
    oracle.gemv("t", 0.5, a.view().copy(order="F"), x.view().ravel().copy(), -0.25, ref)
    err = np.linalg.norm((yv.device_view().ravel() - ref).astype(np.float64)) / (
        0.5 * np.linalg.norm(a.view().astype(np.float64)) * np.linalg.norm(xs)
        + 0.25 * np.linalg.norm(yv0.astype(np.float64)))
    assert err <= 4 * n * unit_roundoff(dtype), err


def _blas_walk(v, n, inc):
    """Logical elements of a BLAS vector argument: negative increments start
    at the last element."""
    step = abs(inc)
    idx = np.arange(n) * step
    return idx[::-1] if inc < 0 else idx


@pytest.mark.parametrize("ch", ["d", "s"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (-1, 2), (3, -2), (-2, -3)])
def test_blas_axpy_dot_negative_increments(blis, ch, incx, incy):
    n = 777
    dtype = np.float64 if ch == "d" else np.float32
    T = _T(dtype)
    say(f"{ch}axpy_ / {ch}dot_ incx={incx} incy={incy}")
    xm = Mat(abs(incx) * (n - 1) + 1, 1, dtype=dtype, seed=7).to_device()
    ym = Mat(abs(incy) * (n - 1) + 1, 1, dtype=dtype, seed=8).to_device()
    xh, yh = xm.view().ravel().copy(), ym.view().ravel().copy()
    ni, ix, iy = C.c_int32(n), C.c_int32(incx), C.c_int32(incy)
    # dot first (no side effects), then axpy with alpha = 0.5 (exact products).
    got = getattr(blis, f"{ch}dot_")(C.byref(ni), xm.ptr, C.byref(ix), ym.ptr, C.byref(iy))
    px, py = _blas_walk(xh, n, incx), _blas_walk(yh, n, incy)
    prods = xh[px].astype(np.float64) * yh[py].astype(np.float64)
    assert abs(got - prods.sum()) <= 4 * n * unit_roundoff(dtype) * np.abs(prods).sum()
    getattr(blis, f"{ch}axpy_")(C.byref(ni), C.byref(T(0.5)), xm.ptr, C.byref(ix), ym.ptr,
                                C.byref(iy))
    want = yh.copy()
    want[py] = (yh[py] + dtype(0.5) * xh[px]).astype(dtype)
    np.testing.assert_array_equal(ym.device_view().ravel(), want)


def test_sgemm_batch_f77_groups(blis, oracle):
    """sgemm_batch_ with three groups of different shapes and transposes."""
    say("sgemm_batch_ three groups")
    dtype = np.float32
    groups = [(3, "N", "T", 33, 20, 17, 1.0, 0.5), (2, "T", "N", 64, 65, 66, -0.5, 0.0),
              (4, "T", "T", 5, 90, 40, 2.0, -1.0)]
    probs = []
    seed = 0
    for (size, ta, tb, m, n, k, al, be) in groups:
        for _ in range(size):
            a = Mat(*((k, m) if ta == "T" else (m, k)), dtype=dtype, seed=seed).to_device()
            b = Mat(*((n, k) if tb == "T" else (k, n)), dtype=dtype, seed=seed + 1).to_device()
            c = Mat(m, n, dtype=dtype, seed=seed + 2)
            c0 = c.view().copy()
            c.to_device()
            probs.append((a, b, c, c0))
            seed += 3
    G = len(groups)
    Fi = C.c_int32 * G
    chars = lambda i: (C.c_char * G)(*[g[i].encode() for g in groups])
    ta, tb = chars(1), chars(2)
    m, n, k = (Fi(*[g[i] for g in groups]) for i in (3, 4, 5))
    alpha = (C.c_float * G)(*[g[6] for g in groups])
    beta = (C.c_float * G)(*[g[7] for g in groups])
    lda = Fi(*[(g[5] if g[1] == "T" else g[3]) for g in groups])
    ldb = Fi(*[(g[4] if g[2] == "T" else g[5]) for g in groups])
    ldc = Fi(*[g[3] for g in groups])
    P = C.c_void_p * len(probs)
    ap, bp, cp = (P(*[p[i].ptr for p in probs]) for i in range(3))
    gc = C.c_int32(G)
    gs = Fi(*[g[0] for g in groups])
    blis.sgemm_batch_(C.cast(ta, C.c_void_p), C.cast(tb, C.c_void_p), C.cast(m, C.c_void_p),
                      C.cast(n, C.c_void_p), C.cast(k, C.c_void_p), C.cast(alpha, C.c_void_p),
                      C.cast(ap, C.c_void_p), C.cast(lda, C.c_void_p), C.cast(bp, C.c_void_p),
                      C.cast(ldb, C.c_void_p), C.cast(beta, C.c_void_p),
                      C.cast(cp, C.c_void_p), C.cast(ldc, C.c_void_p),
                      C.cast(C.pointer(gc), C.c_void_p), C.cast(gs, C.c_void_p))
    idx = 0
    for (size, gta, gtb, gm, gn, gk, al, be) in groups:
        for _ in range(size):
            a, b, c, c0 = probs[idx]
            ref = c0.copy(order="F")
            oracle.gemm(gta.lower(), gtb.lower(), al, a.view(), b.view(), be, ref)
            err = normwise_gemm_err(c.device_view(), ref, a.view(), b.view(), c0, al, be)
            assert err <= 4 * gk * unit_roundoff(dtype), (idx, err)
            idx += 1


@pytest.mark.parametrize("ch", ["d", "s"])
def test_ex_variants_match(blis, ch):
    """Every typed _ex variant against its plain twin, bitwise."""
    dtype = np.float64 if ch == "d" else np.float32
    T = _T(dtype)
    PT = C.POINTER(T)
    P = C.c_void_p
    I = C.c_int64
    say(f"bli_{ch}*_ex variants")
    # gemm
    m, n, k = 70, 45, 33
    a, b = Mat(m, k, dtype=dtype, seed=1).to_device(), Mat(k, n, dtype=dtype, seed=2).to_device()
    c1, c2 = Mat(m, n, dtype=dtype, seed=3).to_device(), Mat(m, n, dtype=dtype, seed=3).to_device()
    fex = getattr(blis, f"bli_{ch}gemm_ex")
    fex.argtypes = [C.c_int, C.c_int, I, I, I, PT, P, I, I, P, I, I, PT, P, I, I, P, P]
    al, be = T(1.25), T(-0.5)
    getattr(blis, f"bli_{ch}gemm")(0, 0, m, n, k, C.byref(al), a.ptr, 1, m, b.ptr, 1, k,
                                   C.byref(be), c1.ptr, 1, m)
    fex(0, 0, m, n, k, C.byref(al), a.ptr, 1, m, b.ptr, 1, k, C.byref(be), c2.ptr, 1, m,
        None, None)
    np.testing.assert_array_equal(c1.device_view(), c2.device_view())
    # trsm
    t = _tri(40, LOWER, False, dtype, seed=9).to_device()
    b1, b2 = Mat(40, 12, dtype=dtype, seed=4).to_device(), Mat(40, 12, dtype=dtype, seed=4).to_device()
    fex = getattr(blis, f"bli_{ch}trsm_ex")
    fex.argtypes = [C.c_int] * 4 + [I, I, PT, P, I, I, P, I, I, P, P]
    getattr(blis, f"bli_{ch}trsm")(0, LOWER, 0, 0, 40, 12, C.byref(al), t.ptr, 1, 40, b1.ptr, 1, 40)
    fex(0, LOWER, 0, 0, 40, 12, C.byref(al), t.ptr, 1, 40, b2.ptr, 1, 40, None, None)
    np.testing.assert_array_equal(b1.device_view(), b2.device_view())
    # axpyv, dotv, gemv
    nv = 501
    x = Mat(nv, 1, dtype=dtype, seed=5).to_device()
    y1, y2 = Mat(nv, 1, dtype=dtype, seed=6).to_device(), Mat(nv, 1, dtype=dtype, seed=6).to_device()
    fex = getattr(blis, f"bli_{ch}axpyv_ex")
    fex.argtypes = [C.c_int, I, PT, P, I, P, I, P, P]
    getattr(blis, f"bli_{ch}axpyv")(0, nv, C.byref(al), x.ptr, 1, y1.ptr, 1)
    fex(0, nv, C.byref(al), x.ptr, 1, y2.ptr, 1, None, None)
    np.testing.assert_array_equal(y1.device_view(), y2.device_view())
    r1, r2 = T(0), T(0)
    fex = getattr(blis, f"bli_{ch}dotv_ex")
    fex.argtypes = [C.c_int, C.c_int, I, P, I, P, I, P, P, P]
    getattr(blis, f"bli_{ch}dotv")(0, 0, nv, x.ptr, 1, y1.ptr, 1, C.addressof(r1))
    fex(0, 0, nv, x.ptr, 1, y1.ptr, 1, C.addressof(r2), None, None)
    assert r1.value == r2.value
    g = Mat(nv, 20, dtype=dtype, seed=7).to_device()
    z1, z2 = Mat(20, 1, dtype=dtype, seed=8).to_device(), Mat(20, 1, dtype=dtype, seed=8).to_device()
    fex = getattr(blis, f"bli_{ch}gemv_ex")
    fex.argtypes = [C.c_int, C.c_int, I, I, PT, P, I, I, P, I, PT, P, I, P, P]
    getattr(blis, f"bli_{ch}gemv")(TR["t"], 0, nv, 20, C.byref(al), g.ptr, 1, nv, x.ptr, 1,
                                   C.byref(be), z1.ptr, 1)
    fex(TR["t"], 0, nv, 20, C.byref(al), g.ptr, 1, nv, x.ptr, 1, C.byref(be), z2.ptr, 1, None, None)
    np.testing.assert_array_equal(z1.device_view(), z2.device_view())


def test_gemm_ex_object(blis):
    """bli_gemm_ex against bli_gemm, bitwise."""
    say("bli_gemm_ex object")
    m, n, k = 50, 60, 70
    a, b = Mat(m, k, seed=1).to_device(), Mat(k, n, seed=2).to_device()
    c1, c2 = Mat(m, n, seed=3).to_device(), Mat(m, n, seed=3).to_device()
    alo, k1 = _scalar(blis, np.float64, 0.75)
    beo, k2 = _scalar(blis, np.float64, 0.25)
    ao, bo = _obj(blis, np.float64, a), _obj(blis, np.float64, b)
    blis.bli_gemm(C.byref(alo), C.byref(ao), C.byref(bo), C.byref(beo),
                  C.byref(_obj(blis, np.float64, c1)))
    blis.bli_gemm_ex.argtypes = [C.c_void_p] * 7
    blis.bli_gemm_ex(C.byref(alo), C.byref(ao), C.byref(bo), C.byref(beo),
                     C.byref(_obj(blis, np.float64, c2)), None, None)
    np.testing.assert_array_equal(c1.device_view(), c2.device_view())
