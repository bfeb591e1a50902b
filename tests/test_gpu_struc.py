"""Structured level-3 (gemmt, syrk, syr2k, symm, trmm, trmm3) on the GPU
through the C ABI, against the oracle.

Bars:
  * every stored element within the GEMM bar, normwise error
    <= 4 (k + 2) u against the oracle (the same products on the same data);
  * the unstored triangle of C (gemmt, syrk, syr2k) is NaN on entry and must
    come back NaN: it is never written; the unstored triangle of A (symm,
    trmm, trmm3) and a unit diagonal are NaN: they are never read;
  * at 4096 (the size where the 128 x 128 tile runs), the structured kernels
    against the plain GEMM on the explicitly structured dense matrices,
    bit for bit: the k-range skip of trmm only drops leading/trailing exact
    zero terms, and gemmt's masked store writes the same values GEMM does.
The reference's netlib ?blat3 driver exercises ?symm_/?trmm_/?syrk_/?syr2k_
(test_gpu_netlib.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import blis_amd
from tests.conftest import Mat, say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}
LOWER, UPPER = 0xC0, 0x60
LEFT, RIGHT = 0, 1
UNIT = 0x100
DT = {np.dtype(np.float64): blis_amd.BLIS_DOUBLE, np.dtype(np.float32): blis_amd.BLIS_FLOAT}
SHAPES = [(1, 1), (7, 5), (67, 45), (200, 130), (300, 33)]


def _ch(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def _T(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def _mask(m, uplo):
    one = np.ones((m, m), bool)
    return np.tril(one) if uplo == LOWER else np.triu(one)


def _err(got, ref, scale):
    d = (got.astype(np.float64) - ref.astype(np.float64))
    return np.linalg.norm(d) / scale if scale > 0 else np.linalg.norm(d)


def _nan_outside(mat: Mat, m, uplo):
    v = mat.view()
    v[~_mask(m, uplo)] = np.nan


def _struct_a(mn, uplo, unit, dtype, seed, rs=None, cs=None):
    """Square operand whose unstored triangle (and unit diagonal) is NaN."""
    a = Mat(mn, mn, rs=rs, cs=cs, dtype=dtype, seed=seed)
    v = a.view()
    v[~_mask(mn, uplo)] = np.nan
    if unit:
        v[np.diag_indices(mn)] = np.nan
    return a


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("tb", ["n", "t"])
def test_gemmt(blis, oracle, dtype, uplo, ta, tb):
    u = unit_roundoff(dtype)
    T = _T(dtype)
    for (m, k) in SHAPES:
        say(f"gemmt {np.dtype(dtype).name} uplo={uplo:#x} {ta}{tb} m={m} k={k}")
        a = Mat(k if ta == "t" else m, m if ta == "t" else k, dtype=dtype, seed=m + k)
        b = Mat(m if tb == "t" else k, k if tb == "t" else m, dtype=dtype, seed=m * 3 + k)
        c = Mat(m, m, dtype=dtype, seed=m * 5 + k)
        _nan_outside(c, m, uplo)
        c0 = c.view().copy()
        for x in (a, b, c):
            x.to_device()
        getattr(blis, f"bli_{_ch(dtype)}gemmt")(uplo, TR[ta], TR[tb], m, k, C.byref(T(0.7)),
                                                a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs,
                                                C.byref(T(-1.3)), c.ptr, c.rs, c.cs)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.gemmt(uplo, ta, tb, 0.7, a.view().copy(order="F"), b.view().copy(order="F"),
                     -1.3, ref)
        mk = _mask(m, uplo)
        assert np.isnan(got[~mk]).all(), "gemmt wrote outside the stored triangle"
        scale = 0.7 * np.linalg.norm(a.view().astype(np.float64)) * \
            np.linalg.norm(b.view().astype(np.float64)) + 1.3 * np.linalg.norm(c0[mk])
        assert _err(got[mk], ref[mk], scale) <= 4 * (k + 2) * u


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layout", ["row", "general", "host"])
def test_gemmt_layouts(blis, oracle, dtype, layout):
    """Row-stored C (the operation transpose swaps the triangle), general
    strides everywhere (operands packed), and host pointers (staged)."""
    u = unit_roundoff(dtype)
    T = _T(dtype)
    m, k = 150, 70
    for uplo in (LOWER, UPPER):
        say(f"gemmt {layout} {np.dtype(dtype).name} uplo={uplo:#x}")
        if layout == "row":
            a = Mat(m, k, rs=k, cs=1, dtype=dtype, seed=1)
            b = Mat(k, m, rs=m, cs=1, dtype=dtype, seed=2)
            c = Mat(m, m, rs=m + 3, cs=1, dtype=dtype, seed=3)
        elif layout == "general":
            a = Mat(m, k, rs=2, cs=2 * m + 1, dtype=dtype, seed=1)
            b = Mat(k, m, rs=3, cs=3 * k, dtype=dtype, seed=2)
            c = Mat(m, m, rs=2, cs=2 * m + 2, dtype=dtype, seed=3)
        else:
            a = Mat(m, k, dtype=dtype, seed=1)
            b = Mat(k, m, dtype=dtype, seed=2)
            c = Mat(m, m, dtype=dtype, seed=3)
        _nan_outside(c, m, uplo)
        c0 = c.view().copy()
        host = layout == "host"
        if not host:
            for x in (a, b, c):
                x.to_device()
        ptr = (lambda x: x.hptr) if host else (lambda x: x.ptr)
        getattr(blis, f"bli_{_ch(dtype)}gemmt")(uplo, 0, 0, m, k, C.byref(T(1.0)), ptr(a),
                                                a.rs, a.cs, ptr(b), b.rs, b.cs, C.byref(T(0.5)),
                                                ptr(c), c.rs, c.cs)
        got = c.view().copy() if host else c.device_view()
        ref = c0.copy(order="F")
        oracle.gemmt(uplo, "n", "n", 1.0, a.view().copy(order="F"), b.view().copy(order="F"),
                     0.5, ref)
        mk = _mask(m, uplo)
        assert np.isnan(got[~mk]).all()
        scale = np.linalg.norm(a.view().astype(np.float64)) * \
            np.linalg.norm(b.view().astype(np.float64)) + 0.5 * np.linalg.norm(c0[mk])
        assert _err(got[mk], ref[mk], scale) <= 4 * (k + 2) * u


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
def test_syrk_syr2k(blis, oracle, dtype, uplo, ta):
    u = unit_roundoff(dtype)
    T = _T(dtype)
    ch = _ch(dtype)
    for (m, k) in [(5, 3), (90, 40), (257, 129)]:
        say(f"syrk/syr2k {np.dtype(dtype).name} uplo={uplo:#x} {ta} m={m} k={k}")
        shape = (k, m) if ta == "t" else (m, k)
        a = Mat(*shape, dtype=dtype, seed=m)
        b = Mat(*shape, dtype=dtype, seed=m + 1)
        c = Mat(m, m, dtype=dtype, seed=m + 2)
        _nan_outside(c, m, uplo)
        c0 = c.view().copy()
        for x in (a, b, c):
            x.to_device()
        getattr(blis, f"bli_{ch}syrk")(uplo, TR[ta], m, k, C.byref(T(0.5)), a.ptr, a.rs, a.cs,
                                       C.byref(T(2.0)), c.ptr, c.rs, c.cs)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.syrk(uplo, ta, 0.5, a.view().copy(order="F"), 2.0, ref)
        mk = _mask(m, uplo)
        na = np.linalg.norm(a.view().astype(np.float64))
        nb = np.linalg.norm(b.view().astype(np.float64))
        assert np.isnan(got[~mk]).all()
        assert _err(got[mk], ref[mk], 0.5 * na * na + 2 * np.linalg.norm(c0[mk])) <= 4 * (k + 2) * u
        # syr2k on the same C
        c.dbuf.copy_(__import__("torch").from_numpy(c.buf.copy()))
        getattr(blis, f"bli_{ch}syr2k")(uplo, TR[ta], TR[ta], m, k, C.byref(T(0.5)), a.ptr,
                                        a.rs, a.cs, b.ptr, b.rs, b.cs, C.byref(T(2.0)), c.ptr,
                                        c.rs, c.cs)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.syr2k(uplo, ta, ta, 0.5, a.view().copy(order="F"), b.view().copy(order="F"),
                     2.0, ref)
        assert np.isnan(got[~mk]).all()
        assert _err(got[mk], ref[mk], na * nb + 2 * np.linalg.norm(c0[mk])) <= 8 * (k + 2) * u


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("tb", ["n", "t"])
def test_symm(blis, oracle, dtype, side, uplo, tb):
    u = unit_roundoff(dtype)
    T = _T(dtype)
    for (m, n) in SHAPES:
        say(f"symm {np.dtype(dtype).name} side={side} uplo={uplo:#x} {tb} m={m} n={n}")
        mn = m if side == LEFT else n
        a = _struct_a(mn, uplo, False, dtype, seed=mn)
        b = Mat(n if tb == "t" else m, m if tb == "t" else n, dtype=dtype, seed=m + n)
        c = Mat(m, n, dtype=dtype, seed=m + 2 * n)
        c0 = c.view().copy()
        for x in (a, b, c):
            x.to_device()
        getattr(blis, f"bli_{_ch(dtype)}symm")(side, uplo, 0, TR[tb], m, n, C.byref(T(0.9)),
                                               a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs,
                                               C.byref(T(0.25)), c.ptr, c.rs, c.cs)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.symm(side, uplo, tb, 0.9, a.view().copy(order="F"), b.view().copy(order="F"),
                    0.25, ref)
        assert np.isfinite(got).all(), "symm read the unstored triangle"
        av = np.nan_to_num(a.view().astype(np.float64))
        scale = 0.9 * 2 * np.linalg.norm(av) * np.linalg.norm(b.view().astype(np.float64)) + \
            0.25 * np.linalg.norm(c0)
        assert _err(got, ref, scale) <= 4 * (mn + 2) * u


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("diag", [0, UNIT])
def test_trmm(blis, oracle, dtype, side, uplo, ta, diag):
    u = unit_roundoff(dtype)
    T = _T(dtype)
    for (m, n) in SHAPES:
        say(f"trmm {np.dtype(dtype).name} side={side} uplo={uplo:#x} {ta} diag={diag:#x} "
            f"m={m} n={n}")
        mn = m if side == LEFT else n
        a = _struct_a(mn, uplo, diag == UNIT, dtype, seed=mn + 1)
        b = Mat(m, n, dtype=dtype, seed=m + 7 * n)
        b0 = b.view().copy()
        a.to_device()
        b.to_device()
        getattr(blis, f"bli_{_ch(dtype)}trmm")(side, uplo, TR[ta], diag, m, n,
                                               C.byref(T(1.5)), a.ptr, a.rs, a.cs, b.ptr, b.rs,
                                               b.cs)
        got = b.device_view()
        ref = b0.copy(order="F")
        oracle.trmm(side, uplo, ta, diag, 1.5, a.view().copy(order="F"), ref)
        assert np.isfinite(got).all(), "trmm read the unstored triangle or the unit diagonal"
        av = np.nan_to_num(a.view().astype(np.float64))
        scale = 1.5 * (np.linalg.norm(av) + np.sqrt(mn)) * np.linalg.norm(b0.astype(np.float64))
        assert _err(got, ref, scale) <= 4 * (mn + 2) * u


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [LEFT, RIGHT])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("tb", ["n", "t"])
def test_trmm3(blis, oracle, dtype, side, uplo, ta, tb):
    u = unit_roundoff(dtype)
    T = _T(dtype)
    for (m, n), diag in [((7, 5), 0), ((67, 45), UNIT), ((200, 130), 0)]:
        say(f"trmm3 {np.dtype(dtype).name} side={side} uplo={uplo:#x} {ta}{tb} diag={diag:#x} "
            f"m={m} n={n}")
        mn = m if side == LEFT else n
        a = _struct_a(mn, uplo, diag == UNIT, dtype, seed=mn + 2)
        b = Mat(n if tb == "t" else m, m if tb == "t" else n, dtype=dtype, seed=3 * m + n)
        c = Mat(m, n, dtype=dtype, seed=m + n)
        c0 = c.view().copy()
        for x in (a, b, c):
            x.to_device()
        getattr(blis, f"bli_{_ch(dtype)}trmm3")(side, uplo, TR[ta], diag, TR[tb], m, n,
                                                C.byref(T(-0.5)), a.ptr, a.rs, a.cs, b.ptr,
                                                b.rs, b.cs, C.byref(T(1.25)), c.ptr, c.rs, c.cs)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.trmm3(side, uplo, ta, diag, tb, -0.5, a.view().copy(order="F"),
                     b.view().copy(order="F"), 1.25, ref)
        assert np.isfinite(got).all()
        av = np.nan_to_num(a.view().astype(np.float64))
        scale = 0.5 * (np.linalg.norm(av) + np.sqrt(mn)) * \
            np.linalg.norm(b.view().astype(np.float64)) + 1.25 * np.linalg.norm(c0)
        assert _err(got, ref, scale) <= 4 * (mn + 2) * u


def test_struc_alpha_zero_and_empty(blis):
    """alpha == 0: gemmt scales only the stored triangle (beta = 0 writes
    zeros without reading NaNs there), trmm sets B := 0; m == 0 is a no-op."""
    m, k = 40, 9
    a = Mat(m, k, seed=1).to_device()
    c = Mat(m, m, seed=2)
    c.view()[:] = np.nan
    _nan_outside(c, m, LOWER)
    c.to_device()
    say("dgemmt alpha=0 beta=0")
    blis.bli_dgemmt(LOWER, 0, 8, m, k, C.byref(C.c_double(0.0)), a.ptr, a.rs, a.cs, a.ptr,
                    a.cs, a.rs, C.byref(C.c_double(0.0)), c.ptr, c.rs, c.cs)
    got = c.device_view()
    mk = _mask(m, LOWER)
    assert (got[mk] == 0).all() and np.isnan(got[~mk]).all()
    say("dtrmm alpha=0")
    t = _struct_a(m, UPPER, False, np.float64, seed=4).to_device()
    b = Mat(m, 6, seed=5).to_device()
    blis.bli_dtrmm(LEFT, UPPER, 0, 0, m, 6, C.byref(C.c_double(0.0)), t.ptr, t.rs, t.cs, b.ptr,
                   b.rs, b.cs)
    assert (b.device_view() == 0).all()
    blis.bli_dtrmm(LEFT, UPPER, 0, 0, 0, 6, C.byref(C.c_double(1.0)), t.ptr, t.rs, t.cs, b.ptr,
                   b.rs, b.cs)


def _normz(x):
    return np.where(x == 0, 0.0, x)   # +0 and -0 compare as one value


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_struc_full_size_bitwise_vs_gemm(blis, dtype):
    """At 4096 the structured kernels run the large tile; they must equal the
    plain GEMM (already pinned) on the explicitly structured dense operands."""
    import torch
    n = 4096
    T = _T(dtype)
    ch = _ch(dtype)
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    g = torch.Generator(device="cuda").manual_seed(7)
    a = torch.rand((n, n), device="cuda", dtype=tdt, generator=g) * 2 - 1   # column-major A^T
    b = torch.rand((n, n), device="cuda", dtype=tdt, generator=g) * 2 - 1
    c0 = torch.rand((n, n), device="cuda", dtype=tdt, generator=g) * 2 - 1
    # Column-major views: element (i, j) of X is X_t[j, i].
    a_lower_dense = torch.tril(a.T).T.contiguous()   # lower triangle of A, column-major
    one, half = C.byref(T(1.0)), C.byref(T(0.5))

    say(f"{ch}trmm L/L/N/N {n}x{n} vs {ch}gemm on tril(A)")
    x = b.clone()
    getattr(blis, f"bli_{ch}trmm")(LEFT, LOWER, 0, 0, n, n, one, a.data_ptr(), 1, n,
                                   x.data_ptr(), 1, n)
    y = torch.zeros_like(b)
    getattr(blis, f"bli_{ch}gemm")(0, 0, n, n, n, one, a_lower_dense.data_ptr(), 1, n,
                                   b.data_ptr(), 1, n, C.byref(T(0.0)), y.data_ptr(), 1, n)
    torch.cuda.synchronize()
    assert np.array_equal(_normz(x.cpu().numpy()), _normz(y.cpu().numpy()))

    say(f"{ch}gemmt lower {n} vs {ch}gemm")
    cg = c0.clone()
    getattr(blis, f"bli_{ch}gemmt")(LOWER, 0, 0, n, n, one, a.data_ptr(), 1, n, b.data_ptr(),
                                    1, n, half, cg.data_ptr(), 1, n)
    cf = c0.clone()
    getattr(blis, f"bli_{ch}gemm")(0, 0, n, n, n, one, a.data_ptr(), 1, n, b.data_ptr(), 1, n,
                                   half, cf.data_ptr(), 1, n)
    torch.cuda.synchronize()
    cg_np, cf_np, c0_np = cg.cpu().numpy().T, cf.cpu().numpy().T, c0.cpu().numpy().T
    mk = np.tril(np.ones((n, n), bool))
    assert np.array_equal(cg_np[mk], cf_np[mk])
    assert np.array_equal(cg_np[~mk], c0_np[~mk])

    say(f"{ch}symm L/U {n} vs {ch}gemm on the mirrored A")
    a_sym = torch.triu(a.T)
    a_sym = (a_sym + torch.triu(a_sym, 1).T).T.contiguous()   # column-major full symmetric
    cs = c0.clone()
    getattr(blis, f"bli_{ch}symm")(LEFT, UPPER, 0, 0, n, n, one, a.data_ptr(), 1, n,
                                   b.data_ptr(), 1, n, half, cs.data_ptr(), 1, n)
    cf = c0.clone()
    getattr(blis, f"bli_{ch}gemm")(0, 0, n, n, n, one, a_sym.data_ptr(), 1, n, b.data_ptr(), 1,
                                   n, half, cf.data_ptr(), 1, n)
    torch.cuda.synchronize()
    assert np.array_equal(cs.cpu().numpy(), cf.cpu().numpy())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_struc_object_api(blis, oracle, dtype):
    """bli_gemmt/syrk/syr2k/symm/trmm/trmm3 read uplo, structure, diag and
    trans from the obj_t info bits and match the typed entry points bit for bit."""
    T = _T(dtype)
    ch = _ch(dtype)
    m, n, k = 45, 31, 23

    def obj(mat, uplo=None, struc=None, diag=None, trans=False):
        o = blis_amd.obj_t()
        blis.bli_obj_create_with_attached_buffer(DT[np.dtype(dtype)], mat.m, mat.n, mat.ptr,
                                                 mat.rs, mat.cs, C.byref(o))
        if uplo is not None:
            o.info = (o.info & ~(0x7 << 5)) | uplo
        if struc is not None:
            o.info = (o.info & ~(0x3 << 20)) | struc
        if diag is not None:
            o.info = (o.info & ~0x100) | diag
        if trans:
            o.info ^= 0x8
        return o

    def scal(v):
        x = T(v)
        o = blis_amd.obj_t()
        blis.bli_obj_create_1x1_with_attached_buffer(DT[np.dtype(dtype)], C.byref(x), C.byref(o))
        return o, x

    al, alv = scal(0.75)
    be, bev = scal(-0.5)

    # gemmt (C upper, A transposed)
    a = Mat(k, m, dtype=dtype, seed=1).to_device()
    b = Mat(k, m, dtype=dtype, seed=2).to_device()
    c1 = Mat(m, m, dtype=dtype, seed=3).to_device()
    c2 = Mat(m, m, dtype=dtype, seed=3).to_device()
    say(f"bli_gemmt object {ch}")
    blis.bli_gemmt(C.byref(al), C.byref(obj(a, trans=True)), C.byref(obj(b)), C.byref(be),
                   C.byref(obj(c1, uplo=UPPER)))
    getattr(blis, f"bli_{ch}gemmt")(UPPER, 8, 0, m, k, C.byref(alv), a.ptr, a.rs, a.cs, b.ptr,
                                    b.rs, b.cs, C.byref(bev), c2.ptr, c2.rs, c2.cs)
    assert np.array_equal(c1.device_view(), c2.device_view())

    # syrk / syr2k (C lower, A m x k)
    a = Mat(m, k, dtype=dtype, seed=4).to_device()
    b = Mat(m, k, dtype=dtype, seed=5).to_device()
    for op in ("syrk", "syr2k"):
        c1 = Mat(m, m, dtype=dtype, seed=6).to_device()
        c2 = Mat(m, m, dtype=dtype, seed=6).to_device()
        say(f"bli_{op} object {ch}")
        if op == "syrk":
            blis.bli_syrk(C.byref(al), C.byref(obj(a)), C.byref(be), C.byref(obj(c1, LOWER)))
            getattr(blis, f"bli_{ch}syrk")(LOWER, 0, m, k, C.byref(alv), a.ptr, a.rs, a.cs,
                                           C.byref(bev), c2.ptr, c2.rs, c2.cs)
        else:
            blis.bli_syr2k(C.byref(al), C.byref(obj(a)), C.byref(obj(b)), C.byref(be),
                           C.byref(obj(c1, LOWER)))
            getattr(blis, f"bli_{ch}syr2k")(LOWER, 0, 0, m, k, C.byref(alv), a.ptr, a.rs, a.cs,
                                            b.ptr, b.rs, b.cs, C.byref(bev), c2.ptr, c2.rs, c2.cs)
        assert np.array_equal(c1.device_view(), c2.device_view())

    # symm right, lower
    s = _struct_a(n, LOWER, False, dtype, seed=7).to_device()
    b = Mat(m, n, dtype=dtype, seed=8).to_device()
    c1 = Mat(m, n, dtype=dtype, seed=9).to_device()
    c2 = Mat(m, n, dtype=dtype, seed=9).to_device()
    say(f"bli_symm object {ch}")
    blis.bli_symm(RIGHT, C.byref(al), C.byref(obj(s, LOWER, blis_amd.BLIS_SYMMETRIC)),
                  C.byref(obj(b)), C.byref(be), C.byref(obj(c1)))
    getattr(blis, f"bli_{ch}symm")(RIGHT, LOWER, 0, 0, m, n, C.byref(alv), s.ptr, s.rs, s.cs,
                                   b.ptr, b.rs, b.cs, C.byref(bev), c2.ptr, c2.rs, c2.cs)
    assert np.array_equal(c1.device_view(), c2.device_view())

    # trmm left upper transposed unit; trmm3 right lower
    t = _struct_a(m, UPPER, True, dtype, seed=10).to_device()
    b1 = Mat(m, n, dtype=dtype, seed=11).to_device()
    b2 = Mat(m, n, dtype=dtype, seed=11).to_device()
    say(f"bli_trmm object {ch}")
    blis.bli_trmm(LEFT, C.byref(al), C.byref(obj(t, UPPER, blis_amd.BLIS_TRIANGULAR, UNIT,
                                                 trans=True)), C.byref(obj(b1)))
    getattr(blis, f"bli_{ch}trmm")(LEFT, UPPER, 8, UNIT, m, n, C.byref(alv), t.ptr, t.rs, t.cs,
                                   b2.ptr, b2.rs, b2.cs)
    assert np.array_equal(b1.device_view(), b2.device_view())
    t = _struct_a(n, LOWER, False, dtype, seed=12).to_device()
    b = Mat(m, n, dtype=dtype, seed=13).to_device()
    c1 = Mat(m, n, dtype=dtype, seed=14).to_device()
    c2 = Mat(m, n, dtype=dtype, seed=14).to_device()
    say(f"bli_trmm3 object {ch}")
    blis.bli_trmm3(RIGHT, C.byref(al), C.byref(obj(t, LOWER, blis_amd.BLIS_TRIANGULAR)),
                   C.byref(obj(b)), C.byref(be), C.byref(obj(c1)))
    getattr(blis, f"bli_{ch}trmm3")(RIGHT, LOWER, 0, 0, 0, m, n, C.byref(alv), t.ptr, t.rs,
                                    t.cs, b.ptr, b.rs, b.cs, C.byref(bev), c2.ptr, c2.rs, c2.cs)
    assert np.array_equal(c1.device_view(), c2.device_view())
