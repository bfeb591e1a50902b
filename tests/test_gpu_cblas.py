"""The CBLAS layer on the GPU, row-major and column-major, against the oracle.

A row-major CBLAS matrix X (M x N, leading dimension ldx) has element (i, j)
at X[i*ldx + j]: a C-ordered numpy array. The oracle works on element
semantics with any strides, so the reference result is the oracle's routine
on the same arrays with the same side/uplo/trans; the CBLAS layer must
produce it by the row-major mapping of frame/compat/cblas/src (operands
swapped, side and uplo flipped, transposes toggled). Bar: the routine's
normwise GEMM bar 4 (k + 2) u against the oracle; TRSM: 8 m u relative.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import say

pytestmark = pytest.mark.gpu

ROW, COL = 101, 102
NOTR, TR = 111, 112
UP, LO = 121, 122
NONUNIT, UNITD = 131, 132
LEFT, RIGHT = 141, 142
F = C.c_int32


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x.ravel(order="K"))).to("cuda")


class Arr:
    """A numpy matrix in row- or column-major order with a device copy."""

    def __init__(self, rows, cols, order, dtype, seed, ld_pad=0):
        rng = np.random.default_rng(seed)
        ld = (cols if order == ROW else rows) + ld_pad
        flat = rng.uniform(-1, 1, ld * (rows if order == ROW else cols)).astype(dtype)
        self.order, self.ld, self.rows, self.cols = order, ld, rows, cols
        self.flat = flat
        es = flat.itemsize
        strides = (ld * es, es) if order == ROW else (es, ld * es)
        self.a = np.lib.stride_tricks.as_strided(flat, (rows, cols), strides)
        self.d = None

    def to_device(self):
        import torch
        self.d = torch.from_numpy(self.flat.copy()).to("cuda")
        return self

    @property
    def ptr(self):
        return C.c_void_p(self.d.data_ptr())

    def result(self):
        import torch
        torch.cuda.synchronize()
        flat = self.d.cpu().numpy()
        es = flat.itemsize
        strides = (self.ld * es, es) if self.order == ROW else (es, self.ld * es)
        return np.lib.stride_tricks.as_strided(flat, (self.rows, self.cols), strides).copy()


def _T(dtype):
    return C.c_double if dtype == np.float64 else C.c_float


def _u(dtype):
    return 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24


def _err(got, ref, scale):
    return np.linalg.norm((got - ref).astype(np.float64)) / scale


def _fn(blis, name, restype=None):
    f = getattr(blis, name)
    f.restype = restype
    return f


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", [ROW, COL])
def test_cblas_gemm_gemmt(blis, oracle, dtype, order):
    T = _T(dtype)
    ch = "d" if dtype == np.float64 else "s"
    u = _u(dtype)
    m, n, k = 75, 61, 43
    for ta in (NOTR, TR):
        for tb in (NOTR, TR):
            say(f"cblas_{ch}gemm order={order} {ta} {tb}")
            a = Arr(*((k, m) if ta == TR else (m, k)), order, dtype, 1, ld_pad=3).to_device()
            b = Arr(*((n, k) if tb == TR else (k, n)), order, dtype, 2).to_device()
            c = Arr(m, n, order, dtype, 3, ld_pad=1).to_device()
            c0 = c.a.copy()
            _fn(blis, f"cblas_{ch}gemm")(order, ta, tb, m, n, k, T(0.5), a.ptr, a.ld, b.ptr, b.ld,
                                         T(1.5), c.ptr, c.ld)
            ref = c0.copy(order="F")
            oracle.gemm("t" if ta == TR else "n", "t" if tb == TR else "n", 0.5, a.a, b.a, 1.5,
                        ref)
            scale = 0.5 * np.linalg.norm(a.a) * np.linalg.norm(b.a) + 1.5 * np.linalg.norm(c0)
            assert _err(c.result(), ref, scale) <= 4 * (k + 2) * u
    for uplo in (UP, LO):
        say(f"cblas_{ch}gemmt order={order} uplo={uplo}")
        a = Arr(m, k, order, dtype, 4).to_device()
        b = Arr(k, m, order, dtype, 5).to_device()
        c = Arr(m, m, order, dtype, 6).to_device()
        c0 = c.a.copy()
        _fn(blis, f"cblas_{ch}gemmt")(order, uplo, NOTR, NOTR, m, k, T(1.0), a.ptr, a.ld, b.ptr,
                                      b.ld, T(0.5), c.ptr, c.ld)
        ref = c0.copy(order="F")
        oracle.gemmt(0x60 if uplo == UP else 0xC0, "n", "n", 1.0, a.a, b.a, 0.5, ref)
        got = c.result()
        mk = np.triu(np.ones((m, m), bool)) if uplo == UP else np.tril(np.ones((m, m), bool))
        assert np.array_equal(got[~mk], c0[~mk])
        scale = np.linalg.norm(a.a) * np.linalg.norm(b.a) + 0.5 * np.linalg.norm(c0)
        assert _err(got[mk], ref[mk], scale) <= 4 * (k + 2) * u


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", [ROW, COL])
def test_cblas_structured(blis, oracle, dtype, order):
    T = _T(dtype)
    ch = "d" if dtype == np.float64 else "s"
    u = _u(dtype)
    m, n, k = 57, 39, 26
    L = {UP: 0x60, LO: 0xC0}
    for side in (LEFT, RIGHT):
        for uplo in (UP, LO):
            mn = m if side == LEFT else n
            say(f"cblas_{ch}symm/trmm/trsm order={order} side={side} uplo={uplo}")
            a = Arr(mn, mn, order, dtype, 7).to_device()
            b = Arr(m, n, order, dtype, 8).to_device()
            c = Arr(m, n, order, dtype, 9).to_device()
            c0 = c.a.copy()
            _fn(blis, f"cblas_{ch}symm")(order, side, uplo, m, n, T(0.5), a.ptr, a.ld, b.ptr,
                                         b.ld, T(2.0), c.ptr, c.ld)
            ref = c0.copy(order="F")
            oracle.symm(0 if side == LEFT else 1, L[uplo], "n", 0.5, a.a, b.a, 2.0, ref)
            scale = 0.5 * 2 * np.linalg.norm(a.a) * np.linalg.norm(b.a) + 2 * np.linalg.norm(c0)
            assert _err(c.result(), ref, scale) <= 4 * (mn + 2) * u
            for ta in (NOTR, TR):
                bt = Arr(m, n, order, dtype, 10).to_device()
                b0 = bt.a.copy()
                _fn(blis, f"cblas_{ch}trmm")(order, side, uplo, ta, NONUNIT, m, n, T(1.5), a.ptr,
                                             a.ld, bt.ptr, bt.ld)
                ref = b0.copy(order="F")
                oracle.trmm(0 if side == LEFT else 1, L[uplo], "t" if ta == TR else "n", 0,
                            1.5, a.a, ref)
                scale = 1.5 * np.linalg.norm(a.a) * np.linalg.norm(b0)
                assert _err(bt.result(), ref, scale) <= 4 * (mn + 2) * u
                # trsm with a well-conditioned triangle
                ad = Arr(mn, mn, order, dtype, 11)
                ad.a[...] = ad.a / mn + np.eye(mn, dtype=dtype) * 2
                ad.to_device()
                bs = Arr(m, n, order, dtype, 12).to_device()
                b0 = bs.a.copy()
                _fn(blis, f"cblas_{ch}trsm")(order, side, uplo, ta, UNITD, m, n, T(1.0), ad.ptr,
                                             ad.ld, bs.ptr, bs.ld)
                ref = b0.copy(order="F")
                oracle.trsm(0 if side == LEFT else 1, L[uplo], "t" if ta == TR else "n", 0x100,
                            1.0, ad.a, ref)
                got = bs.result()
                assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 8 * max(m, n) * u
    for uplo in (UP, LO):
        for tr in (NOTR, TR):
            say(f"cblas_{ch}syrk/syr2k order={order} uplo={uplo} trans={tr}")
            a = Arr(*((k, m) if tr == TR else (m, k)), order, dtype, 13).to_device()
            b = Arr(*((k, m) if tr == TR else (m, k)), order, dtype, 14).to_device()
            mk = np.triu(np.ones((m, m), bool)) if uplo == UP else np.tril(np.ones((m, m), bool))
            for op in ("syrk", "syr2k"):
                c = Arr(m, m, order, dtype, 15).to_device()
                c0 = c.a.copy()
                if op == "syrk":
                    _fn(blis, f"cblas_{ch}syrk")(order, uplo, tr, m, k, T(0.5), a.ptr, a.ld,
                                                 T(2.0), c.ptr, c.ld)
                    ref = c0.copy(order="F")
                    oracle.syrk(L[uplo], "t" if tr == TR else "n", 0.5, a.a, 2.0, ref)
                else:
                    _fn(blis, f"cblas_{ch}syr2k")(order, uplo, tr, m, k, T(0.5), a.ptr, a.ld,
                                                  b.ptr, b.ld, T(2.0), c.ptr, c.ld)
                    ref = c0.copy(order="F")
                    t = "t" if tr == TR else "n"
                    oracle.syr2k(L[uplo], t, t, 0.5, a.a, b.a, 2.0, ref)
                got = c.result()
                assert np.array_equal(got[~mk], c0[~mk])
                scale = np.linalg.norm(a.a) * np.linalg.norm(b.a) + 2 * np.linalg.norm(c0)
                assert _err(got[mk], ref[mk], scale) <= 8 * (k + 2) * u


@pytest.mark.parametrize("order", [ROW, COL])
def test_cblas_level12_and_batch(blis, oracle, order):
    import torch
    m, n = 300, 170
    say(f"cblas_dgemv order={order}")
    a = Arr(m, n, order, np.float64, 20, ld_pad=2).to_device()
    for ta in (NOTR, TR):
        nx, ny = (n, m) if ta == NOTR else (m, n)
        x = np.random.default_rng(21).uniform(-1, 1, nx)
        y = np.random.default_rng(22).uniform(-1, 1, ny)
        dx, dy = _dev(x), _dev(y)
        _fn(blis, "cblas_dgemv")(order, ta, m, n, C.c_double(0.5), a.ptr, a.ld,
                                 C.c_void_p(dx.data_ptr()), 1, C.c_double(1.5),
                                 C.c_void_p(dy.data_ptr()), 1)
        ref = y.copy()
        oracle.gemv("t" if ta == TR else "n", 0.5, a.a, x, 1.5, ref)
        torch.cuda.synchronize()
        assert np.allclose(dy.cpu().numpy(), ref, rtol=0, atol=1e-12)
    say("cblas_ddot / cblas_daxpy")
    x = np.random.default_rng(23).uniform(-1, 1, 1000)
    y = np.random.default_rng(24).uniform(-1, 1, 1000)
    dx, dy = _dev(x), _dev(y)
    dot = _fn(blis, "cblas_ddot", C.c_double)(1000, C.c_void_p(dx.data_ptr()), 1,
                                              C.c_void_p(dy.data_ptr()), 1)
    assert abs(dot - oracle.dotv(x, y)) <= 1e-12
    _fn(blis, "cblas_daxpy")(500, C.c_double(2.0), C.c_void_p(dx.data_ptr()), 2,
                             C.c_void_p(dy.data_ptr()), 2)
    ref = y.copy()
    oracle.axpyv(2.0, x, ref, 2, 2, n=500)
    torch.cuda.synchronize()
    assert np.allclose(dy.cpu().numpy(), ref, rtol=0, atol=1e-15)

    say(f"cblas_dgemm_batch order={order}")
    groups = [(NOTR, TR, 40, 30, 20, 2), (TR, NOTR, 25, 35, 15, 3)]
    ta_arr = (C.c_int * 2)(*[g[0] for g in groups])
    tb_arr = (C.c_int * 2)(*[g[1] for g in groups])
    M = (F * 2)(*[g[2] for g in groups])
    N = (F * 2)(*[g[3] for g in groups])
    K = (F * 2)(*[g[4] for g in groups])
    GS = (F * 2)(*[g[5] for g in groups])
    alphas = (C.c_double * 2)(0.5, -1.0)
    betas = (C.c_double * 2)(1.0, 0.25)
    As, Bs, Cs, refs, LDA, LDB, LDC = [], [], [], [], [], [], []
    for gi, (ta, tb, mm, nn, kk, cnt) in enumerate(groups):
        for s in range(cnt):
            av = Arr(*((kk, mm) if ta == TR else (mm, kk)), order, np.float64, 30 + s).to_device()
            bv = Arr(*((nn, kk) if tb == TR else (kk, nn)), order, np.float64, 40 + s).to_device()
            cv = Arr(mm, nn, order, np.float64, 50 + s).to_device()
            ref = cv.a.copy(order="F")
            oracle.gemm("t" if ta == TR else "n", "t" if tb == TR else "n", alphas[gi], av.a,
                        bv.a, betas[gi], ref)
            As.append(av); Bs.append(bv); Cs.append(cv); refs.append(ref)
        LDA.append(As[-1].ld); LDB.append(Bs[-1].ld); LDC.append(Cs[-1].ld)
    PA = (C.c_void_p * len(As))(*[x.d.data_ptr() for x in As])
    PB = (C.c_void_p * len(Bs))(*[x.d.data_ptr() for x in Bs])
    PC = (C.c_void_p * len(Cs))(*[x.d.data_ptr() for x in Cs])
    _fn(blis, "cblas_dgemm_batch")(order, ta_arr, tb_arr, M, N, K, alphas, PA, (F * 2)(*LDA), PB,
                                   (F * 2)(*LDB), betas, PC, (F * 2)(*LDC), 2, GS)
    for cv, ref in zip(Cs, refs):
        assert np.allclose(cv.result(), ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("order", [ROW, COL])
def test_cblas_zgemm(blis, oracle, order):
    import torch

    class CX(C.Structure):
        _fields_ = [("real", C.c_double), ("imag", C.c_double)]

    m, n, k = 37, 29, 21
    rng = np.random.default_rng(60)

    def rnd(r, c):
        x = rng.uniform(-1, 1, (r, c)) + 1j * rng.uniform(-1, 1, (r, c))
        return np.ascontiguousarray(x) if order == ROW else np.asfortranarray(x)

    for ta, tstr in ((NOTR, "n"), (113, "c")):
        say(f"cblas_zgemm order={order} transa={ta}")
        a = rnd(*((k, m) if tstr == "c" else (m, k)))
        b, c0 = rnd(k, n), rnd(m, n)
        da, db, dc = _dev(a), _dev(b), _dev(c0)
        lda = a.shape[1] if order == ROW else a.shape[0]
        ldb = b.shape[1] if order == ROW else b.shape[0]
        ldc = c0.shape[1] if order == ROW else c0.shape[0]
        al, be = CX(0.5, -0.25), CX(1.0, 0.5)
        _fn(blis, "cblas_zgemm")(order, ta, NOTR, m, n, k, C.byref(al), C.c_void_p(da.data_ptr()),
                                 lda, C.c_void_p(db.data_ptr()), ldb, C.byref(be),
                                 C.c_void_p(dc.data_ptr()), ldc)
        torch.cuda.synchronize()
        flat = dc.cpu().numpy()
        got = flat.reshape((m, n)) if order == ROW else flat.reshape((n, m)).T
        ref = np.asfortranarray(c0.copy())
        oracle.zgemm(tstr, "n", 0.5 - 0.25j, np.asfortranarray(a), np.asfortranarray(b),
                     1.0 + 0.5j, ref)
        err = np.linalg.norm(got - ref) / (np.linalg.norm(a) * np.linalg.norm(b) +
                                           np.linalg.norm(c0))
        assert err <= 8 * (k + 2) * 2.0 ** -53
