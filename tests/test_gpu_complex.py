"""Complex GEMM (1m) and mixed-datatype GEMM on the GPU through the C ABI.

Bars:
  * complex GEMM against the oracle's 1m GEMM: normwise error
    ||C - C_ref||_F / (|alpha| ||A|| ||B|| + |beta| ||C0||) <= 8 (k + 2) u
    (a complex product is 2k real products deep);
  * at 2048 (no oracle), the probe residual of the testsuite
    (testsuite/src/test_gemm.c:303-414): ||C x - (beta C0 x + alpha A (B x))||
    relative to the norms, within 16 k u;
  * the BLAS, object and CBLAS forms give the same bits as the typed call;
  * the reference's netlib ZGEMM/CGEMM drivers run against the library in
    test_gpu_netlib.py;
  * mixed datatypes against the oracle's restatement of BLIS's rules
    (oracle.gemm_md), within the bar of the computation precision.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import blis_amd
from tests.conftest import say

pytestmark = pytest.mark.gpu

TRB = {"n": 0x0, "t": 0x8, "r": 0x10, "c": 0x18}
DT = {np.dtype(np.float32): 0, np.dtype(np.float64): 2, np.dtype(np.complex64): 1,
      np.dtype(np.complex128): 3}


class CX(C.Structure):
    _fields_ = [("real", C.c_double), ("imag", C.c_double)]


class SX(C.Structure):
    _fields_ = [("real", C.c_float), ("imag", C.c_float)]


def _sx(dtype, v):
    v = complex(v)
    return (CX if np.dtype(dtype) == np.complex128 else SX)(v.real, v.imag)


def _u(dtype):
    return 2.0 ** -53 if np.dtype(dtype) in (np.complex128, np.float64) else 2.0 ** -24


class ZMat:
    """Strided complex (or real) matrix in a flat buffer, mirrored on the GPU."""

    def __init__(self, m, n, rs=None, cs=None, dtype=np.complex128, seed=0, off=0):
        if rs is None and cs is None:
            rs, cs = 1, max(m, 1)
        self.m, self.n, self.rs, self.cs, self.off = m, n, rs, cs, off
        self.dtype = np.dtype(dtype)
        size = off + (max(m, 1) - 1) * rs + (max(n, 1) - 1) * cs + 1
        rng = np.random.default_rng(seed)
        buf = rng.uniform(-1, 1, size)
        if np.iscomplexobj(np.zeros(1, self.dtype)):
            buf = buf + 1j * rng.uniform(-1, 1, size)
        self.buf = buf.astype(self.dtype)
        self.dbuf = None

    def view(self, buf=None):
        b = self.buf if buf is None else buf
        es = self.dtype.itemsize
        return np.lib.stride_tricks.as_strided(b[self.off:], shape=(self.m, self.n),
                                               strides=(self.rs * es, self.cs * es))

    def to_device(self):
        import torch
        self.dbuf = torch.from_numpy(self.buf.copy()).to("cuda")
        return self

    @property
    def ptr(self):
        return self.dbuf.data_ptr() + self.off * self.dtype.itemsize

    @property
    def hptr(self):
        return self.buf.ctypes.data + self.off * self.dtype.itemsize

    def device_view(self):
        import torch
        torch.cuda.synchronize()
        return self.view(self.dbuf.cpu().numpy())


def _zerr(got, ref, a, b, c0, alpha, beta):
    num = np.linalg.norm((got - ref).astype(np.complex128))
    den = abs(alpha) * np.linalg.norm(a.astype(np.complex128)) * \
        np.linalg.norm(b.astype(np.complex128)) + abs(beta) * np.linalg.norm(
            np.nan_to_num(c0.astype(np.complex128)))
    return num / den


def _call_typed(blis, dtype, ta, tb, m, n, k, alpha, a, b, beta, c, host=False):
    ch = "z" if np.dtype(dtype) == np.complex128 else "c"
    p = (lambda x: x.hptr) if host else (lambda x: x.ptr)
    al, be = _sx(dtype, alpha), _sx(dtype, beta)
    getattr(blis, f"bli_{ch}gemm")(TRB[ta], TRB[tb], C.c_int64(m), C.c_int64(n), C.c_int64(k),
                                   C.byref(al), C.c_void_p(p(a)), C.c_int64(a.rs),
                                   C.c_int64(a.cs), C.c_void_p(p(b)), C.c_int64(b.rs),
                                   C.c_int64(b.cs), C.byref(be), C.c_void_p(p(c)),
                                   C.c_int64(c.rs), C.c_int64(c.cs))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("ta", ["n", "t", "r", "c"])
@pytest.mark.parametrize("tb", ["n", "t", "r", "c"])
def test_zgemm_all_trans(blis, oracle, dtype, ta, tb):
    u = _u(dtype)
    for (m, n, k) in [(1, 1, 1), (7, 5, 3), (67, 45, 33), (130, 97, 70)]:
        say(f"{np.dtype(dtype).name} gemm {ta}{tb} m={m} n={n} k={k}")
        a = ZMat(*((k, m) if ta in "tc" else (m, k)), dtype=dtype, seed=m + k)
        b = ZMat(*((n, k) if tb in "tc" else (k, n)), dtype=dtype, seed=3 * n + k)
        c = ZMat(m, n, dtype=dtype, seed=m * n + 1)
        c0 = c.view().copy()
        for x in (a, b, c):
            x.to_device()
        alpha, beta = 0.7 - 0.9j, 1.3 - 1.1j
        _call_typed(blis, dtype, ta, tb, m, n, k, alpha, a, b, beta, c)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.zgemm(ta, tb, alpha, a.view().copy(order="F"), b.view().copy(order="F"), beta,
                     ref)
        assert _zerr(got, ref, a.view(), b.view(), c0, alpha, beta) <= 8 * (k + 2) * u


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("case", ["row", "general", "host", "real_alpha", "beta0_nan",
                                  "real_beta", "alpha0"])
def test_zgemm_layouts_and_scalars(blis, oracle, dtype, case):
    u = _u(dtype)
    m, n, k = 90, 70, 50
    alpha, beta = 0.5 + 0.25j, -0.75 + 0.5j
    say(f"{np.dtype(dtype).name} gemm {case}")
    if case == "row":
        a = ZMat(m, k, rs=k + 1, cs=1, dtype=dtype, seed=1)
        b = ZMat(k, n, rs=n, cs=1, dtype=dtype, seed=2)
        c = ZMat(m, n, rs=n + 2, cs=1, dtype=dtype, seed=3)
    elif case == "general":
        a = ZMat(m, k, rs=2, cs=2 * m + 1, dtype=dtype, seed=1)
        b = ZMat(k, n, rs=3, cs=3 * k, dtype=dtype, seed=2)
        c = ZMat(m, n, rs=2, cs=2 * m + 4, dtype=dtype, seed=3)
    else:
        a = ZMat(m, k, dtype=dtype, seed=1)
        b = ZMat(k, n, dtype=dtype, seed=2)
        c = ZMat(m, n, dtype=dtype, seed=3)
    if case == "real_alpha":
        alpha = 1.5
    if case == "real_beta":
        beta = 0.5
    if case == "alpha0":
        alpha = 0.0
    if case == "beta0_nan":
        beta = 0.0
        c.buf[:] = np.nan
    c0 = c.view().copy()
    host = case == "host"
    if not host:
        for x in (a, b, c):
            x.to_device()
    _call_typed(blis, dtype, "n", "n", m, n, k, alpha, a, b, beta, c, host=host)
    got = c.view().copy() if host else c.device_view()
    ref = np.nan_to_num(c0).copy(order="F") if beta == 0 else c0.copy(order="F")
    oracle.zgemm("n", "n", alpha, a.view().copy(order="F"), b.view().copy(order="F"), beta, ref)
    assert np.isfinite(got).all()
    assert _zerr(got, ref, a.view(), b.view(), c0, alpha, beta) <= 8 * (k + 2) * u


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_zgemm_full_size_probe(blis, dtype):
    """2048^3 through the large real tiles: the probe residual."""
    import torch
    n = 2048
    u = _u(dtype)
    tdt = torch.complex128 if np.dtype(dtype) == np.complex128 else torch.complex64
    g = torch.Generator(device="cuda").manual_seed(11)

    def rnd():
        re = torch.rand((n, n), device="cuda", generator=g, dtype=torch.float64) * 2 - 1
        im = torch.rand((n, n), device="cuda", generator=g, dtype=torch.float64) * 2 - 1
        return torch.complex(re, im).to(tdt)

    a, b, c = rnd(), rnd(), rnd()   # column-major views: X(i, j) = X_t[j, i]
    c0 = c.clone()
    alpha, beta = 0.7 - 0.2j, 0.5 + 0.5j
    ch = "z" if np.dtype(dtype) == np.complex128 else "c"
    say(f"{ch}gemm {n}^3 probe residual")
    i64, vp = C.c_int64, C.c_void_p
    getattr(blis, f"bli_{ch}gemm")(0, 0, i64(n), i64(n), i64(n), C.byref(_sx(dtype, alpha)),
                                   vp(a.data_ptr()), i64(1), i64(n), vp(b.data_ptr()), i64(1),
                                   i64(n), C.byref(_sx(dtype, beta)), vp(c.data_ptr()), i64(1),
                                   i64(n))
    torch.cuda.synchronize()
    A, B = a.cpu().numpy().T.astype(np.complex128), b.cpu().numpy().T.astype(np.complex128)
    Cn, C0 = c.cpu().numpy().T.astype(np.complex128), c0.cpu().numpy().T.astype(np.complex128)
    x = np.random.default_rng(1).uniform(-1, 1, n) + 0j
    lhs = Cn @ x
    rhs = beta * (C0 @ x) + alpha * (A @ (B @ x))
    res = np.linalg.norm(lhs - rhs) / (abs(alpha) * np.linalg.norm(A) * np.linalg.norm(B) *
                                       np.linalg.norm(x) + abs(beta) * np.linalg.norm(C0) *
                                       np.linalg.norm(x))
    assert res <= 16 * n * u, res


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_zgemm_blas_object_batch_bitwise(blis, dtype):
    """?gemm_, the object API (conjugate-transpose from the obj_t bits) and
    ?gemm_batch_ give the typed call's bits."""
    ch = "z" if np.dtype(dtype) == np.complex128 else "c"
    m, n, k = 33, 29, 17
    alpha, beta = 0.3 + 0.8j, -0.6 + 0.1j
    a = ZMat(k, m, dtype=dtype, seed=5).to_device()
    b = ZMat(k, n, dtype=dtype, seed=6).to_device()
    c_ref = ZMat(m, n, dtype=dtype, seed=7).to_device()
    say(f"{ch}gemm typed C^H N")
    _call_typed(blis, dtype, "c", "n", m, n, k, alpha, a, b, beta, c_ref)
    ref = c_ref.device_view()

    say(f"{ch}gemm_ BLAS")
    c1 = ZMat(m, n, dtype=dtype, seed=7).to_device()
    f = C.c_int32
    fn = getattr(blis, f"{ch}gemm_")
    fn.argtypes = None
    fn(C.c_char_p(b"C"), C.c_char_p(b"N"), C.byref(f(m)), C.byref(f(n)), C.byref(f(k)),
       C.byref(_sx(dtype, alpha)), C.c_void_p(a.ptr), C.byref(f(a.cs)), C.c_void_p(b.ptr),
       C.byref(f(b.cs)), C.byref(_sx(dtype, beta)), C.c_void_p(c1.ptr), C.byref(f(c1.cs)))
    assert np.array_equal(c1.device_view(), ref)

    say(f"bli_gemm object {ch}")
    c2 = ZMat(m, n, dtype=dtype, seed=7).to_device()
    dtc = DT[np.dtype(dtype)]

    def obj(x, trans=0):
        o = blis_amd.obj_t()
        blis.bli_obj_create_with_attached_buffer(dtc, x.m, x.n, C.c_void_p(x.ptr), x.rs, x.cs,
                                                 C.byref(o))
        o.info |= trans
        return o

    al, be = _sx(dtype, alpha), _sx(dtype, beta)
    oal, obe = blis_amd.obj_t(), blis_amd.obj_t()
    blis.bli_obj_create_1x1_with_attached_buffer(dtc, C.byref(al), C.byref(oal))
    blis.bli_obj_create_1x1_with_attached_buffer(dtc, C.byref(be), C.byref(obe))
    blis.bli_gemm(C.byref(oal), C.byref(obj(a, 0x18)), C.byref(obj(b)), C.byref(obe),
                  C.byref(obj(c2)))
    assert np.array_equal(c2.device_view(), ref)

    say(f"{ch}gemm_batch_ two groups")
    cs = [ZMat(m, n, dtype=dtype, seed=7).to_device() for _ in range(3)]
    PA = (C.c_void_p * 3)(*[a.ptr] * 3)
    PB = (C.c_void_p * 3)(*[b.ptr] * 3)
    PC = (C.c_void_p * 3)(*[x.ptr for x in cs])
    S = CX if np.dtype(dtype) == np.complex128 else SX
    alphas = (S * 2)(_sx(dtype, alpha), _sx(dtype, alpha))
    betas = (S * 2)(_sx(dtype, beta), _sx(dtype, beta))
    I = f * 2
    bf = getattr(blis, f"{ch}gemm_batch_")
    bf.argtypes = None
    bf((C.c_char * 2)(b"C", b"C"), (C.c_char * 2)(b"N", b"N"), I(m, m), I(n, n), I(k, k), alphas,
       PA, I(a.cs, a.cs), PB, I(b.cs, b.cs), betas, PC, I(m, m), C.byref(f(2)), I(1, 2))
    for x in cs:
        assert np.array_equal(x.device_view(), ref)


MD_CASES = [
    # (A dtype, B dtype, C dtype, computation in double)
    (np.float32, np.float32, np.float64, True),     # single operands, double C and compute
    (np.float64, np.float64, np.float32, True),     # double compute, single C
    (np.float64, np.float32, np.float64, False),    # single-precision compute
    (np.float64, np.float64, np.float64, False),    # all double, single compute
    (np.complex128, np.complex128, np.float64, True),   # R += C * C (real part kept)
    (np.float64, np.float64, np.complex128, True),      # C += R * R (real parts only)
    (np.float64, np.complex128, np.complex128, True),   # C += R * C
    (np.complex128, np.float64, np.complex128, True),   # C += C * R
    (np.float64, np.complex128, np.float64, True),      # R += R * C (real part of B)
    (np.complex64, np.complex128, np.complex64, True),  # complex, mixed precision
]


@pytest.mark.parametrize("da,db,dc,comp_d", MD_CASES)
def test_gemm_mixed_datatypes(blis, oracle, da, db, dc, comp_d):
    """bli_gemm with operands of different datatypes (docs/MixedDatatypes.md)
    against the oracle's restatement of the rules."""
    m, n, k = 45, 38, 29
    say(f"bli_gemm md A={np.dtype(da).name} B={np.dtype(db).name} C={np.dtype(dc).name} "
        f"comp={'d' if comp_d else 's'}")
    a = ZMat(k, m, dtype=da, seed=1).to_device()      # used transposed
    b = ZMat(k, n, dtype=db, seed=2).to_device()
    c = ZMat(m, n, dtype=dc, seed=3).to_device()
    c0 = c.view().copy()

    def obj(x, trans=0):
        o = blis_amd.obj_t()
        blis.bli_obj_create_with_attached_buffer(DT[x.dtype], x.m, x.n, C.c_void_p(x.ptr),
                                                 x.rs, x.cs, C.byref(o))
        o.info |= trans
        return o

    oa, ob, oc = obj(a, 0x8), obj(b), obj(c)
    oc.info = (oc.info & ~(0x3 << 22)) | ((2 if comp_d else 0) << 22)
    al, be = CX(0.75, 0.25), CX(-0.5, 0.0)
    oal, obe = blis_amd.obj_t(), blis_amd.obj_t()
    blis.bli_obj_create_1x1_with_attached_buffer(3, C.byref(al), C.byref(oal))
    blis.bli_obj_create_1x1_with_attached_buffer(3, C.byref(be), C.byref(obe))
    blis.bli_gemm(C.byref(oal), C.byref(oa), C.byref(ob), C.byref(obe), C.byref(oc))
    got = c.device_view()
    comp_cx = (np.iscomplexobj(a.buf) and np.iscomplexobj(b.buf)) or \
        (np.iscomplexobj(c.buf) and (np.iscomplexobj(a.buf) or np.iscomplexobj(b.buf)))
    alpha = complex(0.75, 0.25) if comp_cx else 0.75
    ref = oracle.gemm_md("t", "n", alpha, a.view().copy(order="F"), b.view().copy(order="F"),
                         -0.5, c0.copy(order="F"), comp_d)
    u = 2.0 ** -53 if comp_d else 2.0 ** -24
    # The result is rounded to C's storage precision at the end.
    u = max(u, 2.0 ** -24 if np.dtype(dc) in (np.float32, np.complex64) else 2.0 ** -53)
    scale = np.linalg.norm(a.view().astype(np.complex128)) * \
        np.linalg.norm(b.view().astype(np.complex128)) + np.linalg.norm(c0.astype(np.complex128))
    err = np.linalg.norm((got - ref).astype(np.complex128)) / scale
    assert err <= 8 * (k + 2) * u, err
    if np.iscomplexobj(c.buf) and not comp_cx:
        assert np.array_equal(got.imag, c0.imag), "imaginary parts must stay untouched"
