"""GEMM parity on the GPU through the C ABI (bli_?gemm, dgemm_/sgemm_,
bli_gemm object API) against the CPU oracle.

Bar: normwise error ||C_gpu - C_oracle||_F / (|alpha| ||A||_F ||B||_F +
|beta| ||C0||_F) <= 4 k u, with u = 2^-53 (d) or 2^-24 (s) — the c*k*u bound
the GEMM error analysis gives for either result, doubled for the difference
of two such results. At full size (8192) the oracle is too slow, so the
size-independent testsuite probe residual (testsuite/src/test_gemm.c:
303-414) is used instead, against the testsuite's thresholds.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import Mat, normwise_gemm_err, say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}


def _ct(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def run_gemm(blis, dtype, ta, tb, m, n, k, alpha, beta, a: Mat, b: Mat, c: Mat):
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    T = _ct(dtype)
    al, be = T(alpha), T(beta)
    getattr(blis, f"bli_{ch}gemm")(TR[ta], TR[tb], m, n, k, C.byref(al), a.ptr, a.rs, a.cs,
                                   b.ptr, b.rs, b.cs, C.byref(be), c.ptr, c.rs, c.cs)


def make_operands(dtype, ta, tb, m, n, k, seed, layout="col"):
    am, an = (k, m) if ta == "t" else (m, k)
    bm, bn = (n, k) if tb == "t" else (k, n)
    if layout == "col":
        a = Mat(am, an, dtype=dtype, seed=seed)
        b = Mat(bm, bn, dtype=dtype, seed=seed + 1)
        c = Mat(m, n, dtype=dtype, seed=seed + 2)
    elif layout == "row":
        a = Mat(am, an, rs=an, cs=1, dtype=dtype, seed=seed)
        b = Mat(bm, bn, rs=bn, cs=1, dtype=dtype, seed=seed + 1)
        c = Mat(m, n, rs=n, cs=1, dtype=dtype, seed=seed + 2)
    elif layout == "padded":
        # lda/ldb/ldc padded and odd offsets (misaligned for 16-byte vectors)
        a = Mat(am, an, rs=1, cs=am + 3, dtype=dtype, seed=seed, off=1)
        b = Mat(bm, bn, rs=1, cs=bm + 5, dtype=dtype, seed=seed + 1, off=3)
        c = Mat(m, n, rs=1, cs=m + 1, dtype=dtype, seed=seed + 2, off=1)
    else:  # general strides: neither unit row nor unit column stride
        a = Mat(am, an, rs=2, cs=2 * am + 1, dtype=dtype, seed=seed)
        b = Mat(bm, bn, rs=bn + 2, cs=3, dtype=dtype, seed=seed + 1)
        c = Mat(m, n, rs=3, cs=3 * m + 2, dtype=dtype, seed=seed + 2)
    return a, b, c


def check_parity(blis, oracle, dtype, ta, tb, m, n, k, alpha, beta, layout="col", seed=1):
    say(f"gemm {np.dtype(dtype).name} {ta}{tb} m={m} n={n} k={k} alpha={alpha} beta={beta} {layout}")
    a, b, c = make_operands(dtype, ta, tb, m, n, k, seed, layout)
    c0 = c.view().copy()
    for x in (a, b, c):
        x.to_device()
    run_gemm(blis, dtype, ta, tb, m, n, k, alpha, beta, a, b, c)
    got = c.device_view()
    ref = c0.copy(order="F")
    oracle.gemm(ta, tb, alpha, a.view(), b.view(), beta, ref)
    err = normwise_gemm_err(got, ref, a.view(), b.view(), c0, alpha, beta)
    assert err <= 4 * max(k, 1) * unit_roundoff(dtype), (ta, tb, m, n, k, layout, err)
    return err


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("tb", ["n", "t"])
def test_gemm_trans_edges(blis, oracle, dtype, ta, tb):
    for (m, n, k) in [(1, 1, 1), (7, 5, 3), (67, 45, 33), (128, 128, 128), (257, 130, 70),
                      (33, 300, 17)]:
        check_parity(blis, oracle, dtype, ta, tb, m, n, k, 0.7, 1.3)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layout", ["row", "padded", "general"])
def test_gemm_layouts(blis, oracle, dtype, layout):
    for ta, tb in [("n", "n"), ("t", "n"), ("n", "t"), ("t", "t")]:
        check_parity(blis, oracle, dtype, ta, tb, 77, 61, 45, -1.5, 0.25, layout=layout)


def test_gemm_config1_dgemm_f77(blis, oracle):
    """Config 1: DGEMM NN column-major m=n=k=512, alpha=1, beta=0.5, through
    the BLAS symbol dgemm_ with device pointers."""
    m = n = k = 512
    say("dgemm_ config1 512^3")
    a = Mat(m, k, seed=100).to_device()
    b = Mat(k, n, seed=101).to_device()
    c = Mat(m, n, seed=102)
    c0 = c.view().copy()
    c.to_device()
    mi, ni, ki, ld = (C.c_int32(v) for v in (m, n, k, 512))
    al, be = C.c_double(1.0), C.c_double(0.5)
    blis.dgemm_(b"N", b"N", C.byref(mi), C.byref(ni), C.byref(ki), C.byref(al), a.ptr,
                C.byref(ld), b.ptr, C.byref(ld), C.byref(be), c.ptr, C.byref(ld))
    got = c.device_view()
    ref = c0.copy(order="F")
    oracle.gemm("n", "n", 1.0, a.view(), b.view(), 0.5, ref)
    err = normwise_gemm_err(got, ref, a.view(), b.view(), c0, 1.0, 0.5)
    say(f"config1 normwise err {err:.3e}")
    assert err <= 4 * k * unit_roundoff(np.float64)


def test_gemm_host_pointers(blis, oracle):
    """Host (non-device) operands are staged through device memory."""
    m, n, k = 90, 70, 50
    say("dgemm host pointers")
    a, b, c = Mat(m, k, seed=5), Mat(k, n, seed=6), Mat(m, n, seed=7)
    c0 = c.view().copy()
    al, be = C.c_double(1.25), C.c_double(-0.5)
    blis.bli_dgemm(0, 0, m, n, k, C.byref(al), a.hptr, a.rs, a.cs, b.hptr, b.rs, b.cs,
                   C.byref(be), c.hptr, c.rs, c.cs)
    ref = c0.copy(order="F")
    oracle.gemm("n", "n", 1.25, a.view(), b.view(), -0.5, ref)
    assert normwise_gemm_err(c.view(), ref, a.view(), b.view(), c0, 1.25, -0.5) <= \
        4 * k * unit_roundoff(np.float64)


def test_gemm_beta_zero_nan_and_alpha_zero(blis):
    m, n, k = 40, 30, 20
    say("dgemm beta=0 over NaN, alpha=0 with NaN operands")
    a, b = Mat(m, k, seed=1).to_device(), Mat(k, n, seed=2).to_device()
    c = Mat(m, n, fill=np.nan).to_device()
    al, be = C.c_double(1.0), C.c_double(0.0)
    blis.bli_dgemm(0, 0, m, n, k, C.byref(al), a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs,
                   C.byref(be), c.ptr, c.rs, c.cs)
    got = c.device_view()
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, a.view() @ b.view(), rtol=1e-13, atol=1e-13)
    # alpha == 0 must not touch A or B: C := beta * C
    an = Mat(m, k, fill=np.nan).to_device()
    bn = Mat(k, n, fill=np.nan).to_device()
    c2 = Mat(m, n, seed=9)
    c20 = c2.view().copy()
    c2.to_device()
    al, be = C.c_double(0.0), C.c_double(3.0)
    blis.bli_dgemm(0, 0, m, n, k, C.byref(al), an.ptr, an.rs, an.cs, bn.ptr, bn.rs, bn.cs,
                   C.byref(be), c2.ptr, c2.rs, c2.cs)
    np.testing.assert_array_equal(c2.device_view(), 3.0 * c20)
    # k == 0: C := beta * C
    blis.bli_dgemm(0, 0, m, n, 0, C.byref(C.c_double(1.0)), an.ptr, an.rs, an.cs, bn.ptr,
                   bn.rs, bn.cs, C.byref(C.c_double(0.5)), c2.ptr, c2.rs, c2.cs)
    np.testing.assert_array_equal(c2.device_view(), 1.5 * c20)


def test_gemm_object_api(blis, oracle):
    import blis_amd
    m, n, k = 65, 66, 67
    say("bli_gemm object API")
    a, b, c = Mat(k, m, seed=21), Mat(k, n, seed=22), Mat(m, n, seed=23)
    c0 = c.view().copy()
    for x in (a, b, c):
        x.to_device()
    ao, bo, co, alo, beo = (blis_amd.obj_t() for _ in range(5))
    blis.bli_obj_create_with_attached_buffer(blis_amd.BLIS_DOUBLE, k, m, a.ptr, a.rs, a.cs, C.byref(ao))
    blis.bli_obj_create_with_attached_buffer(blis_amd.BLIS_DOUBLE, k, n, b.ptr, b.rs, b.cs, C.byref(bo))
    blis.bli_obj_create_with_attached_buffer(blis_amd.BLIS_DOUBLE, m, n, c.ptr, c.rs, c.cs, C.byref(co))
    alpha, beta = C.c_double(2.0), C.c_double(1.2)
    blis.bli_obj_create_1x1_with_attached_buffer(blis_amd.BLIS_DOUBLE, C.byref(alpha), C.byref(alo))
    blis.bli_obj_create_1x1_with_attached_buffer(blis_amd.BLIS_DOUBLE, C.byref(beta), C.byref(beo))
    ao.info = (ao.info & ~0x18) | blis_amd.BLIS_TRANSPOSE      # op(A) = A^T
    blis.bli_gemm(C.byref(alo), C.byref(ao), C.byref(bo), C.byref(beo), C.byref(co))
    got = c.device_view()
    ref = c0.copy(order="F")
    oracle.gemm("t", "n", 2.0, a.view(), b.view(), 1.2, ref)
    assert normwise_gemm_err(got, ref, a.view(), b.view(), c0, 2.0, 1.2) <= 4 * k * unit_roundoff(np.float64)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
def test_gemm_every_tile(blis, oracle, tile):
    """Every entry of the tile table computes the same product."""
    try:
        for dtype in (np.float64, np.float32):
            blis.bli_amd_gemm_set_config(tile, 8, 1, 1)
            for ta, tb in [("n", "n"), ("t", "t")]:
                check_parity(blis, oracle, dtype, ta, tb, 300, 200, 100, 1.0, 0.5)
    finally:
        blis.bli_amd_gemm_set_config(-1, 8, 1, 1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layout", ["col", "row", "padded", "general"])
def test_gemm_split_k(blis, oracle, dtype, layout):
    """k partitions into workspace plus one ordered reduction: the same
    product as one pass, for ragged k, every transpose pair and every
    storage of C."""
    blis.bli_amd_gemm_set_config.argtypes = [C.c_int] * 4
    try:
        # tile -1: the heuristic picks the tile, the forced split still holds
        for tile, split, (m, n, k) in [(4, 3, (67, 45, 333)), (3, 8, (130, 65, 1000)),
                                       (8, 2, (5, 129, 31)), (1, 64, (40, 30, 20)),
                                       (-1, 5, (90, 70, 640))]:
            blis.bli_amd_gemm_set_config(tile, 8, 1, split)
            for ta, tb in [("n", "n"), ("t", "n"), ("n", "t"), ("t", "t")]:
                check_parity(blis, oracle, dtype, ta, tb, m, n, k, -0.5, 0.75, layout)
    finally:
        blis.bli_amd_gemm_set_config(-1, 8, 1, 1)


def test_gemm_split_k_beta_zero_over_nan(blis):
    """The reduction stores C without reading it when beta == 0."""
    blis.bli_amd_gemm_set_config.argtypes = [C.c_int] * 4
    m, n, k = 70, 33, 500
    say("dgemm split-k beta=0 over NaN")
    a, b = Mat(m, k, seed=5).to_device(), Mat(k, n, seed=6).to_device()
    c = Mat(m, n, fill=np.nan).to_device()
    al, be = C.c_double(1.0), C.c_double(0.0)
    try:
        blis.bli_amd_gemm_set_config(3, 8, 1, 4)
        blis.bli_dgemm(0, 0, m, n, k, C.byref(al), a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs,
                       C.byref(be), c.ptr, c.rs, c.cs)
    finally:
        blis.bli_amd_gemm_set_config(-1, 8, 1, 1)
    got = c.device_view()
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, a.view() @ b.view(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemm_default_split_for_deep_products(blis, oracle, dtype):
    """A small output over a long k (m, n <= k/4, k >= 2048) is split by the
    built-in rule (here into 8 partitions); the result matches the oracle,
    and every tile given the same split gives bitwise the same result."""
    blis.bli_amd_gemm_set_config.argtypes = [C.c_int] * 4
    m, n, k = 64, 48, 8192
    check_parity(blis, oracle, dtype, "t", "n", m, n, k, 1.25, -0.5)
    a, b, c = make_operands(dtype, "n", "t", m, n, k, 31)
    for x in (a, b):
        x.to_device()
    outs = []
    try:
        for tile in (-1, 0, 1, 2, 3, 4):
            blis.bli_amd_gemm_set_config(tile, 8, 1, 8 if tile >= 0 else 1)
            cc = Mat(m, n, dtype=dtype, seed=33).to_device()
            run_gemm(blis, dtype, "n", "t", m, n, k, 1.0, 0.5, a, b, cc)
            outs.append(cc.device_view().copy())
    finally:
        blis.bli_amd_gemm_set_config(-1, 8, 1, 1)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def probe_resid(a, b, c_hat, c0, alpha, beta, ta, tb, seed=0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(-1, 1, c_hat.shape[1])
    A = a.T if ta == "t" else a
    B = b.T if tb == "t" else b
    v = c_hat.astype(np.float64) @ t
    z = alpha * (A.astype(np.float64) @ (B.astype(np.float64) @ t)) + beta * (c0.astype(np.float64) @ t)
    scale = (abs(alpha) * np.linalg.norm(A.astype(np.float64)) * np.linalg.norm(B.astype(np.float64))
             + abs(beta) * np.linalg.norm(c0.astype(np.float64))) * np.linalg.norm(t)
    return np.linalg.norm(v - z) / scale


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(4096, 4096, 4096), (8192, 8192, 256)])
def test_gemm_large_probe_residual(blis, dtype, shape):
    """Full-size parity via the size-independent probe residual."""
    import torch
    m, n, k = shape
    for ta, tb in [("n", "n"), ("t", "n"), ("n", "t"), ("t", "t")]:
        say(f"large gemm {np.dtype(dtype).name} {ta}{tb} {shape}")
        a, b, c = make_operands(dtype, ta, tb, m, n, k, seed=31)
        c0 = c.view().copy()
        for x in (a, b, c):
            x.to_device()
        run_gemm(blis, dtype, ta, tb, m, n, k, 1.0, 0.5, a, b, c)
        r = probe_resid(a.view(), b.view(), c.device_view(), c0, 1.0, 0.5, ta, tb)
        assert r <= 4 * k * unit_roundoff(dtype), (ta, tb, r)
        del a, b, c
        torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemm_k_contiguous_operands_match_rows_contiguous(blis, dtype):
    """A k-contiguous operand (op(A) = A^T, op(B) = B) and a rows-contiguous
    one must give bit-identical products, at ragged sizes and padded leading
    dimensions, close to an FP64 product on sampled rows. For FP64 the
    128 x 128 tile of sixteen 32 x 32 waves is pinned: on it, large products
    move a k-contiguous operand to the rows-contiguous layout with one
    transpose (64 x 64 tiles with partial edges here) before the tile kernel
    runs."""
    import torch
    m, n, k = 6000, 4100, 3001
    if np.dtype(dtype) == np.float64:
        blis.bli_amd_gemm_set_config.argtypes = [C.c_int] * 4
        blis.bli_amd_gemm_set_config(13, 8, 1, 1)
    try:
        _k_contiguous_case(blis, dtype, m, n, k)
    finally:
        blis.bli_amd_gemm_set_config(-1, 8, 1, 1)


def _k_contiguous_case(blis, dtype, m, n, k):
    import torch
    dev = torch.device("cuda")
    g = torch.Generator(device="cpu").manual_seed(77)
    A = (torch.rand(m, k, generator=g, dtype=torch.float64) * 2 - 1).to(dtype=_torch_dt(dtype))
    B = (torch.rand(k, n, generator=g, dtype=torch.float64) * 2 - 1).to(dtype=_torch_dt(dtype))
    C0 = (torch.rand(m, n, generator=g, dtype=torch.float64) * 2 - 1).to(dtype=_torch_dt(dtype))

    def colmajor(x, pad):
        # column-major storage with leading dimension rows + pad
        buf = torch.zeros(x.shape[1], x.shape[0] + pad, dtype=x.dtype)
        buf[:, :x.shape[0]] = x.t()
        return buf.to(dev), x.shape[0] + pad

    T = _ct(dtype)
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    al, be = T(0.75), T(-0.5)
    outs = {}
    for case in ("nn", "tn", "nt"):
        say(f"k-contiguous layouts {ch} {case} {m}x{n}x{k}")
        if case[0] == "n":
            a, lda = colmajor(A, 5)
        else:
            a, lda = colmajor(A.t().contiguous(), 5)
        if case[1] == "n":
            b, ldb = colmajor(B, 3)
        else:
            b, ldb = colmajor(B.t().contiguous(), 3)
        c, ldc = colmajor(C0, 0)
        getattr(blis, f"bli_{ch}gemm")(TR[case[0]], TR[case[1]], m, n, k, C.byref(al),
                                       a.data_ptr(), 1, lda, b.data_ptr(), 1, ldb, C.byref(be),
                                       c.data_ptr(), 1, ldc)
        torch.cuda.synchronize()
        outs[case] = c[:, :m].cpu().t().numpy().copy()
        del a, b, c
    assert np.array_equal(outs["nn"], outs["nt"])
    assert np.array_equal(outs["tn"], outs["nt"])
    rows = np.arange(0, m, 97)
    An, Bn, Cn = (x.double().numpy() for x in (A, B, C0))
    ref = 0.75 * (An[rows] @ Bn) - 0.5 * Cn[rows]
    err = np.linalg.norm(outs["nn"][rows] - ref) / (
        0.75 * np.linalg.norm(An[rows]) * np.linalg.norm(Bn) + 0.5 * np.linalg.norm(Cn[rows]))
    assert err <= 4 * k * unit_roundoff(dtype), err
    torch.cuda.empty_cache()


def _torch_dt(dtype):
    import torch
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
