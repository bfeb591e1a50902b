"""The registered gemm microkernels (bli_?gemm_ukernel, the BLIS_GEMM_UKR
entries of bli_gks_query_cntx) on the GPU: BLIS's microkernel contract
(frame/3/bli_l3_ukr_params.h:38-48, ref_kernels/3/bli_gemm_ref.c:163-317) on
packed micro-panels -- A an MR x k micro-panel stored column-major with ld
MR, B a k x NR micro-panel stored row-major with ld NR -- with edge sizes
m <= MR, n <= NR, any C strides, and beta == 0 not reading C. Checked
against the oracle on the unpacked operands (normwise 4 (k + 2) u).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import say, unit_roundoff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,k", [(64, 32, 16), (5, 3, 7), (64, 64, 200), (1, 1, 1)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_gemm_ukernel_packed_panels(blis, oracle, dtype, m, n, k, beta):
    import torch
    ch = "d" if dtype == np.float64 else "s"
    T = C.c_double if dtype == np.float64 else C.c_float
    num = 2 if dtype == np.float64 else 0
    cntx = blis.bli_gks_query_cntx
    cntx.restype = C.c_void_p
    ctx = cntx()
    q = blis.bli_cntx_get_blksz_def_dt
    q.restype = C.c_int64
    q.argtypes = [C.c_int, C.c_int, C.c_void_p]
    mr, nr = q(num, 1, ctx), q(num, 2, ctx)
    if m > mr or n > nr:
        pytest.skip("shape beyond the registered MR x NR")
    say(f"bli_{ch}gemm_ukernel m={m} n={n} k={k} beta={beta} (MR {mr}, NR {nr})")
    rng = np.random.default_rng(m * 100 + n * 10 + k)
    a = rng.uniform(-1, 1, (m, k)).astype(dtype)
    b = rng.uniform(-1, 1, (k, n)).astype(dtype)
    ap = np.zeros((k, mr), dtype)            # column-major micro-panel: a[i + p*MR]
    ap[:, :m] = a.T
    bp = np.zeros((k, nr), dtype)            # row-major micro-panel: b[p*NR + j]
    bp[:, :n] = b
    c = rng.uniform(-1, 1, (m, n)).astype(dtype)
    if beta == 0.0:
        c[:] = np.nan
    rs_c, cs_c = 3 * n + 1, 3          # C with general strides
    cbuf = np.full((m - 1) * rs_c + (n - 1) * cs_c + 1, 7.0, dtype)
    cview = np.lib.stride_tricks.as_strided(cbuf, (m, n), (rs_c * cbuf.itemsize,
                                                             cs_c * cbuf.itemsize))
    cview[...] = c
    dA = torch.from_numpy(ap.ravel()).to("cuda")
    dB = torch.from_numpy(bp.ravel()).to("cuda")
    dC = torch.from_numpy(cbuf.copy()).to("cuda")
    fn = getattr(blis, f"bli_{ch}gemm_ukernel")
    fn.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
    al, be = T(0.75), T(beta)
    fn(m, n, k, C.addressof(al), dA.data_ptr(), dB.data_ptr(), C.addressof(be), dC.data_ptr(),
       rs_c, cs_c, None, ctx)
    torch.cuda.synchronize()
    out = dC.cpu().numpy()
    got = np.lib.stride_tricks.as_strided(out, (m, n), (rs_c * out.itemsize,
                                                         cs_c * out.itemsize)).copy()
    ref = np.zeros((m, n), dtype, order="F") if beta == 0 else c.copy(order="F")
    oracle.gemm("n", "n", 0.75, np.asfortranarray(a), np.asfortranarray(b), beta, ref)
    u = unit_roundoff(dtype)
    scale = 0.75 * np.linalg.norm(a) * np.linalg.norm(b) + beta * np.linalg.norm(np.nan_to_num(c))
    assert np.isfinite(got).all()
    assert np.linalg.norm(got - ref) / scale <= 4 * (k + 2) * u
    # elements of the buffer outside C are untouched
    mask = np.ones(out.shape, bool)
    idx = (np.arange(m)[:, None] * rs_c + np.arange(n)[None, :] * cs_c).ravel()
    mask[idx] = False
    assert (out[mask] == 7.0).all()
