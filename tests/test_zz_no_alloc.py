"""bli_amd_gemm_noalloc: GEMM whose scratch comes from the library's fixed
per-GPU array, with no device allocation. Through the C ABI it must give the
bits of bli_dgemm / bli_sgemm (same kernels, same default configuration) on
two calls in a row, the second smaller, with shapes that use the scratch:
split-k partitions (small outputs over a long k) and a packed
general-stride operand. Oversized requests are refused before any work.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import Mat, say

pytestmark = pytest.mark.gpu

BLIS_SUCCESS = -1
BLIS_INSUFFICIENT_STACK_BUF_SIZE = -132
BLIS_FAILURE = -2


def _run_pair(blis, dtype, m, n, k, b_general, seed):
    import torch
    ch = "d" if dtype == np.float64 else "s"
    T = C.c_double if dtype == np.float64 else C.c_float
    num_t = 2 if dtype == np.float64 else 0
    a = Mat(m, k, dtype=dtype, seed=seed).to_device()
    if b_general:
        b = Mat(k, n, rs=2, cs=2 * k + 1, dtype=dtype, seed=seed + 1).to_device()
    else:
        b = Mat(k, n, dtype=dtype, seed=seed + 1).to_device()
    c1 = Mat(m, n, dtype=dtype, seed=seed + 2).to_device()
    c2 = Mat(m, n, dtype=dtype, seed=seed + 2).to_device()
    al, be = T(0.75), T(-0.5)
    say(f"{ch}gemm {m}x{n}x{k} general_b={b_general}: bli_{ch}gemm")
    getattr(blis, f"bli_{ch}gemm")(0, 0, m, n, k, C.byref(al), a.ptr, a.rs, a.cs, b.ptr, b.rs,
                                   b.cs, C.byref(be), c1.ptr, c1.rs, c1.cs)
    say(f"{ch}gemm {m}x{n}x{k} general_b={b_general}: bli_amd_gemm_noalloc")
    stream = torch.cuda.current_stream().cuda_stream
    rc = blis.bli_amd_gemm_noalloc(num_t, 0, 0, m, n, k, C.byref(al), a.ptr, a.rs, a.cs, b.ptr,
                                   b.rs, b.cs, C.byref(be), c2.ptr, c2.rs, c2.cs, stream)
    assert rc == BLIS_SUCCESS
    torch.cuda.synchronize()
    assert np.array_equal(c1.device_view(), c2.device_view())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_gemm_twice(blis, dtype):
    # Split-k: 4*max(m, n) <= k and k >= 2048 -> partitions in scratch.
    _run_pair(blis, dtype, 160, 96, 8192, b_general=True, seed=1)
    # Second call, smaller, again through the scratch (fewer partitions).
    _run_pair(blis, dtype, 64, 40, 4096, b_general=True, seed=11)
    # A plain product (no scratch) and alpha == 0 / empty shapes.
    _run_pair(blis, dtype, 300, 200, 100, b_general=False, seed=21)


def test_noalloc_rejects_oversized_and_host_operands(blis):
    import torch
    al, be = C.c_double(1.0), C.c_double(0.0)
    # A general-stride A of 4096 x 4096 doubles needs a 128 MiB packed copy:
    # refused before anything runs. The buffers are sized for the whole
    # problem all the same.
    n, k = 4096, 4096
    lda = 2 * n + 1
    a = torch.zeros(2 * (n - 1) + (k - 1) * lda + 1, dtype=torch.float64, device="cuda")
    b = torch.zeros(k * 64, dtype=torch.float64, device="cuda")
    c = torch.zeros(n * 64, dtype=torch.float64, device="cuda")
    say("noalloc oversized request")
    rc = blis.bli_amd_gemm_noalloc(2, 0, 0, n, 64, k, C.byref(al), a.data_ptr(), 2, lda,
                                   b.data_ptr(), 1, k, C.byref(be), c.data_ptr(), 1, n, None)
    assert rc == BLIS_INSUFFICIENT_STACK_BUF_SIZE
    torch.cuda.synchronize()
    assert float(c.abs().sum()) == 0.0
    say("noalloc host operands")
    h = np.zeros(16)
    rc = blis.bli_amd_gemm_noalloc(2, 0, 0, 4, 4, 4, C.byref(al), h.ctypes.data, 1, 4,
                                   h.ctypes.data, 1, 4, C.byref(be), h.ctypes.data, 1, 4, None)
    assert rc == BLIS_FAILURE
