"""The single-call small-problem path (sup.hip; BLIS's sup, frame/3/
bli_l3_sup.c:37-130): GEMMs with m, n, k <= 512 and no configuration given.

* Against the oracle for every transpose pair, edge shapes and both types
  (normwise 4 (k + 2) u), including config 1 (512^3, alpha 1, beta 0.5).
* Bit for bit equal to the table's 32 x 32 tile (forced with
  bli_amd_gemm_set_config): the sup kernel runs the same MFMA k-steps in the
  same order with deeper k slabs.
* beta == 0 never reads C (NaN C stays out of the result).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import Mat, normwise_gemm_err, say, unit_roundoff

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (5, 7, 3), (31, 33, 65), (64, 64, 64), (100, 37, 200), (257, 129, 500),
          (512, 512, 512)]


def _gemm(blis, dtype, ta, tb, m, n, k, alpha, a, b, beta, c):
    ch = "d" if dtype == np.float64 else "s"
    T = C.c_double if dtype == np.float64 else C.c_float
    getattr(blis, f"bli_{ch}gemm")(8 if ta else 0, 8 if tb else 0, m, n, k, C.byref(T(alpha)),
                                   a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs, C.byref(T(beta)), c.ptr,
                                   c.rs, c.cs)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
def test_sup_vs_oracle_and_table_tile(blis, oracle, dtype, ta, tb):
    u = unit_roundoff(dtype)
    for (m, n, k) in SHAPES:
        say(f"sup {np.dtype(dtype).name} ta={ta} tb={tb} {m}x{n}x{k}")
        a = Mat(*((k, m) if ta else (m, k)), dtype=dtype, seed=m + k).to_device()
        b = Mat(*((n, k) if tb else (k, n)), dtype=dtype, seed=n + 3 * k).to_device()
        c = Mat(m, n, dtype=dtype, seed=m * n).to_device()
        c2 = Mat(m, n, dtype=dtype, seed=m * n).to_device()
        c0 = c.view().copy()
        _gemm(blis, dtype, ta, tb, m, n, k, 1.0, a, b, 0.5, c)
        got = c.device_view()
        ref = c0.copy(order="F")
        oracle.gemm("t" if ta else "n", "t" if tb else "n", 1.0, a.view().copy(order="F"),
                    b.view().copy(order="F"), 0.5, ref)
        assert normwise_gemm_err(got, ref, a.view(), b.view(), c0, 1.0, 0.5) <= 4 * (k + 2) * u
        # The same product on the table's 32 x 32 tile: the same bits.
        blis.bli_amd_gemm_set_config(4, 8, 1, 1)
        try:
            _gemm(blis, dtype, ta, tb, m, n, k, 1.0, a, b, 0.5, c2)
        finally:
            blis.bli_amd_gemm_set_config(-1, 8, 1, 1)
        assert np.array_equal(got, c2.device_view())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sup_beta_zero_ignores_nan(blis, dtype):
    m, n, k = 70, 45, 90
    say(f"sup beta=0 NaN C {np.dtype(dtype).name}")
    a = Mat(m, k, dtype=dtype, seed=1).to_device()
    b = Mat(k, n, dtype=dtype, seed=2).to_device()
    c = Mat(m, n, dtype=dtype, seed=3, fill=np.nan).to_device()
    _gemm(blis, dtype, 0, 0, m, n, k, 2.0, a, b, 0.0, c)
    got = c.device_view()
    ref = 2.0 * a.view().astype(np.float64) @ b.view().astype(np.float64)
    assert np.isfinite(got).all()
    assert np.allclose(got, ref, rtol=0, atol=8 * k * unit_roundoff(dtype) * 4)
