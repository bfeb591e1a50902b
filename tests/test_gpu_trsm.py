"""TRSM parity on the GPU through the C ABI against the CPU oracle.

Bar: ||X_gpu - X_oracle||_F / ||X_oracle||_F <= 8 m u for the small cases
(both solves are backward stable and the test matrices are well
conditioned), plus the residual ||op(A) X - alpha B|| / (||A|| ||X|| n u)
<= 8 at the full config size (BLIS testsuite style, testsuite/src/
test_trsm.c:295-389) where the oracle would be too slow.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import Mat, say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}
LOWER, UPPER = 0xC0, 0x60


def tri_matrix(mn, uplo, unit, dtype, seed, diag_shift=None):
    """Well-conditioned triangle; the unused triangle (and a unit diagonal) is
    NaN so any read of it shows up."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (mn, mn)) / max(mn, 1)
    a[np.diag_indices(mn)] = (1.0 if diag_shift is None else diag_shift) + rng.uniform(0, 1, mn)
    full = a.copy()
    if uplo == LOWER:
        a[np.triu_indices(mn, 1)] = np.nan
        full = np.tril(full)
    else:
        a[np.tril_indices(mn, -1)] = np.nan
        full = np.triu(full)
    if unit:
        a[np.diag_indices(mn)] = np.nan
        np.fill_diagonal(full, 1.0)
    m = Mat(mn, mn, dtype=dtype, seed=seed)
    m.set(a.astype(dtype))
    return m, full


def run_trsm(blis, dtype, side, uplo, ta, diag, m, n, alpha, a: Mat, b: Mat):
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    T = C.c_double if ch == "d" else C.c_float
    getattr(blis, f"bli_{ch}trsm")(side, uplo, TR[ta], diag, m, n, C.byref(T(alpha)), a.ptr,
                                   a.rs, a.cs, b.ptr, b.rs, b.cs)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("diag", [0, 0x100])
def test_trsm_all_cases(blis, oracle, dtype, side, uplo, ta, diag):
    u = unit_roundoff(dtype)
    for (m, n) in [(1, 1), (7, 5), (67, 45), (200, 130)]:
        say(f"trsm {np.dtype(dtype).name} side={side} uplo={uplo:#x} {ta} diag={diag:#x} m={m} n={n}")
        mn = m if side == 0 else n
        a, _ = tri_matrix(mn, uplo, diag == 0x100, dtype, seed=m * 7 + n)
        b = Mat(m, n, dtype=dtype, seed=m + 3 * n)
        b0 = b.view().copy()
        a.to_device()
        b.to_device()
        run_trsm(blis, dtype, side, uplo, ta, diag, m, n, 2.0, a, b)
        got = b.device_view()
        ref = b0.copy(order="F")
        oracle.trsm(side, uplo, ta, diag, 2.0, a.view().copy(order="F"), ref)
        err = np.linalg.norm((got - ref).astype(np.float64)) / np.linalg.norm(ref.astype(np.float64))
        assert np.isfinite(got).all()
        assert err <= 8 * max(m, n) * u, (m, n, err)


def test_trsm_row_stored_and_host(blis, oracle):
    m, n = 90, 33
    say("dtrsm row-stored B, host pointers")
    a, _ = tri_matrix(m, LOWER, False, np.float64, seed=4)
    b = Mat(m, n, rs=n, cs=1, seed=8)
    b0 = b.view().copy()
    blis.bli_dtrsm(0, LOWER, 0, 0, m, n, C.byref(C.c_double(1.5)), a.hptr, a.rs, a.cs,
                   b.hptr, b.rs, b.cs)
    ref = b0.copy(order="F")
    oracle.trsm(0, LOWER, "n", 0, 1.5, a.view().copy(order="F"), ref)
    np.testing.assert_allclose(b.view(), ref, rtol=0, atol=1e-13)


def test_dtrsm_f77_alpha_zero(blis):
    m, n = 20, 10
    say("dtrsm_ alpha=0")
    a = Mat(m, m, fill=np.nan).to_device()
    b = Mat(m, n, seed=3).to_device()
    mi, ni, ld = C.c_int32(m), C.c_int32(n), C.c_int32(m)
    blis.dtrsm_(b"L", b"L", b"N", b"N", C.byref(mi), C.byref(ni), C.byref(C.c_double(0.0)),
                a.ptr, C.byref(ld), b.ptr, C.byref(ld))
    np.testing.assert_array_equal(b.device_view(), np.zeros((m, n)))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
def test_trsm_config2_residual(blis, dtype, uplo):
    """Config 2 shape: left/no-trans/non-unit, m=4096, n=1024, A with
    diagonal 4096 + U[0,1], off-diagonal U[-1,1], alpha = 1; lower as in the
    benchmark, and upper (every aligned 512-row block solved by one product
    with its composed inverse either way)."""
    m, n = 4096, 1024
    say(f"trsm config2 {np.dtype(dtype).name} uplo={uplo:#x}")
    rng = np.random.default_rng(77)
    full = rng.uniform(-1, 1, (m, m))
    full = np.tril(full) if uplo == LOWER else np.triu(full)
    full[np.diag_indices(m)] = 4096 + rng.uniform(0, 1, m)
    a = Mat(m, m, dtype=dtype)
    a.set(full.astype(dtype))
    b = Mat(m, n, dtype=dtype, seed=78)
    b0 = b.view().astype(np.float64)
    a.to_device()
    b.to_device()
    run_trsm(blis, dtype, 0, uplo, "n", 0, m, n, 1.0, a, b)
    x = b.device_view().astype(np.float64)
    af = a.view().astype(np.float64)
    resid = np.linalg.norm(af @ x - b0) / (np.linalg.norm(af) * np.linalg.norm(x) * n * unit_roundoff(dtype))
    say(f"trsm config2 scaled residual {resid:.3e}")
    assert resid <= 8.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("diag", [0, 0x100])
@pytest.mark.parametrize("size", ["ragged", "whole"])
def test_trsm_many_leaves(blis, oracle, dtype, side, uplo, ta, diag, size):
    """Solves with several diagonal blocks and many right-hand sides, where
    diagonal blocks that pass the dominance check are solved by products with
    their inverses: a dominant triangle (every block takes the product path)
    and one whose rows are not dominant (every block falls back to
    substitution). "ragged": 700 rows, the last block partial (256-row
    inverses); "whole": 1024 rows, whole 512-row blocks (inverses composed
    from 64-row ones in the band store). Same bar as the small cases."""
    u = unit_roundoff(dtype)
    if size == "ragged":
        m, n = (700, 530) if side == 0 else (530, 700)
    else:
        m, n = (1024, 600) if side == 0 else (600, 1024)
    mn = m if side == 0 else n
    for kind in ("dominant", "not dominant"):
        say(f"trsm many leaves {kind} {np.dtype(dtype).name} side={side} uplo={uplo:#x} {ta} "
            f"diag={diag:#x} m={m} n={n}")
        if kind == "dominant":
            a, _ = tri_matrix(mn, uplo, diag == 0x100, dtype, seed=mn + 5)
        else:
            # off-diagonal entries of size 2/sqrt(mn): row sums far above the
            # diagonal, yet a well-conditioned triangle
            rng = np.random.default_rng(mn + 6)
            full = rng.uniform(-1, 1, (mn, mn)) * (2.0 / np.sqrt(mn))
            full[np.diag_indices(mn)] = 1.0 + rng.uniform(0, 1, mn)
            if uplo == LOWER:
                full[np.triu_indices(mn, 1)] = np.nan
            else:
                full[np.tril_indices(mn, -1)] = np.nan
            if diag == 0x100:
                full[np.diag_indices(mn)] = np.nan
            a = Mat(mn, mn, dtype=dtype, seed=1)
            a.set(full.astype(dtype))
        b = Mat(m, n, dtype=dtype, seed=m + 3 * n)
        b0 = b.view().copy()
        a.to_device()
        b.to_device()
        run_trsm(blis, dtype, side, uplo, ta, diag, m, n, 2.0, a, b)
        got = b.device_view()
        ref = b0.copy(order="F")
        oracle.trsm(side, uplo, ta, diag, 2.0, a.view().copy(order="F"), ref)
        err = np.linalg.norm((got - ref).astype(np.float64)) / np.linalg.norm(ref.astype(np.float64))
        assert np.isfinite(got).all()
        assert err <= 8 * max(m, n) * u, (kind, err)


def test_trsm_many_leaves_async(blis, oracle):
    """The same solves with the library in asynchronous mode (the solve waits
    once, for the check of the diagonal blocks)."""
    import torch
    m, n = 700, 530
    for kind, scale in (("dominant", None), ("not dominant", 2.0)):
        say(f"trsm many leaves async {kind}")
        if scale is None:
            a, _ = tri_matrix(m, LOWER, False, np.float64, seed=m + 5)
        else:
            rng = np.random.default_rng(m + 6)
            full = rng.uniform(-1, 1, (m, m)) * (scale / np.sqrt(m))
            full[np.diag_indices(m)] = 1.0 + rng.uniform(0, 1, m)
            full[np.triu_indices(m, 1)] = np.nan
            a = Mat(m, m, seed=1)
            a.set(full)
        b = Mat(m, n, seed=m + 3 * n)
        b0 = b.view().copy()
        a.to_device()
        b.to_device()
        blis.bli_amd_set_stream(None, 0)
        try:
            run_trsm(blis, np.float64, 0, LOWER, "n", 0, m, n, 2.0, a, b)
            blis.bli_amd_synchronize()
        finally:
            blis.bli_amd_set_stream(None, 1)
        torch.cuda.synchronize()
        got = b.device_view()
        ref = b0.copy(order="F")
        oracle.trsm(0, LOWER, "n", 0, 2.0, a.view().copy(order="F"), ref)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= 8 * max(m, n) * unit_roundoff(np.float64), (kind, err)


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
@pytest.mark.parametrize("ta", ["n", "t"])
def test_trsm_general_stride_dominant(blis, oracle, uplo, ta):
    """A dominant triangle with neither a unit row nor a unit column stride,
    large enough for 512-row blocks (700 rows): every diagonal block passes
    the check, and the composed 512-row inverse, which multiplies
    off-diagonal blocks in place, must not be used for such a triangle."""
    m, n = 700, 530
    say(f"dtrsm general-stride A uplo={uplo:#x} {ta} m={m} n={n}")
    _, full = tri_matrix(m, uplo, False, np.float64, seed=m + 5)
    a = Mat(m, m, rs=2, cs=2 * m, seed=3)
    a.set(full)
    b = Mat(m, n, seed=m + 3 * n)
    b0 = b.view().copy()
    a.to_device()
    b.to_device()
    run_trsm(blis, np.float64, 0, uplo, ta, 0, m, n, 2.0, a, b)
    got = b.device_view()
    ref = b0.copy(order="F")
    oracle.trsm(0, uplo, ta, 0, 2.0, a.view().copy(order="F"), ref)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 8 * max(m, n) * unit_roundoff(np.float64), err


@pytest.mark.parametrize("uplo", [LOWER, UPPER])
def test_trsm_inf_in_b_substitution(blis, oracle, uplo):
    """An Inf in B on the substitution path: rows solved before it stay
    finite and the Inf/NaN pattern of the rest follows the oracle (the
    reference's trsm micro-kernel never updates rows already solved)."""
    m, n = 40, 8
    say(f"dtrsm Inf in B uplo={uplo:#x}")
    a, _ = tri_matrix(m, uplo, False, np.float64, seed=9)
    b = Mat(m, n, seed=10)
    bv = b.view()
    bv[17, 2] = np.inf
    bv[5, 6] = -np.inf
    b0 = bv.copy()
    a.to_device()
    b.to_device()
    run_trsm(blis, np.float64, 0, uplo, "n", 0, m, n, 1.0, a, b)
    got = b.device_view()
    ref = b0.copy(order="F")
    oracle.trsm(0, uplo, "n", 0, 1.0, a.view().copy(order="F"), ref)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(np.isinf(got), np.isinf(ref))
    fin = np.isfinite(ref)
    np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-12, atol=1e-12)
