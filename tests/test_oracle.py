"""Pin the CPU oracle before trusting it (CPU only).

1. The reference's own known-answer drivers: netlib ?blat3 (GEMM, SYMM, TRMM,
   TRSM, SYRK, SYR2K) and
   ?blat2 (GEMV), compiled from the reference tree by oracle/Makefile and
   linked against the oracle's BLAS symbols. They check every trans/side/
   uplo/diag combination, alpha/beta in {0, 1, 0.7}/{0, 1, 1.3}, edge sizes
   {0,1,2,3,5,9}, the error exits (xerbla INFO codes) and untouched padding
   (blastest/src/dblat3.c:4109-4317, dmmch_).
2. Mathematical identities that fix each function independently of how it is
   written: identity and rank-1 GEMMs, brute force in extended precision on
   tiny edge shapes, the BLIS testsuite's probe-vector residual
   (testsuite/src/test_gemm.c:303-414), TRSM on B := A*X with known X, dot
   products of constant vectors, axpy closed forms.
3. Planted mistakes: each checker used above must reject a result with a
   plausible error (a dropped term, a transposed operand, a wrong sign).
"""
from __future__ import annotations

import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import ROOT, normwise_gemm_err, unit_roundoff

REF_DIR = ROOT / "oracle" / "_ref"
NETLIB_IN = ROOT / "tests" / "golden" / "netlib"


def _ensure_ref_built():
    drivers = ("dblat3", "dblat2", "sblat3", "sblat2", "zblat3", "cblat3", "dblat1", "sblat1")
    if not all((REF_DIR / f"{d}_oracle.x").exists() for d in drivers):
        if not Path("/root/reference").exists():
            pytest.skip("netlib drivers not built and the reference tree is absent")
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "all", "ref"], check=True)


def run_netlib(driver: str, suffix: str) -> str:
    exe = REF_DIR / f"{driver}_{suffix}.x"
    deck = NETLIB_IN / f"{driver}.in"
    with tempfile.TemporaryDirectory() as td:
        # The level-1 drivers take no input deck.
        with open(deck if deck.exists() else os.devnull) as fin:
            p = subprocess.run([str(exe)], stdin=fin, cwd=td, capture_output=True, text=True,
                               timeout=600)
        out_file = Path(td) / f"out.{driver}"
        text = out_file.read_text() if out_file.exists() else ""
    assert p.returncode == 0, p.stdout + p.stderr
    return text + p.stdout


NETLIB_CASES = [
    ("dblat3", ["DGEMM", "DSYMM", "DTRMM", "DTRSM", "DSYRK", "DSYR2K"],
     {"DGEMM": 17496, "DSYMM": 1296, "DTRMM": 2592, "DTRSM": 2592, "DSYRK": 1944,
      "DSYR2K": 1944}),
    ("sblat3", ["SGEMM", "SSYMM", "STRMM", "STRSM", "SSYRK", "SSYR2K"],
     {"SGEMM": 17496, "SSYMM": 1296, "STRMM": 2592, "STRSM": 2592, "SSYRK": 1944,
      "SSYR2K": 1944}),
    ("zblat3", ["ZGEMM"], {"ZGEMM": 17496}),
    ("cblat3", ["CGEMM"], {"CGEMM": 17496}),
    ("dblat2", ["DGEMV"], {"DGEMV": 3460}),
    ("sblat2", ["SGEMV"], {"SGEMV": 3460}),
]


def check_netlib_output(text: str, routines, calls):
    for r in routines:
        assert f"{r:<6} PASSED THE TESTS OF ERROR-EXITS" in text, text[-3000:]
        assert f"{r:<6} PASSED THE COMPUTATIONAL TESTS ( {calls[r]:5d} CALLS)" in text or \
            f"{r:<6} PASSED THE COMPUTATIONAL TESTS ({calls[r]:6d} CALLS)" in text, text[-3000:]
    for bad in ("FAILED", "SUSPECT", "******* FATAL ERROR"):
        assert bad not in text, text[-3000:]


@pytest.mark.parametrize("driver,routines,calls", NETLIB_CASES)
def test_netlib_drivers_on_oracle(driver, routines, calls):
    _ensure_ref_built()
    check_netlib_output(run_netlib(driver, "oracle"), routines, calls)


# Level-1 drivers: thirteen routines each, one "Test of subprogram number"
# section per routine ending in "----- PASS -----" when every check held.
NETLIB1_CASES = [
    ("dblat1", ["DDOT", "DAXPY", "DROTG", "DROT", "DCOPY", "DSWAP", "DNRM2", "DASUM", "DSCAL",
                "IDAMAX", "DROTMG", "DROTM", "DSDOT"]),
    ("sblat1", ["SDOT", "SAXPY", "SROTG", "SROT", "SCOPY", "SSWAP", "SNRM2", "SASUM", "SSCAL",
                "ISAMAX", "SROTMG", "SROTM", "SDSDOT"]),
]


def check_netlib1_output(text: str, routines):
    parts = text.split("Test of subprogram number")[1:]
    assert len(parts) == len(routines), text[-3000:]
    for part, name in zip(parts, routines):
        head = part.splitlines()[0].split()
        assert head[-1] == name, (head, name)
        assert "----- PASS -----" in part and "FAIL" not in part, part[:2000]


@pytest.mark.parametrize("driver,routines", NETLIB1_CASES)
def test_netlib1_drivers_on_oracle(driver, routines):
    _ensure_ref_built()
    check_netlib1_output(run_netlib(driver, "oracle"), routines)


# ---------------------------------------------------------------- GEMM --

def _rand(rng, shape, dtype):
    return np.asfortranarray(rng.uniform(-1, 1, shape).astype(dtype))


def _exact_gemm(ta, tb, alpha, a, b, beta, c):
    """Extended-precision triple loop (long double accumulation)."""
    A = (a.T if ta == "t" else a).astype(np.longdouble)
    B = (b.T if tb == "t" else b).astype(np.longdouble)
    out = np.longdouble(alpha) * (A @ B)
    if beta != 0:
        out += np.longdouble(beta) * c.astype(np.longdouble)
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("tb", ["n", "t"])
def test_gemm_bruteforce_edges(oracle, dtype, ta, tb):
    rng = np.random.default_rng(11)
    u = unit_roundoff(dtype)
    for (m, n, k) in [(1, 1, 1), (5, 7, 3), (6, 8, 256), (7, 9, 257), (13, 17, 300), (67, 45, 33)]:
        for alpha, beta in [(1.0, 0.0), (0.7, 1.3), (-2.0, 0.5)]:
            a = _rand(rng, (k, m) if ta == "t" else (m, k), dtype)
            b = _rand(rng, (n, k) if tb == "t" else (k, n), dtype)
            c0 = _rand(rng, (m, n), dtype)
            c = c0.copy(order="F")
            oracle.gemm(ta, tb, alpha, a, b, beta, c)
            ref = _exact_gemm(ta, tb, alpha, a, b, beta, c0)
            err = normwise_gemm_err(c, np.asarray(ref, dtype=np.float64),
                                    a, b, c0, alpha, beta)
            assert err <= 2 * k * u, (m, n, k, alpha, beta, err)


def test_gemm_beta_zero_does_not_read_c(oracle):
    rng = np.random.default_rng(3)
    a, b = _rand(rng, (9, 5), np.float64), _rand(rng, (5, 11), np.float64)
    c = np.full((9, 11), np.nan, order="F")
    oracle.gemm("n", "n", 1.0, a, b, 0.0, c)
    assert np.isfinite(c).all()
    np.testing.assert_allclose(c, a @ b, rtol=1e-14, atol=1e-14)


def test_gemm_alpha_zero_scales_c(oracle):
    a = np.full((4, 3), np.nan, order="F")
    b = np.full((3, 5), np.nan, order="F")
    c = np.asfortranarray(np.arange(20.0).reshape(4, 5))
    oracle.gemm("n", "n", 0.0, a, b, 2.0, c)
    np.testing.assert_array_equal(c, 2.0 * np.arange(20.0).reshape(4, 5))


def test_gemm_identity_and_rank1(oracle):
    rng = np.random.default_rng(5)
    n = 37
    b = _rand(rng, (n, 29), np.float64)
    c0 = _rand(rng, (n, 29), np.float64)
    c = c0.copy(order="F")
    oracle.gemm("n", "n", 2.0, np.asfortranarray(np.eye(n)), b, 0.5, c)
    np.testing.assert_allclose(c, 2.0 * b + 0.5 * c0, rtol=0, atol=4e-16)
    # rank-1: k = 1 gives exact products
    x, y = _rand(rng, (40, 1), np.float64), _rand(rng, (1, 50), np.float64)
    c = np.zeros((40, 50), order="F")
    oracle.gemm("n", "n", 1.0, x, y, 0.0, c)
    np.testing.assert_array_equal(c, x * y)


def probe_residual_gemm(a, b, c, c0, alpha, beta, ta, tb, rng):
    """BLIS testsuite probe-vector residual (testsuite/src/test_gemm.c:303-414):
    v = C t, z = beta C0 t + alpha op(A) (op(B) t), resid = ||v - z||_2."""
    A = a.T if ta == "t" else a
    B = b.T if tb == "t" else b
    t = rng.uniform(-1, 1, c.shape[1])
    t /= 2.0 ** np.ceil(np.log2(np.abs(t).sum()))
    v = c.astype(np.float64) @ t
    z = alpha * (A.astype(np.float64) @ (B.astype(np.float64) @ t)) + beta * (c0.astype(np.float64) @ t)
    return np.linalg.norm(v - z)


def _normalized(rng, shape, dtype):
    """testsuite randomization: U[-1,1] divided by the power of two >= ||X||_1
    (testsuite/src/test_libblis.c:2529-2560)."""
    x = rng.uniform(-1, 1, shape)
    x /= 2.0 ** np.ceil(np.log2(np.abs(x).sum(axis=0).max()))
    return np.asfortranarray(x.astype(dtype))


@pytest.mark.parametrize("dtype,thresh", [(np.float64, 1e-14), (np.float32, 1e-5)])
def test_gemm_testsuite_residual(oracle, dtype, thresh):
    rng = np.random.default_rng(7)
    for (m, n, k) in [(100, 100, 100), (300, 200, 150)]:
        for ta in "nt":
            for tb in "nt":
                a = _normalized(rng, (k, m) if ta == "t" else (m, k), dtype)
                b = _normalized(rng, (n, k) if tb == "t" else (k, n), dtype)
                c0 = _normalized(rng, (m, n), dtype)
                c = c0.copy(order="F")
                oracle.gemm(ta, tb, 2.0, a, b, 1.2, c)
                assert probe_residual_gemm(a, b, c, c0, 2.0, 1.2, ta, tb, rng) < thresh


def test_gemm_general_strides_match_contiguous(oracle):
    rng = np.random.default_rng(9)
    base = rng.uniform(-1, 1, 4000)
    a = np.lib.stride_tricks.as_strided(base, shape=(13, 11), strides=(3 * 8, 50 * 8))
    b = np.asfortranarray(rng.uniform(-1, 1, (11, 7)))
    c_strided = np.zeros((13, 7))  # row-major C
    c_col = np.zeros((13, 7), order="F")
    oracle.gemm("n", "n", 1.0, a, b, 0.0, c_strided)
    oracle.gemm("n", "n", 1.0, np.asfortranarray(a), b, 0.0, c_col)
    np.testing.assert_array_equal(c_strided, c_col)


# ---------------------------------------------------------------- TRSM --

@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("uplo", [0xC0, 0x60])
@pytest.mark.parametrize("ta", ["n", "t"])
@pytest.mark.parametrize("diag", [0, 0x100])
def test_trsm_known_solution(oracle, side, uplo, ta, diag):
    """B := op(A) X (or X op(A)) for a known X; the solve must return X."""
    rng = np.random.default_rng(13)
    for (m, n) in [(7, 5), (67, 45), (300, 20)]:
        mn = m if side == 0 else n
        # Off-diagonal entries of size 1/mn keep the triangle well conditioned
        # also with a unit diagonal.
        a = rng.uniform(-1, 1, (mn, mn)) / mn
        a[np.diag_indices(mn)] = 1.0 + rng.uniform(0, 1, mn)
        tri = np.tril(a) if uplo == 0xC0 else np.triu(a)
        if diag == 0x100:
            np.fill_diagonal(tri, 1.0)
        op = tri.T if ta == "t" else tri
        x = rng.uniform(-1, 1, (m, n))
        alpha = 2.0
        b = (op @ x if side == 0 else x @ op) / alpha
        b = np.asfortranarray(b)
        # The other triangle (and a unit diagonal) must not be read.
        a_in = a.copy()
        if uplo == 0xC0:
            a_in[np.triu_indices(mn, 1)] = np.nan
        else:
            a_in[np.tril_indices(mn, -1)] = np.nan
        if diag == 0x100:
            a_in[np.diag_indices(mn)] = np.nan
        oracle.trsm(side, uplo, ta, diag, alpha, np.asfortranarray(a_in), b)
        np.testing.assert_allclose(b, x, rtol=0, atol=1e-12)


def residual_trsm(a_tri, x, b0, alpha, side):
    """testsuite/src/test_trsm.c:295-389 style check: ||op(A) X - alpha B||."""
    lhs = a_tri @ x if side == 0 else x @ a_tri
    return np.linalg.norm(lhs - alpha * b0) / (np.linalg.norm(a_tri) * np.linalg.norm(x))


# ------------------------------------------------------- level-1v / 2 --

def test_dot_axpy_closed_forms(oracle):
    n = 1000
    x = np.full(n, 0.5)
    y = np.full(n, 3.0)
    assert oracle.dotv(x, y) == 1500.0
    assert oracle.dotv(x[:0], y[:0], n=0) == 0.0
    z = y.copy()
    oracle.axpyv(2.0, x, z)
    np.testing.assert_array_equal(z, np.full(n, 4.0))
    z = y.copy()
    oracle.axpyv(0.0, np.full(n, np.nan), z)
    np.testing.assert_array_equal(z, y)
    # strided with a negative increment walks backwards from the origin
    xs = np.arange(1.0, 11.0)
    ys = np.zeros(10)
    oracle.axpyv(1.0, xs[::-1].copy(), ys, incx=1, incy=1)
    np.testing.assert_array_equal(ys, xs[::-1])


@pytest.mark.parametrize("ta", ["n", "t"])
def test_gemv_matches_extended(oracle, ta):
    rng = np.random.default_rng(17)
    a = np.asfortranarray(rng.uniform(-1, 1, (123, 77)))
    nx = 123 if ta == "t" else 77
    my = 77 if ta == "t" else 123
    x = rng.uniform(-1, 1, nx)
    y0 = rng.uniform(-1, 1, my)
    y = y0.copy()
    oracle.gemv(ta, 0.7, a, x, 1.3, y)
    A = (a.T if ta == "t" else a).astype(np.longdouble)
    ref = 0.7 * (A @ x.astype(np.longdouble)) + 1.3 * y0
    np.testing.assert_allclose(y, np.asarray(ref, dtype=np.float64), rtol=0, atol=1e-13)
    # row-stored A takes the other variant; same result within rounding
    y2 = y0.copy()
    oracle.gemv(ta, 0.7, np.ascontiguousarray(a), x, 1.3, y2)
    np.testing.assert_allclose(y2, y, rtol=0, atol=1e-13)


# ---------------------------------------------------- planted mistakes --

def test_checkers_reject_planted_mistakes(oracle):
    """Each check above must fail on a plausible bug, planted in the oracle's
    output on the CPU: a dropped k term, a transposed operand, a sign flip."""
    rng = np.random.default_rng(23)
    m = n = k = 64
    u = unit_roundoff(np.float64)
    a, b, c0 = (_rand(rng, s, np.float64) for s in [(m, k), (k, n), (m, n)])
    c = c0.copy(order="F")
    oracle.gemm("n", "n", 1.0, a, b, 0.5, c)
    ref = np.asarray(_exact_gemm("n", "n", 1.0, a, b, 0.5, c0), dtype=np.float64)
    assert normwise_gemm_err(c, ref, a, b, c0, 1.0, 0.5) <= 2 * k * u
    dropped = c - np.outer(a[:, -1], b[-1, :])          # last rank-1 term missing
    assert normwise_gemm_err(dropped, ref, a, b, c0, 1.0, 0.5) > 1e-3
    transposed = c0 * 0.5 + a.T @ b                       # A used transposed
    assert normwise_gemm_err(transposed, ref, a, b, c0, 1.0, 0.5) > 1e-3
    sign = c0 * 0.5 - a @ b                               # wrong sign on alpha
    assert normwise_gemm_err(sign, ref, a, b, c0, 1.0, 0.5) > 1e-3
    # probe residual catches a single corrupted entry
    bad = c.copy()
    bad[3, 5] += 1e-9
    r = np.random.default_rng(1)
    assert probe_residual_gemm(a, b, bad, c0, 1.0, 0.5, "n", "n", r) > 1e-12
    # trsm residual catches a solve that skipped the diagonal scaling
    t = np.tril(rng.uniform(-1, 1, (32, 32))) + 32 * np.eye(32)
    bb = np.asfortranarray(rng.uniform(-1, 1, (32, 8)))
    x = bb.copy(order="F")
    oracle.trsm(0, 0xC0, "n", 0, 1.0, np.asfortranarray(t), x)
    assert residual_trsm(t, x, bb, 1.0, 0) < 32 * u
    assert residual_trsm(t, x * 32, bb, 1.0, 0) > 1e-3


def test_double_double_gemm_is_correctly_rounded(oracle):
    """oracle_dgemm_dd_rows (the reference product for split precision)
    against exact rational arithmetic: every entry of alpha*A*B + beta*C
    must be the exact value rounded to the nearest double (up to one ulp on
    rare near-ties), also with catastrophic cancellation."""
    from fractions import Fraction
    rng = np.random.default_rng(12)
    m, n, k = 5, 4, 60
    a = np.asfortranarray(rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-20, 20, (m, k))))
    b = np.asfortranarray(rng.uniform(-1, 1, (k, n)))
    b[:, 0] = 0.0
    b[0, 0], b[1, 0] = 1.0, 1.0
    a[:, 1] = -a[:, 0]          # column 0 of C cancels exactly to zero
    c0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    alpha, beta = 0.75, -1.5
    got = np.asfortranarray(c0.copy())
    oracle.gemm_dd_rows("n", "n", alpha, a, b, beta, got, 0, m)
    for i in range(m):
        for j in range(n):
            exact = Fraction(alpha) * sum(Fraction(a[i, p]) * Fraction(b[p, j]) for p in range(k)) \
                + Fraction(beta) * Fraction(c0[i, j])
            want = float(exact)          # correctly rounded
            assert abs(got[i, j] - want) <= np.spacing(abs(want)), (i, j, got[i, j], want)
    assert np.all(got[:, 0] == beta * c0[:, 0])


# ---------------------------------------------------------- gemm_batch --

@pytest.mark.parametrize("ch", ["d", "s"])
def test_oracle_gemm_batch_is_a_loop_of_gemm(oracle, ch):
    """oracle_?gemm_batch follows the group semantics of ?gemm_batch_
    (frame/compat/extra/bla_gemm_batch.c:172-241): group g's parameters apply
    to its group_size[g] problems, whose pointers follow one another in the
    arrays (an empty group consumes none). Pinned by bitwise equality with
    one oracle gemm call per problem."""
    import ctypes as C
    dt = np.float64 if ch == "d" else np.float32
    T = C.c_double if ch == "d" else C.c_float
    rng = np.random.default_rng(3)
    groups = [(2, "n", "t", 7, 5, 9, 1.5, 0.0), (0, "t", "t", 3, 3, 3, 1.0, 1.0),
              (3, "t", "n", 12, 4, 6, -0.5, 0.25), (1, "n", "n", 1, 8, 2, 2.0, -1.0)]
    probs = []
    for (size, ta, tb, m, n, k, al, be) in groups:
        for _ in range(size):
            a = np.asfortranarray(rng.uniform(-1, 1, (k, m) if ta == "t" else (m, k)).astype(dt))
            b = np.asfortranarray(rng.uniform(-1, 1, (n, k) if tb == "t" else (k, n)).astype(dt))
            c = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dt))
            probs.append((a, b, c))
    G = len(groups)
    TRV = {"n": 0x0, "t": 0x8}
    I = C.c_int64 * G
    ta = (C.c_int * G)(*[TRV[g[1]] for g in groups])
    tb = (C.c_int * G)(*[TRV[g[2]] for g in groups])
    m, n, k = (I(*[g[i] for g in groups]) for i in (3, 4, 5))
    alpha = (T * G)(*[g[6] for g in groups])
    beta = (T * G)(*[g[7] for g in groups])
    lda = I(*[(g[5] if g[1] == "t" else g[3]) for g in groups])
    ldb = I(*[(g[4] if g[2] == "t" else g[5]) for g in groups])
    ldc = I(*[g[3] for g in groups])
    gs = I(*[g[0] for g in groups])
    P = C.c_void_p * len(probs)
    outs = [p[2].copy(order="F") for p in probs]
    aptr = P(*[p[0].ctypes.data for p in probs])
    bptr = P(*[p[1].ctypes.data for p in probs])
    cptr = P(*[o.ctypes.data for o in outs])
    fn = getattr(oracle.lib(), f"oracle_{ch}gemm_batch")
    fn.restype = None
    fn.argtypes = [C.c_void_p] * 13 + [C.c_int64, C.c_void_p]
    fn(C.cast(ta, C.c_void_p), C.cast(tb, C.c_void_p), C.cast(m, C.c_void_p),
       C.cast(n, C.c_void_p), C.cast(k, C.c_void_p), C.cast(alpha, C.c_void_p),
       C.cast(aptr, C.c_void_p), C.cast(lda, C.c_void_p), C.cast(bptr, C.c_void_p),
       C.cast(ldb, C.c_void_p), C.cast(beta, C.c_void_p), C.cast(cptr, C.c_void_p),
       C.cast(ldc, C.c_void_p), G, C.cast(gs, C.c_void_p))
    idx = 0
    for (size, gta, gtb, gm, gn, gk, al, be) in groups:
        for _ in range(size):
            a, b, c = probs[idx]
            ref = c.copy(order="F")
            oracle.gemm(gta, gtb, al, a, b, be, ref)
            np.testing.assert_array_equal(outs[idx], ref)
            idx += 1


# ---- structured level-3: extended-precision pins ---------------------------

def _tri(a, lower, unit=False):
    t = np.tril(a) if lower else np.triu(a)
    if unit:
        np.fill_diagonal(t, 1.0)
    return t


def _sym(a, lower):
    t = np.tril(a) if lower else np.triu(a)
    return t + (np.tril(t, -1).T if lower else np.triu(t, 1).T)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_struc_oracle_against_extended_precision(oracle, dtype):
    """gemmt/syrk/syr2k/symm/trmm/trmm3 of the oracle against products formed
    in long double from explicitly structured dense matrices; the unstored
    triangle of C and of A (NaN-filled) must stay untouched / unread."""
    rng = np.random.default_rng(11)
    u = unit_roundoff(dtype)
    L, U = oracle.LOWER, oracle.UPPER
    ld = np.longdouble

    def close(got, ref, scale, k):
        err = np.max(np.abs(got.astype(ld) - ref)) / scale
        return err <= 4 * (k + 2) * u

    for m, k in [(1, 1), (7, 5), (33, 17)]:
        a = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dtype))
        b = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dtype))
        for uplo in (L, U):
            low = uplo == L
            mask = np.tril(np.ones((m, m), bool)) if low else np.triu(np.ones((m, m), bool))
            c0 = np.asfortranarray(rng.uniform(-1, 1, (m, m)).astype(dtype))
            c0[~mask] = np.nan
            # gemmt with B given transposed (A * B^T)
            c = c0.copy(order="F")
            oracle.gemmt(uplo, "n", "t", 0.5, a, b, 1.5, c)
            ref = 0.5 * (a.astype(ld) @ b.astype(ld).T) + 1.5 * c0.astype(ld)
            assert np.isnan(c[~mask]).all()
            assert close(c[mask], ref[mask], 2.0 * k + 2, k)
            # syrk and syr2k
            c = c0.copy(order="F")
            oracle.syrk(uplo, "n", 0.5, a, 1.5, c)
            ref = 0.5 * (a.astype(ld) @ a.astype(ld).T) + 1.5 * c0.astype(ld)
            assert close(c[mask], ref[mask], 2.0 * k + 2, k) and np.isnan(c[~mask]).all()
            c = c0.copy(order="F")
            oracle.syr2k(uplo, "n", "n", 0.5, a, b, 1.5, c)
            ref = 0.5 * (a.astype(ld) @ b.astype(ld).T + b.astype(ld) @ a.astype(ld).T) \
                + 1.5 * c0.astype(ld)
            assert close(c[mask], ref[mask], 4.0 * k + 2, 2 * k) and np.isnan(c[~mask]).all()
            # symm / trmm / trmm3 with the unstored triangle of A set to NaN
            sq = np.asfortranarray(rng.uniform(-1, 1, (m, m)).astype(dtype))
            sq_nan = sq.copy(order="F")
            sq_nan[~mask] = np.nan
            bb = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dtype))
            cc0 = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(dtype))
            cc = cc0.copy(order="F")
            oracle.symm(oracle.LEFT, uplo, "n", 0.5, sq_nan, bb, 1.5, cc)
            ref = 0.5 * (_sym(sq.astype(ld), low) @ bb.astype(ld)) + 1.5 * cc0.astype(ld)
            assert close(cc, ref, 2.0 * m + 2, m)
            for unit in (False, True):
                tri_nan = sq_nan.copy(order="F")
                if unit:
                    np.fill_diagonal(tri_nan, np.nan)
                bt = bb.copy(order="F")
                oracle.trmm(oracle.LEFT, uplo, "t", oracle.UNIT if unit else oracle.NONUNIT,
                            0.5, tri_nan, bt)
                ref = 0.5 * (_tri(sq.astype(ld), low, unit).T @ bb.astype(ld))
                assert close(bt, ref, 2.0 * m, m)
                cr0 = np.asfortranarray(rng.uniform(-1, 1, (k, m)).astype(dtype))
                cr = cr0.copy(order="F")
                oracle.trmm3(oracle.RIGHT, uplo, "n", oracle.UNIT if unit else oracle.NONUNIT,
                             "t", 0.5, tri_nan, bb, 1.5, cr)
                ref = 0.5 * (bb.astype(ld).T @ _tri(sq.astype(ld), low, unit)) + \
                    1.5 * cr0.astype(ld)
                assert close(cr, ref, 2.0 * m + 2, m)


def test_struc_checks_reject_planted_mistakes(oracle):
    """The extended-precision check above rejects a trmm that uses the wrong
    triangle and a symm that mirrors the stored triangle the wrong way."""
    rng = np.random.default_rng(5)
    m, n = 9, 4
    sq = np.asfortranarray(rng.uniform(-1, 1, (m, m)))
    bb = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    bt = bb.copy(order="F")
    oracle.trmm(oracle.LEFT, oracle.LOWER, "n", oracle.NONUNIT, 1.0, sq, bt)
    wrong_tri = np.triu(sq) @ bb
    assert np.max(np.abs(bt - wrong_tri)) > 1e-3
    cc = np.zeros((m, n), order="F")
    oracle.symm(oracle.LEFT, oracle.LOWER, "n", 1.0, sq, bb, 0.0, cc)
    wrong_sym = (np.tril(sq) + np.tril(sq, -1)) @ bb   # mirror added without transposing
    assert np.max(np.abs(cc - wrong_sym)) > 1e-3



# ---- complex (1m) and mixed-datatype GEMM: extended-precision pins --------

def _op(x, t):
    return {"n": x, "t": x.T, "r": np.conj(x), "c": np.conj(x).T}[t]


@pytest.mark.parametrize("cdt", [np.complex128, np.complex64])
def test_zgemm_oracle_against_extended_precision(oracle, cdt):
    """The 1m oracle against the product formed in complex long double, for
    every transpose/conjugate combination, a complex alpha and beta, beta = 0
    over a NaN C, and a C with general strides."""
    rng = np.random.default_rng(3)
    u = unit_roundoff(np.float64 if cdt == np.complex128 else np.float32)
    ld = np.clongdouble

    def rnd(r, c):
        return (rng.uniform(-1, 1, (r, c)) + 1j * rng.uniform(-1, 1, (r, c))).astype(cdt)

    for (m, n, k) in [(1, 1, 1), (7, 5, 3), (19, 23, 31)]:
        for ta in "ntrc":
            for tb in "ntrc":
                a = np.asfortranarray(rnd(*((k, m) if ta in "tc" else (m, k))))
                b = np.asfortranarray(rnd(*((n, k) if tb in "tc" else (k, n))))
                c0 = np.asfortranarray(rnd(m, n))
                al, be = 0.7 - 0.9j, 1.3 - 1.1j
                c = c0.copy(order="F")
                oracle.zgemm(ta, tb, al, a, b, be, c)
                ref = al * (_op(a.astype(ld), ta) @ _op(b.astype(ld), tb)) + be * c0.astype(ld)
                err = np.max(np.abs(c.astype(ld) - ref))
                assert err <= 8 * (k + 2) * u * (2 * k + 2), (m, n, k, ta, tb, err)
    # beta = 0 ignores a NaN C; general strides
    a, b = np.asfortranarray(rnd(6, 4)), np.asfortranarray(rnd(4, 5))
    big = np.full((12, 15), np.nan, dtype=cdt)
    cv = big[::2, ::3]
    oracle.zgemm("n", "n", 1.0, a, b, 0.0, cv)
    assert np.allclose(cv, a @ b, atol=50 * u)
    assert np.isnan(big[1::2]).all()


def test_zgemm_check_rejects_planted_mistake(oracle):
    """The check above rejects a product with the conjugation dropped."""
    rng = np.random.default_rng(4)
    a = np.asfortranarray(rng.uniform(-1, 1, (5, 4)) + 1j * rng.uniform(-1, 1, (5, 4)))
    b = np.asfortranarray(rng.uniform(-1, 1, (4, 3)) + 1j * rng.uniform(-1, 1, (4, 3)))
    c = np.zeros((5, 3), dtype=complex, order="F")
    oracle.zgemm("r", "n", 1.0, a, b, 0.0, c)
    assert np.max(np.abs(c - a @ b)) > 1e-3


def test_gemm_md_oracle_rules(oracle):
    """The mixed-datatype oracle against the documented rules, formed
    directly in long double: real C += complex A * complex B keeps the real
    part of the complex product; complex C += real A * real B updates only
    the real parts; float C from double A, B in double precision."""
    rng = np.random.default_rng(9)
    ld, cld = np.longdouble, np.clongdouble
    m, n, k = 9, 7, 5
    az = rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))
    bz = rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))
    cr = rng.uniform(-1, 1, (m, n))
    got = oracle.gemm_md("n", "n", 0.5, az, bz, 2.0, cr, True)
    ref = 0.5 * (az.astype(cld) @ bz.astype(cld)).real + 2.0 * cr.astype(ld)
    assert np.max(np.abs(got - ref)) < 1e-13
    ar, br = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n))
    cz = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    got = oracle.gemm_md("n", "n", 0.5, ar, br, 2.0, cz, True)
    assert np.max(np.abs(got.real - (0.5 * ar.astype(ld) @ br.astype(ld) + 2 * cz.real))) < 1e-13
    assert np.array_equal(got.imag, cz.imag)
    cf = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    got = oracle.gemm_md("t", "n", 1.0, ar.T.copy(), br, 1.0, cf, True)
    ref = (ar.astype(ld) @ br.astype(ld) + cf.astype(ld)).astype(np.float32)
    assert got.dtype == np.float32 and np.array_equal(got, ref)
