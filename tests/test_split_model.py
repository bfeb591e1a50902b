"""CPU checks of the split-precision model (oracle/ozaki.py) and policy.

The model is the bit-exact checker of bli_amd_?gemm_split. The checks here pin
it with facts that do not depend on how it is written:
  * the digits reconstruct each scaled entry exactly, with |digit| <= 127;
  * the result lies within the a-priori truncation bound of the
    double-double oracle product (oracle_dgemm_dd_rows), and with 12 slices it
    is exact to FP64 rounding;
  * a model with a planted mistake (one weight class dropped) breaks the
    bound;
  * at the policy's slice count, the error vs the double-double oracle is no
    larger than the native FP64 GEMM's (oracle dgemm) on the exponent-spread
    workload (u * 2^round(N(0, sigma)));
  * the C library's policy entry point agrees with the model's policy.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import ozaki


def _dd(oracle, a, b):
    m, n = a.shape[0], b.shape[1]
    ref = np.zeros((m, n), order="F")
    oracle.gemm_dd_rows("n", "n", 1.0, np.asfortranarray(a), np.asfortranarray(b), 0.0, ref,
                        0, m, nthreads=4)
    return ref


def test_digits_reconstruct_exactly():
    rng = np.random.default_rng(3)
    x = ozaki.gen_spread(rng, (40, 300), 6.0)
    x[5, :] = 0.0
    s = 9
    digits, e = ozaki.slices(x, s)
    assert all(np.abs(d).max() <= 127 for d in digits)
    # sum_t A_t 2^-7t + r == ldexp(x, -e); r < 2^-7s in magnitude
    approx = sum(np.ldexp(d.astype(np.float64), -7 * (t + 1)) for t, d in enumerate(digits))
    r = np.ldexp(x, -e[:, None]) - approx
    assert np.all(np.abs(r) < 2.0 ** (-7 * s))
    assert np.all(e[5] == 0) and all(np.all(d[5] == 0) for d in digits)
    # entries sharing the row's exponent have 53 bits below it: 8 digits
    # (56 bits) hold them exactly, 7 digits (49 bits) do not
    y = rng.uniform(0.5, 1.0, (4, 50))
    for s, exact in ((8, True), (7, False)):
        d, e = ozaki.slices(y, s)
        back = sum(np.ldexp(v.astype(np.float64), -7 * (t + 1)) for t, v in enumerate(d))
        assert np.array_equal(np.ldexp(back, e[:, None]), y) == exact


@pytest.mark.parametrize("sigma", [0.0, 4.0])
@pytest.mark.parametrize("s", [3, 5, 8])
def test_model_within_truncation_bound(oracle, sigma, s):
    rng = np.random.default_rng(11 + s)
    m, n, k = 24, 20, 700
    a = ozaki.gen_spread(rng, (m, k), sigma)
    b = ozaki.gen_spread(rng, (k, n), sigma)
    ref = _dd(oracle, a, b)
    got = ozaki.split_gemm(a, b, s)
    bound = ozaki.truncation_bound(a, b, s)
    rnd = s * 2.0 ** -53 * np.linalg.norm(np.abs(a) @ np.abs(b))
    err = np.linalg.norm(got - ref)
    assert err <= bound + rnd, (err, bound)


def test_model_with_12_slices_is_fp64_exact(oracle):
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (16, 512))
    b = rng.uniform(-1, 1, (512, 12))
    ref = _dd(oracle, a, b)
    got = ozaki.split_gemm(a, b, 12)
    # within a few ulps of the correctly rounded product, elementwise
    assert np.all(np.abs(got - ref) <= 4 * 2.0 ** -53 * np.abs(a) @ np.abs(b))


def _split_gemm_dropping_top_class(a, b, s):
    A, ea = ozaki.slices(a, s)
    B, eb = ozaki.slices(np.ascontiguousarray(b.T), s)
    v = np.zeros((a.shape[0], b.shape[1]))
    for t in range(s):
        for u in range(s - t):
            if t + u == 0:
                continue  # planted mistake: the weight-2 class is lost
            v += np.ldexp((A[t] @ B[u].T).astype(np.float64), -7 * (t + u + 2))
    return np.ldexp(v, ea[:, None] + eb[None, :])


def test_planted_mistake_breaks_the_bound(oracle):
    rng = np.random.default_rng(8)
    a = rng.uniform(-1, 1, (12, 256))
    b = rng.uniform(-1, 1, (256, 10))
    ref = _dd(oracle, a, b)
    bad = _split_gemm_dropping_top_class(a, b, 6)
    assert np.linalg.norm(bad - ref) > 1e6 * ozaki.truncation_bound(a, b, 6)


@pytest.mark.parametrize("sigma", [0.0, 4.0])
def test_policy_matches_native_error(oracle, sigma):
    rng = np.random.default_rng(int(sigma) + 100)
    m = n = 24
    k = 4096
    a = ozaki.gen_spread(rng, (m, k), sigma)
    b = ozaki.gen_spread(rng, (k, n), sigma)
    ref = _dd(oracle, a, b)
    nat = np.zeros((m, n), order="F")
    oracle.gemm("n", "n", 1.0, np.asfortranarray(a), np.asfortranarray(b), 0.0, nat)
    gap = max(ozaki.mean_gap(a), ozaki.mean_gap(b.T))
    s = ozaki.choose_slices(True, gap)
    den = np.linalg.norm(a) * np.linalg.norm(b)
    err_split = np.linalg.norm(ozaki.split_gemm(a, b, s) - ref) / den
    err_native = np.linalg.norm(nat - ref) / den
    assert err_split <= err_native, (s, err_split, err_native)


def test_policy_slice_counts():
    # uniform [-1,1] entries sit 1/ln2 bits below their row max on average
    assert ozaki.choose_slices(True, 1.0 / np.log(2)) == 8
    assert ozaki.choose_slices(False, 1.0 / np.log(2)) == 4
    assert ozaki.choose_slices(True, 0.0) == 8
    assert ozaki.choose_slices(True, 16.5) == 10


def test_library_policy_agrees_with_model():
    import blis_amd
    L = blis_amd.lib()
    for is_double in (0, 1):
        for gap in (0.0, 1.44, 3.0, 7.9, 16.5, 31.4, 60.0):
            assert L.bli_amd_gemm_split_slices(is_double, C.c_double(gap)) == \
                ozaki.choose_slices(bool(is_double), gap)


# ---- the modular scheme (Ozaki scheme II) ----------------------------------------

def test_crt_constants_are_crt_weights():
    """w_l = 1 mod p_l and 0 mod every other modulus, w_l < M, the 39-bit
    parts of w_l and of M add back to them exactly, and the moduli are
    pairwise coprime."""
    import math
    for n in (2, 9, 16, 24, 32):
        cc = ozaki.crt_const(n)
        ps, M = cc["p"], cc["M"]
        for i in range(n):
            for j in range(i + 1, n):
                assert math.gcd(ps[i], ps[j]) == 1
        for l, parts in enumerate(cc["w"]):
            w = sum(int(v) for v in parts)
            assert 0 <= w < M
            for j, p in enumerate(ps):
                assert w % p == (1 if j == l else 0)
            assert all(float(v).is_integer() and v < 2.0 ** (M.bit_length() + 1) for v in parts)
        assert sum(int(v) for v in cc["mpart"]) == M


@pytest.mark.parametrize("sigma", [0.0, 4.0, 8.0])
def test_crt_reconstructs_the_exact_integer_product(sigma):
    """The FP64 reconstruction gives the exact integer product of the
    operands' integer images (Python integers) to within one rounding."""
    rng = np.random.default_rng(int(sigma) + 40)
    a = ozaki.gen_spread(rng, (7, 90), sigma)
    b = ozaki.gen_spread(rng, (90, 6), sigma)
    a[2, :] = 0.0
    gap = max(ozaki.mean_gap(a), ozaki.mean_gap(b.T))
    n, sh = ozaki.crt_choose(True, gap, 90)
    assert n > 0
    exact = []
    c = ozaki.crt_gemm(a, b, n, sh, exact_out=exact)
    _, ea = ozaki.crt_integers(a, sh)
    _, eb = ozaki.crt_integers(b.T, sh)
    X = np.ldexp(c, -(ea[:, None] + eb[None, :] - 2 * sh))
    for i in range(7):
        for j in range(6):
            e = exact[0][i][j]
            assert abs(int(X[i, j]) - e) <= max(1, abs(e)) * 2.0 ** -52, (i, j)
    assert np.all(c[2] == 0.0)


@pytest.mark.parametrize("sigma", [0.0, 4.0, 8.0])
def test_crt_within_bound_and_native_error(oracle, sigma):
    """The policy's moduli and shift: error vs the double-double oracle within
    the a-priori bound (plus FP64 rounding) and no larger than native GEMM's."""
    rng = np.random.default_rng(int(sigma) + 200)
    m = n = 16
    k = 2048
    a = ozaki.gen_spread(rng, (m, k), sigma)
    b = ozaki.gen_spread(rng, (k, n), sigma)
    ref = _dd(oracle, a, b)
    nat = np.zeros((m, n), order="F")
    oracle.gemm("n", "n", 1.0, np.asfortranarray(a), np.asfortranarray(b), 0.0, nat)
    gap = max(ozaki.mean_gap(a), ozaki.mean_gap(b.T))
    nm, sh = ozaki.crt_choose(True, gap, k)
    got = ozaki.crt_gemm(a, b, nm, sh)
    rnd = 4 * 2.0 ** -53 * np.linalg.norm(np.abs(a) @ np.abs(b))
    assert np.linalg.norm(got - ref) <= ozaki.crt_bound(a, b, sh) + rnd
    den = np.linalg.norm(a) * np.linalg.norm(b)
    assert np.linalg.norm(got - ref) / den <= np.linalg.norm(nat - ref) / den


def test_crt_planted_mistake_too_few_moduli(oracle):
    """Two moduli fewer than the shift needs (M below 4 k 2^(2 sh)) wraps the
    product around M: the result leaves the bound by orders of magnitude."""
    rng = np.random.default_rng(9)
    a = rng.uniform(-1, 1, (10, 300))
    b = rng.uniform(-1, 1, (300, 8))
    ref = _dd(oracle, a, b)
    n, sh = ozaki.crt_choose(True, 1.44, 300)
    good = ozaki.crt_gemm(a, b, n, sh)
    bad = ozaki.crt_gemm(a, b, n - 2, sh)
    bound = ozaki.crt_bound(a, b, sh)
    assert np.linalg.norm(good - ref) <= bound + 4 * 2.0 ** -53 * np.linalg.norm(np.abs(a) @ np.abs(b))
    assert np.linalg.norm(bad - ref) > 1e6 * bound


def test_crt_policy_counts():
    # uniform [-1,1] data at k = 8192: sh = ceil(50 + 1.44) = 52 needs
    # 2*52 + 13 + 2 = 119 bits of M; 15 moduli give 117.6 bits, 16 give 125.2
    assert ozaki.crt_choose(True, 1.0 / np.log(2), 8192) == (16, 52)
    assert ozaki.crt_choose(True, 0.0, 1 << 18) == (0, 50)
    # and the shift supported by n moduli inverts that rule
    assert ozaki.crt_shift(16, 8192) >= 52 and ozaki.crt_shift(15, 8192) < 52


def test_library_crt_policy_agrees_with_model():
    import blis_amd
    L = blis_amd.lib()
    L.bli_amd_gemm_split_moduli.argtypes = [C.c_int, C.c_double, C.c_int64, C.POINTER(C.c_int)]
    L.bli_amd_gemm_split_moduli.restype = C.c_int
    for is_double in (0, 1):
        for gap in (0.0, 1.44, 3.0, 7.9, 16.5, 31.4, 60.0, 120.0):
            for k in (64, 4096, 8192, 100000, 1 << 18):
                sh = C.c_int(-1)
                n = L.bli_amd_gemm_split_moduli(is_double, gap, k, C.byref(sh))
                assert (n, sh.value) == ozaki.crt_choose(bool(is_double), gap, k), (gap, k)
