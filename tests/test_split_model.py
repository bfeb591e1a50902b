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
