"""Split-precision GEMM (bli_amd_?gemm_split) on the GPU.

Bars:
  * beta == 0, one pass over k: bit-identical to the CPU model of the same
    algorithm (oracle/ozaki.py). That model is pinned by the double-double
    oracle and the truncation bound in test_split_model.py;
  * beta != 0, or several passes over k: within the reported truncation bound
    plus FP64 rounding (8 u |A||B|) of the double-double oracle;
  * policy: on u * 2^round(N(0, sigma)) data, the automatic slice count gives
    an error vs the double-double oracle no larger than native bli_dgemm's;
  * Inf/NaN operands and slice counts past break-even run the native path
    (return 0, result identical to bli_?gemm).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import ozaki
from tests.conftest import Mat, say

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}


def _T(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def _ch(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def split(blis, dtype, ta, tb, m, n, k, alpha, beta, a, b, c, s):
    T = _T(dtype)
    eb = C.c_double(-1.0)
    got = getattr(blis, f"bli_amd_{_ch(dtype)}gemm_split")(
        TR[ta], TR[tb], m, n, k, C.byref(T(alpha)), a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs,
        C.byref(T(beta)), c.ptr, c.rs, c.cs, s, C.byref(eb))
    return got, eb.value


def native(blis, dtype, ta, tb, m, n, k, alpha, beta, a, b, c):
    T = _T(dtype)
    getattr(blis, f"bli_{_ch(dtype)}gemm")(TR[ta], TR[tb], m, n, k, C.byref(T(alpha)), a.ptr,
                                           a.rs, a.cs, b.ptr, b.rs, b.cs, C.byref(T(beta)),
                                           c.ptr, c.rs, c.cs)


def operands(dtype, ta, tb, m, n, k, seed, layout, sigma=0.0):
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
    else:  # general strides and padding
        a = Mat(am, an, rs=2, cs=2 * am + 1, dtype=dtype, seed=seed)
        b = Mat(bm, bn, rs=bn + 2, cs=3, dtype=dtype, seed=seed + 1)
        c = Mat(m, n, rs=1, cs=m + 3, dtype=dtype, seed=seed + 2, off=1)
    if sigma > 0:
        rng = np.random.default_rng(seed + 7)
        a.set(ozaki.gen_spread(rng, (am, an), sigma).astype(dtype))
        b.set(ozaki.gen_spread(rng, (bm, bn), sigma).astype(dtype))
    return a, b, c


def eff(x: Mat, t: str):
    v = x.view().astype(np.float64)
    return v.T if t == "t" else v


def dd_ref(oracle, a_eff, b_eff, alpha=1.0, beta=0.0, c0=None):
    m, n = a_eff.shape[0], b_eff.shape[1]
    ref = np.zeros((m, n), order="F") if c0 is None else np.asfortranarray(c0.astype(np.float64))
    oracle.gemm_dd_rows("n", "n", alpha, np.asfortranarray(a_eff), np.asfortranarray(b_eff),
                        beta, ref, 0, m, nthreads=8)
    return ref


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")])
@pytest.mark.parametrize("layout", ["col", "row", "gen"])
def test_bit_identical_to_model(blis, dtype, ta, tb, layout):
    m, n, k = 70, 90, 300
    s = 8 if dtype == np.float64 else 4
    say(f"split bitexact {np.dtype(dtype).name} {ta}{tb} {layout} {m}x{n}x{k} s={s}")
    a, b, c = operands(dtype, ta, tb, m, n, k, 21, layout, sigma=2.0)
    c.set(np.nan)  # beta == 0 must not read C
    a.to_device(); b.to_device(); c.to_device()
    got, bound = split(blis, dtype, ta, tb, m, n, k, 1.0, 0.0, a, b, c, s)
    assert got == s and bound > 0
    want = ozaki.split_gemm(eff(a, ta), eff(b, tb), s).astype(dtype)
    assert np.array_equal(c.device_view(), want)


@pytest.mark.parametrize("s", [2, 3, 5, 9, 10, 12])
def test_every_slice_count_bit_identical(blis, s):
    m, n, k = 130, 65, 200
    say(f"split bitexact s={s} {m}x{n}x{k}")
    a, b, c = operands(np.float64, "n", "n", m, n, k, 40 + s, "col", sigma=3.0)
    a.to_device(); b.to_device(); c.to_device()
    got, _ = split(blis, np.float64, "n", "n", m, n, k, -0.75, 0.0, a, b, c, s)
    assert got == s
    want = ozaki.split_gemm(eff(a, "n"), eff(b, "n"), s, alpha=-0.75)
    assert np.array_equal(c.device_view(), want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_beta_and_bound(blis, oracle, dtype):
    m, n, k = 96, 80, 1000
    s = 7 if dtype == np.float64 else 3
    say(f"split beta/bound {np.dtype(dtype).name} s={s}")
    a, b, c = operands(dtype, "n", "t", m, n, k, 3, "col", sigma=1.0)
    c0 = c.view().astype(np.float64).copy()
    a.to_device(); b.to_device(); c.to_device()
    alpha, beta = 0.5, -1.25
    got, bound = split(blis, dtype, "n", "t", m, n, k, alpha, beta, a, b, c, s)
    assert got == s
    ae, be = eff(a, "n"), eff(b, "t")
    ref = dd_ref(oracle, ae, be, alpha, beta, c0)
    u = 2.0 ** -53 if dtype == np.float64 else 2.0 ** -24
    rnd = 8 * u * (abs(alpha) * np.linalg.norm(np.abs(ae) @ np.abs(be)) + abs(beta) * np.linalg.norm(c0))
    err = np.linalg.norm(c.device_view().astype(np.float64) - ref)
    assert err <= bound + rnd, (err, bound, rnd)


def test_multi_pass_over_k(blis, oracle):
    # 12 slices allow 173 chunks of 64 per exact int32 pass: k = 11200 takes two
    m, n, k, s = 64, 48, 11200, 12
    say(f"split multipass {m}x{n}x{k} s={s}")
    a, b, c = operands(np.float64, "n", "n", m, n, k, 9, "col")
    a.to_device(); b.to_device(); c.to_device()
    got, bound = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, s)
    assert got == s
    ae, be = eff(a, "n"), eff(b, "n")
    ref = dd_ref(oracle, ae, be)
    err = np.abs(c.device_view() - ref)
    assert np.all(err <= 4 * 2.0 ** -53 * (np.abs(ae) @ np.abs(be)) + bound)


@pytest.mark.parametrize("sigma", [0.0, 4.0])
def test_policy_error_not_above_native(blis, oracle, sigma):
    m, n, k = 128, 96, 4096
    say(f"split policy sigma={sigma} {m}x{n}x{k}")
    a, b, c = operands(np.float64, "n", "n", m, n, k, 77, "col", sigma=sigma)
    cn = Mat(m, n, seed=1)
    a.to_device(); b.to_device(); c.to_device(); cn.to_device()
    gap = max(ozaki.mean_gap(eff(a, "n")), ozaki.mean_gap(eff(b, "n").T))
    s_pol = ozaki.choose_slices(True, gap)
    # the automatic choice: the modular scheme when its moduli are fewer
    # than the slice scheme's products (FP64), else the slice policy
    n_crt, _ = ozaki.crt_choose(True, gap, k)
    s_max = blis.bli_amd_gemm_split_max_slices(1)
    want = s_pol if s_pol <= s_max else 0
    if n_crt > 0 and (want == 0 or n_crt < want * (want + 1) // 2):
        want = -n_crt
    auto, _ = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, 0)
    assert auto == want
    # the policy's slice count, forced, must match native accuracy
    got, bound = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, s_pol)
    assert got == s_pol
    native(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, cn)
    ae, be = eff(a, "n"), eff(b, "n")
    ref = dd_ref(oracle, ae, be)
    e_split = np.linalg.norm(c.device_view() - ref)
    e_nat = np.linalg.norm(cn.device_view() - ref)
    say(f"  s={got} split_err={e_split:.3e} native_err={e_nat:.3e} bound={bound:.3e}")
    assert e_split <= e_nat
    assert e_split <= bound + 8 * 2.0 ** -53 * np.linalg.norm(np.abs(ae) @ np.abs(be))


def test_wide_spread_falls_back_to_native(blis):
    # sigma = 40: entries spread over about 120 bits below their row maxima,
    # past both 8 slices and 32 moduli
    m, n, k = 100, 70, 512
    say("split fallback sigma=40")
    a, b, c = operands(np.float64, "n", "n", m, n, k, 5, "col", sigma=40.0)
    cn = Mat(m, n, seed=3)
    a.to_device(); b.to_device(); c.to_device(); cn.to_device()
    got, bound = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, 0)
    assert got == 0 and bound == 0.0
    native(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, cn)
    assert np.array_equal(c.device_view(), cn.device_view())


@pytest.mark.parametrize("bad", [np.inf, np.nan])
def test_nonfinite_operand_falls_back_to_native(blis, bad):
    m, n, k = 33, 40, 129
    say(f"split nonfinite {bad}")
    a, b, c = operands(np.float64, "n", "n", m, n, k, 6, "col")
    v = a.view().copy(); v[3, 7] = bad; a.set(v)
    cn = Mat(m, n, seed=4)
    a.to_device(); b.to_device(); c.to_device(); cn.to_device()
    got, _ = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, 8)
    assert got == 0
    native(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, cn)
    np.testing.assert_array_equal(c.device_view(), cn.device_view())
    assert not np.all(np.isfinite(c.device_view()[3]))


def test_quick_returns_and_zero_rows(blis):
    say("split quick returns")
    m, n, k = 50, 30, 64
    a, b, c = operands(np.float64, "n", "n", m, n, k, 12, "col")
    c0 = c.view().copy()
    a.to_device(); b.to_device(); c.to_device()
    # alpha == 0: C := beta C
    got, _ = split(blis, np.float64, "n", "n", m, n, k, 0.0, 2.0, a, b, c, 8)
    assert got == 0
    np.testing.assert_array_equal(c.device_view(), 2.0 * c0)
    # k == 0: C := beta C
    got, _ = split(blis, np.float64, "n", "n", m, n, 0, 1.0, 0.5, a, b, c, 8)
    np.testing.assert_array_equal(c.device_view(), c0)
    # m == 0: nothing touched
    got, _ = split(blis, np.float64, "n", "n", 0, n, k, 1.0, 0.0, a, b, c, 8)
    np.testing.assert_array_equal(c.device_view(), c0)
    # all-zero rows and columns give exact zeros
    av = a.view().copy(); av[4, :] = 0; a.set(av)
    bv = b.view().copy(); bv[:, 9] = 0; b.set(bv)
    a.to_device(); b.to_device()
    got, _ = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, 8)
    out = c.device_view()
    assert got == 8 and np.all(out[4, :] == 0) and np.all(out[:, 9] == 0)
    np.testing.assert_array_equal(out, ozaki.split_gemm(eff(a, "n"), eff(b, "n"), 8))


def test_one_slice_runs_as_two(blis):
    m, n, k = 40, 36, 150
    say(f"split one slice {m}x{n}x{k}")
    a, b, c = operands(np.float64, "n", "n", m, n, k, 17, "col", sigma=1.0)
    a.to_device(); b.to_device(); c.to_device()
    got, bound = split(blis, np.float64, "n", "n", m, n, k, 1.0, 0.0, a, b, c, 1)
    assert got == 2 and bound > 0
    assert np.array_equal(c.device_view(), ozaki.split_gemm(eff(a, "n"), eff(b, "n"), 2))


def test_host_pointers_are_staged(blis):
    say("split host pointers")
    m, n, k = 40, 44, 100
    a, b, c = operands(np.float64, "n", "n", m, n, k, 13, "col")
    T = C.c_double
    eb = C.c_double()
    got = blis.bli_amd_dgemm_split(0, 0, m, n, k, C.byref(T(1.0)), a.hptr, a.rs, a.cs, b.hptr,
                                   b.rs, b.cs, C.byref(T(0.0)), c.hptr, c.rs, c.cs, 8,
                                   C.byref(eb))
    assert got == 8
    np.testing.assert_array_equal(c.view(), ozaki.split_gemm(eff(a, "n"), eff(b, "n"), 8))


# ---- the modular scheme (Ozaki scheme II, nslices < 0) ---------------------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("n", "n"), ("t", "n"), ("n", "t")])
@pytest.mark.parametrize("layout", ["col", "row", "gen"])
def test_modular_bit_identical_to_model(blis, dtype, ta, tb, layout):
    """Forced moduli: the result equals the CPU model (oracle/ozaki.py
    crt_gemm) bit for bit, with beta = 0 and beta != 0, across two output
    tiles of 256 in each direction and a partial k chunk."""
    m, n, k = 300, 270, 130
    nmod = 16 if np.dtype(dtype) == np.float64 else 9
    for beta in (0.0, 0.5):
        say(f"split modular {np.dtype(dtype).name} {ta}{tb} {layout} beta={beta} {m}x{n}x{k} "
            f"moduli={nmod}")
        a, b, c = operands(dtype, ta, tb, m, n, k, 31, layout)
        c0 = c.view().astype(np.float64).copy()
        a.to_device(); b.to_device(); c.to_device()
        got, bound = split(blis, dtype, ta, tb, m, n, k, 0.75, beta, a, b, c, -nmod)
        assert got == -nmod and bound > 0
        sh = ozaki.crt_shift(nmod, k)
        ref = ozaki.crt_gemm(eff(a, ta), eff(b, tb), nmod, sh, alpha=0.75, beta=beta, c0=c0)
        np.testing.assert_array_equal(c.device_view(), ref.astype(dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_modular_all_moduli_bit_identical(blis, dtype, layout):
    """All 32 moduli at depth 8: the shift is 117 bits, so the largest
    scaled integers have 117 bits and their residues need 2^e mod p for
    e = 64; the result still equals the CPU model bit for bit."""
    m, n, k = 70, 66, 8
    nmod = 32
    sh = ozaki.crt_shift(nmod, k)
    assert sh > 116
    say(f"split modular all moduli {np.dtype(dtype).name} {layout} {m}x{n}x{k} sh={sh}")
    a, b, c = operands(dtype, "n", "n", m, n, k, 32, layout)
    c0 = c.view().astype(np.float64).copy()
    a.to_device(); b.to_device(); c.to_device()
    got, bound = split(blis, dtype, "n", "n", m, n, k, 1.25, 0.5, a, b, c, -nmod)
    assert got == -nmod and bound > 0
    ref = ozaki.crt_gemm(eff(a, "n"), eff(b, "n"), nmod, sh, alpha=1.25, beta=0.5, c0=c0)
    np.testing.assert_array_equal(c.device_view(), ref.astype(dtype))


@pytest.mark.parametrize("dtype,sigma", [(np.float64, 0.0), (np.float64, 4.0),
                                         (np.float64, 8.0), (np.float32, 0.0),
                                         (np.float32, 4.0)])
def test_modular_policy_error_not_above_native(blis, oracle, dtype, sigma):
    """The automatic choice on u * 2^round(N(0, sigma)) data takes the
    modular scheme with the policy's moduli (fewer INT8 products than the
    slices); its error against the double-double oracle is no larger than
    native bli_?gemm's and within the reported bound plus rounding of the
    output type."""
    m, n, k = 128, 96, 4096
    is_d = np.dtype(dtype) == np.float64
    say(f"split modular policy {np.dtype(dtype).name} sigma={sigma} {m}x{n}x{k}")
    a, b, c = operands(dtype, "n", "n", m, n, k, 78, "col", sigma=sigma)
    cn = Mat(m, n, dtype=dtype, seed=2)
    a.to_device(); b.to_device(); c.to_device(); cn.to_device()
    gap = max(ozaki.mean_gap(eff(a, "n")), ozaki.mean_gap(eff(b, "n").T))
    n_crt, sh = ozaki.crt_choose(is_d, gap, k)
    s_pol = ozaki.choose_slices(is_d, gap)
    assert n_crt > 0 and n_crt < s_pol * (s_pol + 1) // 2
    got, bound = split(blis, dtype, "n", "n", m, n, k, 1.0, 0.0, a, b, c, 0)
    assert got == -n_crt
    native(blis, dtype, "n", "n", m, n, k, 1.0, 0.0, a, b, cn)
    ae, be = eff(a, "n"), eff(b, "n")
    ref = dd_ref(oracle, ae, be)
    e_split = np.linalg.norm(c.device_view().astype(np.float64) - ref)
    e_nat = np.linalg.norm(cn.device_view().astype(np.float64) - ref)
    say(f"  moduli={n_crt} sh={sh} split_err={e_split:.3e} native_err={e_nat:.3e} bound={bound:.3e}")
    assert e_split <= e_nat
    u = 2.0 ** -53 if is_d else 2.0 ** -24
    assert e_split <= bound + 8 * u * np.linalg.norm(np.abs(ae) @ np.abs(be))
