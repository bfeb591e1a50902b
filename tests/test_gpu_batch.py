"""Batched GEMM on the GPU: one persistent launch over uneven problems.

Pins (frame/compat/extra/bla_gemm_batch.c:172-241 defines gemm_batch as a
loop of gemm over groups and problems, and the reference has no other
result check for it):
  * every problem equals the oracle's GEMM within the GEMM bar (4 k u);
  * every problem is bitwise equal to a plain single-GEMM launch of the same
    tile configuration (64x64 tiles, or 32x32 for problems with m, n <= 48),
    which pins the persistent scheduler as pure scheduling;
  * a plan replayed as a HIP graph gives bitwise the same result on every
    replay, with eager work between replays;
  * the BLAS symbol dgemm_batch_ with group semantics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import Mat, normwise_gemm_err, say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}


def make_batch(dtype, count, seed, lo=8, hi=200):
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        m, n, k = (int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3))
        ta, tb = rng.choice(["n", "t"]), rng.choice(["n", "t"])
        am, an = (k, m) if ta == "t" else (m, k)
        bm, bn = (n, k) if tb == "t" else (k, n)
        pad = lambda v: (v + 7) // 8 * 8
        a = Mat(am, an, rs=1, cs=pad(am), dtype=dtype, seed=seed * 1000 + 3 * i)
        b = Mat(bm, bn, rs=1, cs=pad(bm), dtype=dtype, seed=seed * 1000 + 3 * i + 1)
        c = Mat(m, n, rs=1, cs=pad(m), dtype=dtype, seed=seed * 1000 + 3 * i + 2)
        alpha, beta = float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1))
        probs.append(dict(m=m, n=n, k=k, ta=ta, tb=tb, a=a, b=b, c=c, alpha=alpha, beta=beta,
                          c0=c.view().copy()))
    return probs


def typed_batch_args(probs, dtype):
    G = len(probs)
    T = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    I = C.c_int64 * G
    P = C.c_void_p * G
    ta = (C.c_int * G)(*[TR[p["ta"]] for p in probs])
    tb = (C.c_int * G)(*[TR[p["tb"]] for p in probs])
    m, n, k = (I(*[p[x] for p in probs]) for x in "mnk")
    alpha = (T * G)(*[p["alpha"] for p in probs])
    beta = (T * G)(*[p["beta"] for p in probs])
    a, b, c = (P(*[p[x].ptr for p in probs]) for x in "abc")
    rsa, csa = I(*[p["a"].rs for p in probs]), I(*[p["a"].cs for p in probs])
    rsb, csb = I(*[p["b"].rs for p in probs]), I(*[p["b"].cs for p in probs])
    rsc, csc = I(*[p["c"].rs for p in probs]), I(*[p["c"].cs for p in probs])
    gs = I(*([1] * G))
    keep = [ta, tb, m, n, k, alpha, beta, a, b, c, rsa, csa, rsb, csb, rsc, csc, gs]
    args = [C.cast(x, C.c_void_p) for x in (ta, tb, m, n, k, alpha, a, rsa, csa, b, rsb, csb,
                                            beta, c, rsc, csc)] + [G, C.cast(gs, C.c_void_p)]
    return args, keep


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_matches_oracle_and_single_gemm(blis, oracle, dtype):
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    probs = make_batch(dtype, 96, seed=5 if ch == "d" else 6)
    for p in probs:
        for x in "abc":
            p[x].to_device()
    say(f"{ch}gemm_batch typed, {len(probs)} problems")
    args, keep = typed_batch_args(probs, dtype)
    getattr(blis, f"bli_{ch}gemm_batch")(*args)
    u = unit_roundoff(dtype)
    T = C.c_double if ch == "d" else C.c_float
    for i, p in enumerate(probs):
        got = p["c"].device_view().copy()
        ref = p["c0"].copy(order="F")
        oracle.gemm(p["ta"], p["tb"], p["alpha"], p["a"].view(), p["b"].view(), p["beta"], ref)
        err = normwise_gemm_err(got, ref, p["a"].view(), p["b"].view(), p["c0"], p["alpha"], p["beta"])
        assert err <= 4 * p["k"] * u, (i, p["m"], p["n"], p["k"], err)
        # Same problem through a plain launch with the same tile shape.
        tile = 4 if (p["m"] <= 48 and p["n"] <= 48) else 3
        single = Mat(p["m"], p["n"], rs=1, cs=p["c"].cs, dtype=dtype)
        single.buf[:] = 0
        single.set(p["c0"])
        single.to_device()
        blis.bli_amd_gemm_set_config(tile, 8, 1, 1)
        try:
            getattr(blis, f"bli_{ch}gemm")(TR[p["ta"]], TR[p["tb"]], p["m"], p["n"], p["k"],
                                           C.byref(T(p["alpha"])), p["a"].ptr, p["a"].rs,
                                           p["a"].cs, p["b"].ptr, p["b"].rs, p["b"].cs,
                                           C.byref(T(p["beta"])), single.ptr, single.rs,
                                           single.cs)
        finally:
            blis.bli_amd_gemm_set_config(-1, 8, 1, 1)
        np.testing.assert_array_equal(single.device_view(), got)


def test_batch_plan_graph_replay(blis):
    import torch
    probs = make_batch(np.float64, 64, seed=9)
    for p in probs:
        for x in "abc":
            p[x].to_device()
    args, keep = typed_batch_args(probs, np.float64)
    say("batch plan + graph replay")
    plan = blis.bli_amd_dgemm_batch_plan_create(*args)
    assert plan
    try:
        # beta = 0 makes every replay idempotent: C := alpha*A*B
        for p in probs:
            p["beta"] = 0.0
        args, keep2 = typed_batch_args(probs, np.float64)
        plan0 = blis.bli_amd_dgemm_batch_plan_create(*args)
        blis.bli_amd_gemm_batch_plan_execute(plan0, 0)
        eager = [p["c"].device_view().copy() for p in probs]
        for rep in range(3):
            for p in probs:          # eager work between replays: scribble on C
                p["c"].dbuf.fill_(rep + 1.0)
            torch.cuda.synchronize()
            blis.bli_amd_gemm_batch_plan_execute(plan0, 1)
            for p, e in zip(probs, eager):
                np.testing.assert_array_equal(p["c"].device_view(), e)
        assert blis.bli_amd_gemm_batch_plan_work_items(plan0) > len(probs)
        blis.bli_amd_gemm_batch_plan_destroy(plan0)
    finally:
        blis.bli_amd_gemm_batch_plan_destroy(plan)


def test_dgemm_batch_f77_groups(blis, oracle):
    """Two groups of 3 problems each through dgemm_batch_ (group semantics)."""
    say("dgemm_batch_ f77 groups")
    groups = [dict(ta=b"N", tb=b"T", m=33, n=20, k=17, alpha=0.5, beta=1.0),
              dict(ta=b"T", tb=b"N", m=70, n=65, k=40, alpha=-1.0, beta=0.0)]
    probs = []
    for gi, g in enumerate(groups):
        for s in range(3):
            am, an = (g["k"], g["m"]) if g["ta"] == b"T" else (g["m"], g["k"])
            bm, bn = (g["n"], g["k"]) if g["tb"] == b"T" else (g["k"], g["n"])
            a = Mat(am, an, seed=100 + 10 * gi + s).to_device()
            b = Mat(bm, bn, seed=200 + 10 * gi + s).to_device()
            c = Mat(g["m"], g["n"], seed=300 + 10 * gi + s)
            c0 = c.view().copy()
            c.to_device()
            probs.append((g, a, b, c, c0))
    G = 2
    I = C.c_int32 * G
    ta = (C.c_char * G)(*[g["ta"] for g in groups])
    tb = (C.c_char * G)(*[g["tb"] for g in groups])
    m, n, k = (I(*[g[x] for g in groups]) for x in "mnk")
    lda = I(*[(g["k"] if g["ta"] == b"T" else g["m"]) for g in groups])
    ldb = I(*[(g["n"] if g["tb"] == b"T" else g["k"]) for g in groups])
    ldc = I(*[g["m"] for g in groups])
    alpha = (C.c_double * G)(*[g["alpha"] for g in groups])
    beta = (C.c_double * G)(*[g["beta"] for g in groups])
    P6 = C.c_void_p * 6
    a_arr, b_arr, c_arr = (P6(*[p[i].ptr for p in probs]) for i in (1, 2, 3))
    gc, gs = C.c_int32(G), I(3, 3)
    blis.dgemm_batch_(ta, tb, m, n, k, alpha, a_arr, lda, b_arr, ldb, beta, c_arr, ldc,
                      C.byref(gc), gs)
    for (g, a, b, c, c0) in probs:
        ref = c0.copy(order="F")
        oracle.gemm(g["ta"].decode().lower(), g["tb"].decode().lower(), g["alpha"], a.view(),
                    b.view(), g["beta"], ref)
        err = normwise_gemm_err(c.device_view(), ref, a.view(), b.view(), c0, g["alpha"], g["beta"])
        assert err <= 4 * g["k"] * unit_roundoff(np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_trivial_problems(blis, dtype):
    """k == 0 and alpha == 0 problems inside a batch: C := beta*C, A and B
    are not read (their pointers are null here); m == 1 / n == 1 problems."""
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    say(f"{ch}gemm_batch trivial problems")
    probs = make_batch(dtype, 12, seed=21, lo=1, hi=70)
    probs[0].update(k=0)
    probs[3].update(alpha=0.0)
    probs[5].update(m=1)
    probs[7].update(n=1)
    for p in probs:
        if p is probs[5] or p is probs[7]:
            p["c"] = Mat(p["m"], p["n"], rs=1, cs=max(p["m"], 1), dtype=dtype, seed=77)
            p["c0"] = p["c"].view().copy()
            am, an = (p["k"], p["m"]) if p["ta"] == "t" else (p["m"], p["k"])
            bm, bn = (p["n"], p["k"]) if p["tb"] == "t" else (p["k"], p["n"])
            p["a"] = Mat(am, an, dtype=dtype, seed=78)
            p["b"] = Mat(bm, bn, dtype=dtype, seed=79)
        for x in "abc":
            p[x].to_device()
    args, keep = typed_batch_args(probs, dtype)
    # null A and B for the k == 0 and alpha == 0 problems
    aptrs = C.cast(args[6], C.POINTER(C.c_void_p))
    bptrs = C.cast(args[9], C.POINTER(C.c_void_p))
    for i in (0, 3):
        aptrs[i] = None
        bptrs[i] = None
    getattr(blis, f"bli_{ch}gemm_batch")(*args)
    u = unit_roundoff(dtype)
    for i, p in enumerate(probs):
        got = p["c"].device_view()
        if i in (0, 3):
            np.testing.assert_array_equal(got, (p["beta"] * p["c0"]).astype(dtype))
        else:
            a = p["a"].view().astype(np.float64)
            b = p["b"].view().astype(np.float64)
            ref = p["alpha"] * ((a.T if p["ta"] == "t" else a) @ (b.T if p["tb"] == "t" else b)) \
                + p["beta"] * p["c0"].astype(np.float64)
            err = normwise_gemm_err(got, ref, a, b, p["c0"], p["alpha"], p["beta"])
            assert err <= 4 * max(p["k"], 1) * u, (i, err)


NMODE = 9   # batch kernel modes


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_modes_and_autotune(blis, dtype):
    """Every kernel mode gives bit-identical results; the autotuner leaves C
    as it found it and picks a valid mode."""
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    say(f"{ch}gemm_batch modes + autotune")
    probs = make_batch(dtype, 80, seed=31, lo=5, hi=300)
    for p in probs:
        p["beta"] = 0.0            # each execute is then idempotent
        for x in "abc":
            p[x].to_device()
    args, keep = typed_batch_args(probs, dtype)
    plan = getattr(blis, f"bli_amd_{ch}gemm_batch_plan_create")(*args)
    try:
        results = []
        for mode in range(NMODE):
            blis.bli_amd_gemm_batch_plan_set_mode(plan, mode)
            assert blis.bli_amd_gemm_batch_plan_get_mode(plan) == mode
            for p in probs:
                p["c"].dbuf.fill_(7.0)
            blis.bli_amd_gemm_batch_plan_execute(plan, mode % 2)
            results.append([p["c"].device_view().copy() for p in probs])
        for mode in range(1, NMODE):
            for r0, r in zip(results[0], results[mode]):
                np.testing.assert_array_equal(r, r0)
        # the autotuner restores C
        for p in probs:
            p["c"].dbuf.fill_(3.0)
        before = [p["c"].device_view().copy() for p in probs]
        tuned = blis.bli_amd_gemm_batch_plan_autotune(plan, 2)
        assert 0 <= tuned < NMODE and blis.bli_amd_gemm_batch_plan_get_mode(plan) == tuned
        for p, b in zip(probs, before):
            np.testing.assert_array_equal(p["c"].device_view(), b)
        blis.bli_amd_gemm_batch_plan_execute(plan, 1)
        for p, r0 in zip(probs, results[0]):
            np.testing.assert_array_equal(p["c"].device_view(), r0)
    finally:
        blis.bli_amd_gemm_batch_plan_destroy(plan)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_dropin_repeated_calls(blis, dtype):
    """The drop-in bli_?gemm_batch keeps the plan of a call and reuses it
    when the same arguments come again (the second reuse autotunes it, C
    saved and restored). Three calls on the same C restored in between give
    bit-identical results."""
    ch = "d" if np.dtype(dtype) == np.float64 else "s"
    probs = make_batch(dtype, 64, seed=21 if ch == "d" else 22)
    for p in probs:
        for x in "abc":
            p[x].to_device()
    say(f"{ch}gemm_batch drop-in, 3 calls on {len(probs)} problems")
    args, keep = typed_batch_args(probs, dtype)
    f = getattr(blis, f"bli_{ch}gemm_batch")
    import torch
    results = []
    for call in range(3):
        for p in probs:   # restore C in place (the call's pointers stay the same)
            p["c"].set(p["c0"])
            p["c"].dbuf.copy_(torch.from_numpy(p["c"].buf))
        torch.cuda.synchronize()
        f(*args)
        results.append([p["c"].device_view().copy() for p in probs])
    for call in (1, 2):
        for i in range(len(probs)):
            np.testing.assert_array_equal(results[call][i], results[0][i], err_msg=f"call {call} problem {i}")
