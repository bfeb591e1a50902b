"""Every benchmark configuration at its full size, through the C ABI, with
checks that do not depend on size:

  * DGEMM/SGEMM m = n = k = 8192 and the skinny m = n = 8192, k = 256, all
    four transpose combinations: the testsuite probe residual
    (testsuite/src/test_gemm.c:303-414) over the whole result, and the
    normwise error of sampled rows against the oracle (bar 4 k u);
  * DGEMV 16384 x 16384, no-transpose and transpose: sampled rows (columns)
    against the oracle's gemv (bar 4 n u, normwise over the sample);
  * DAXPYV and DDOTV at n = 2^28: an exact identity over the whole vector
    (alpha = 0.5 makes alpha*x exact, so y + alpha*x has one rounding
    whichever way it is evaluated), a closed form for the dot product, and
    the oracle on a slice;
  * 10,000 uneven problems (m, n, k log-uniform in [8, 512], random
    transposes, leading dimensions padded to multiples of 8, alpha and beta
    in U[-1, 1]) through bli_dgemm_batch / bli_sgemm_batch and dgemm_batch_:
    the probe residual of every problem, the oracle on sampled problems,
    bitwise equality with single bli_?gemm calls on sampled problems, and
    bitwise equality of the BLAS and typed entry points.

Operands are generated on the device from a seed (torch's generator) and
copied to the host only where the oracle needs them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.conftest import say, unit_roundoff

pytestmark = pytest.mark.gpu

TR = {"n": 0x0, "t": 0x8}


def _torch():
    import torch
    return torch


def _uniform(shape, dtype, seed):
    torch = _torch()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.empty(shape, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=g)
    return x.to(dtype)


def _tdt(ch):
    torch = _torch()
    return torch.float64 if ch == "d" else torch.float32


def _ct(ch):
    return C.c_double if ch == "d" else C.c_float


# ---------------------------------------------------------------------------
# GEMM 8192
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("ch", ["d", "s"])
@pytest.mark.parametrize("k", [8192, 256])
@pytest.mark.parametrize("trans", ["nn", "tn", "nt", "tt"])
def test_gemm_8192(blis, oracle, ch, k, trans):
    torch = _torch()
    m = n = 8192
    ta, tb = trans[0], trans[1]
    say(f"{ch}gemm config 2 {trans} {m}x{n}x{k}")
    dt = _tdt(ch)
    # Column-major storage: a (cols, rows) row-major tensor.
    am, an = (k, m) if ta == "t" else (m, k)
    bm, bn = (n, k) if tb == "t" else (k, n)
    SA = _uniform((an, am), dt, 101)
    SB = _uniform((bn, bm), dt, 102)
    SC = _uniform((n, m), dt, 103)
    alpha, beta = 1.0, 0.5
    # op(A) and op(B) as (rows, cols) views of the storage.
    opA = SA.t() if ta == "n" else SA          # m x k
    opB = SB.t() if tb == "n" else SB          # k x n
    t = _uniform((n,), torch.float64, 104)
    z0 = SC.t().double() @ t
    rows = np.linspace(0, m - 1, 8).astype(np.int64)
    c0_rows = SC.t()[rows].double().cpu().numpy()
    T = _ct(ch)
    getattr(blis, f"bli_{ch}gemm")(TR[ta], TR[tb], m, n, k, C.byref(T(alpha)), SA.data_ptr(),
                                   1, am, SB.data_ptr(), 1, bm, C.byref(T(beta)),
                                   SC.data_ptr(), 1, m)
    torch.cuda.synchronize()
    u = unit_roundoff(np.float64 if ch == "d" else np.float32)
    # Probe residual over the whole result.
    Ad, Bd = opA.double(), opB.double()
    v = SC.t().double() @ t
    z = alpha * (Ad @ (Bd @ t)) + beta * z0
    scale = (abs(alpha) * torch.linalg.norm(Ad) * torch.linalg.norm(Bd)
             + abs(beta) * torch.linalg.norm(SC.double())) * torch.linalg.norm(t)
    resid = float(torch.linalg.norm(v - z) / scale)
    del Ad, Bd
    say(f"  probe residual {resid:.3e}")
    assert resid <= 4 * k * u, resid
    # Sampled rows against the oracle.
    a_rows = np.asfortranarray(opA[rows].cpu().numpy())
    b_full = np.asfortranarray(opB.cpu().numpy())
    got = SC.t()[rows].cpu().numpy()
    ref = np.asfortranarray(c0_rows.astype(a_rows.dtype))
    oracle.gemm("n", "n", alpha, a_rows, b_full, beta, ref, nthreads=16)
    num = np.linalg.norm(got.astype(np.float64) - ref.astype(np.float64))
    den = abs(alpha) * np.linalg.norm(a_rows.astype(np.float64)) * \
        np.linalg.norm(b_full.astype(np.float64)) + abs(beta) * np.linalg.norm(c0_rows)
    err = num / den
    say(f"  sampled-row error vs oracle {err:.3e}")
    assert err <= 4 * k * u, err
    del SA, SB, SC
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# Level 1/2 at the bandwidth sizes
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("trans", ["n", "t"])
def test_dgemv_16384(blis, oracle, trans):
    torch = _torch()
    m = n = 16384
    say(f"dgemv config 2 {trans} {m}x{n}")
    SA = _uniform((n, m), torch.float64, 201)   # column-major m x n
    lx, ly = (n, m) if trans == "n" else (m, n)
    x = _uniform((lx,), torch.float64, 202)
    y = _uniform((ly,), torch.float64, 203)
    y0 = y.cpu().numpy().copy()
    alpha, beta = 1.0, 0.5
    blis.bli_dgemv(TR[trans], 0, m, n, C.byref(C.c_double(alpha)), SA.data_ptr(), 1, m,
                   x.data_ptr(), 1, C.byref(C.c_double(beta)), y.data_ptr(), 1)
    torch.cuda.synchronize()
    got = y.cpu().numpy()
    idx = np.linspace(0, ly - 1, 64).astype(np.int64)
    if trans == "n":
        sub = np.asfortranarray(SA.t()[idx].cpu().numpy())     # 64 x n
    else:
        sub = np.asfortranarray(SA[idx].cpu().numpy().T)       # m x 64 (columns idx)
    xh = x.cpu().numpy()
    ref = y0[idx].copy()
    oracle.gemv(trans, alpha, sub, xh, beta, ref)
    err = np.linalg.norm(got[idx] - ref) / (abs(alpha) * np.linalg.norm(sub) * np.linalg.norm(xh)
                                            + abs(beta) * np.linalg.norm(y0[idx]))
    say(f"  sampled error vs oracle {err:.3e}")
    assert err <= 4 * max(m, n) * unit_roundoff(np.float64), err
    del SA
    torch.cuda.empty_cache()


def test_daxpyv_2_28(blis, oracle):
    torch = _torch()
    n = 1 << 28
    say(f"daxpyv config 2 n={n}")
    x = _uniform((n,), torch.float64, 301)
    y = _uniform((n,), torch.float64, 302)
    want = y + 0.5 * x        # 0.5*x is exact: one rounding, like a fused multiply-add
    blis.bli_daxpyv(0, n, C.byref(C.c_double(0.5)), x.data_ptr(), 1, y.data_ptr(), 1)
    torch.cuda.synchronize()
    assert torch.equal(y, want)
    # The oracle on a slice (same inputs, recomputed from the generator).
    lo, hi = n // 3, n // 3 + (1 << 20)
    xs = x[lo:hi].cpu().numpy()
    ys = _uniform((n,), torch.float64, 302)[lo:hi].cpu().numpy()
    oracle.axpyv(0.5, xs, ys)
    np.testing.assert_array_equal(y[lo:hi].cpu().numpy(), ys)
    del x, y, want
    torch.cuda.empty_cache()


def test_ddotv_2_28(blis, oracle):
    torch = _torch()
    n = 1 << 28
    say(f"ddotv config 2 n={n}")
    rho = C.c_double(0.0)
    # Closed form: sum of n products 1 * 0.25 is exactly n / 4.
    ones = torch.ones(n, dtype=torch.float64, device="cuda")
    quarter = torch.full((n,), 0.25, dtype=torch.float64, device="cuda")
    blis.bli_ddotv(0, 0, n, ones.data_ptr(), 1, quarter.data_ptr(), 1, C.addressof(rho))
    assert rho.value == n / 4
    del ones, quarter
    x = _uniform((n,), torch.float64, 401)
    y = _uniform((n,), torch.float64, 402)
    blis.bli_ddotv(0, 0, n, x.data_ptr(), 1, y.data_ptr(), 1, C.addressof(rho))
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    ref = oracle.dotv(xh, yh)
    # Extended-precision value, in slices.
    exact = np.longdouble(0)
    absum = 0.0
    for s in range(0, n, 1 << 24):
        p = xh[s:s + (1 << 24)].astype(np.longdouble) * yh[s:s + (1 << 24)]
        exact += p.sum()
        absum += float(np.abs(p).sum())
    u = unit_roundoff(np.float64)
    say(f"  dot {rho.value!r} oracle {ref!r} extended {float(exact)!r}")
    # Both sums are blocked: a few hundred sequential additions per lane,
    # then fixed-order trees (error bound (steps) * u * sum |x_i y_i|).
    assert abs(rho.value - float(exact)) <= 1024 * u * absum
    assert abs(ref - float(exact)) <= 1024 * u * absum
    del x, y
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# 10,000 uneven problems
# ---------------------------------------------------------------------------

def _batch(ch, count=10000, seed=42):
    torch = _torch()
    rng = np.random.default_rng(seed)
    dims = np.rint(np.exp(rng.uniform(np.log(8), np.log(512), (count, 3)))).astype(np.int64)
    ta = rng.integers(0, 2, count)
    tb = rng.integers(0, 2, count)
    alpha = rng.uniform(-1, 1, count)
    beta = rng.uniform(-1, 1, count)
    es = 8 if ch == "d" else 4
    pad = lambda v: (v + 7) // 8 * 8
    m, n, k = dims[:, 0], dims[:, 1], dims[:, 2]
    arows = np.where(ta == 1, k, m)
    acols = np.where(ta == 1, m, k)
    brows = np.where(tb == 1, n, k)
    bcols = np.where(tb == 1, k, n)
    lda, ldb, ldc = pad(arows), pad(brows), pad(m)
    al = 256 // es
    offs = lambda sizes: np.concatenate([[0], np.cumsum((sizes + al - 1) // al * al)])
    oa, ob, oc = offs(lda * acols), offs(ldb * bcols), offs(ldc * n)
    A = _uniform((int(oa[-1]),), _tdt(ch), seed + 1)
    B = _uniform((int(ob[-1]),), _tdt(ch), seed + 2)
    Cm = _uniform((int(oc[-1]),), _tdt(ch), seed + 3)
    return dict(count=count, m=m, n=n, k=k, ta=ta, tb=tb, alpha=alpha, beta=beta, lda=lda,
                ldb=ldb, ldc=ldc, arows=arows, acols=acols, brows=brows, bcols=bcols, oa=oa,
                ob=ob, oc=oc, A=A, B=B, C=Cm, es=es)


def _typed_args(b, ch):
    cnt = b["count"]
    T = _ct(ch)
    I = C.c_int64 * cnt
    Pp = C.c_void_p * cnt
    es = b["es"]
    arrs = [(C.c_int * cnt)(*(b["ta"] * 8).tolist()), (C.c_int * cnt)(*(b["tb"] * 8).tolist()),
            I(*b["m"].tolist()), I(*b["n"].tolist()), I(*b["k"].tolist()),
            (T * cnt)(*b["alpha"].tolist()),
            Pp(*[b["A"].data_ptr() + int(o) * es for o in b["oa"][:-1]]), I(*([1] * cnt)),
            I(*b["lda"].tolist()),
            Pp(*[b["B"].data_ptr() + int(o) * es for o in b["ob"][:-1]]), I(*([1] * cnt)),
            I(*b["ldb"].tolist()),
            (T * cnt)(*b["beta"].tolist()),
            Pp(*[b["C"].data_ptr() + int(o) * es for o in b["oc"][:-1]]), I(*([1] * cnt)),
            I(*b["ldc"].tolist())]
    gs = I(*([1] * cnt))
    return [C.cast(x, C.c_void_p) for x in arrs] + [cnt, C.cast(gs, C.c_void_p)], (arrs, gs)


def _f77_args(b, ch):
    cnt = b["count"]
    T = _ct(ch)
    Fi = C.c_int32 * cnt
    Pp = C.c_void_p * cnt
    es = b["es"]
    chars = lambda t: (C.c_char * cnt)(*[b"T" if v else b"N" for v in t])
    arrs = [chars(b["ta"]), chars(b["tb"]), Fi(*b["m"].tolist()), Fi(*b["n"].tolist()),
            Fi(*b["k"].tolist()), (T * cnt)(*b["alpha"].tolist()),
            Pp(*[b["A"].data_ptr() + int(o) * es for o in b["oa"][:-1]]), Fi(*b["lda"].tolist()),
            Pp(*[b["B"].data_ptr() + int(o) * es for o in b["ob"][:-1]]), Fi(*b["ldb"].tolist()),
            (T * cnt)(*b["beta"].tolist()),
            Pp(*[b["C"].data_ptr() + int(o) * es for o in b["oc"][:-1]]), Fi(*b["ldc"].tolist())]
    gc = C.c_int32(cnt)
    gs = Fi(*([1] * cnt))
    args = [C.cast(x, C.c_void_p) for x in arrs] + [C.cast(C.pointer(gc), C.c_void_p),
                                                   C.cast(gs, C.c_void_p)]
    return args, (arrs, gc, gs)


def _view(buf, off, rows, cols, ld):
    es = buf.itemsize
    return np.lib.stride_tricks.as_strided(buf[int(off):], shape=(int(rows), int(cols)),
                                           strides=(es, int(ld) * es))


@pytest.mark.parametrize("ch", ["d", "s"])
def test_batch_10000(blis, oracle, ch):
    torch = _torch()
    say(f"{ch}gemm_batch config 4: 10,000 problems")
    b = _batch(ch)
    C0 = b["C"].clone()
    args, keep = _typed_args(b, ch)
    getattr(blis, f"bli_{ch}gemm_batch")(*args)
    torch.cuda.synchronize()
    out_typed = b["C"].clone()
    # The BLAS symbol on the same inputs.
    if ch == "d":
        say("  dgemm_batch_ on the same problems")
        b["C"].copy_(C0)
        fargs, fkeep = _f77_args(b, ch)
        blis.dgemm_batch_.argtypes = [C.c_void_p] * 15
        blis.dgemm_batch_(*fargs)
        torch.cuda.synchronize()
        assert torch.equal(b["C"], out_typed)
    Ah, Bh = b["A"].double().cpu().numpy(), b["B"].double().cpu().numpy()
    Ch, C0h = out_typed.double().cpu().numpy(), C0.double().cpu().numpy()
    u = unit_roundoff(np.float64 if ch == "d" else np.float32)
    rng = np.random.default_rng(7)
    worst = 0.0
    for i in range(b["count"]):
        m, n, k = int(b["m"][i]), int(b["n"][i]), int(b["k"][i])
        A = _view(Ah, b["oa"][i], b["arows"][i], b["acols"][i], b["lda"][i])
        Bm = _view(Bh, b["ob"][i], b["brows"][i], b["bcols"][i], b["ldb"][i])
        Ci = _view(Ch, b["oc"][i], m, n, b["ldc"][i])
        C0i = _view(C0h, b["oc"][i], m, n, b["ldc"][i])
        opA = A.T if b["ta"][i] else A
        opB = Bm.T if b["tb"][i] else Bm
        t = rng.uniform(-1, 1, n)
        al, be = b["alpha"][i], b["beta"][i]
        v = Ci @ t
        z = al * (opA @ (opB @ t)) + be * (C0i @ t)
        scale = (abs(al) * np.linalg.norm(opA) * np.linalg.norm(opB)
                 + abs(be) * np.linalg.norm(C0i)) * np.linalg.norm(t)
        r = np.linalg.norm(v - z) / scale / (4 * k * u)
        worst = max(worst, r)
        assert r <= 1.0, (i, m, n, k, r)
    say(f"  worst probe residual / (4 k u) over all problems {worst:.3e}")
    # Sampled problems: the oracle, and single bli_?gemm calls bitwise.
    T = _ct(ch)
    npdt = np.float64 if ch == "d" else np.float32
    es = b["es"]
    for i in rng.choice(b["count"], 40, replace=False):
        m, n, k = int(b["m"][i]), int(b["n"][i]), int(b["k"][i])
        ta, tb = "t" if b["ta"][i] else "n", "t" if b["tb"][i] else "n"
        A = _view(Ah, b["oa"][i], b["arows"][i], b["acols"][i], b["lda"][i]).astype(npdt)
        Bm = _view(Bh, b["ob"][i], b["brows"][i], b["bcols"][i], b["ldb"][i]).astype(npdt)
        C0i = _view(C0h, b["oc"][i], m, n, b["ldc"][i]).astype(npdt)
        ref = np.asfortranarray(C0i.copy())
        oracle.gemm(ta, tb, float(b["alpha"][i]), A, Bm, float(b["beta"][i]), ref)
        got = _view(Ch, b["oc"][i], m, n, b["ldc"][i])
        num = np.linalg.norm(got - ref.astype(np.float64))
        den = abs(b["alpha"][i]) * np.linalg.norm(A.astype(np.float64)) * \
            np.linalg.norm(Bm.astype(np.float64)) + abs(b["beta"][i]) * np.linalg.norm(C0i.astype(np.float64))
        assert num / den <= 4 * k * u, (i, num / den)
        single = C0[int(b["oc"][i]):int(b["oc"][i + 1])].clone()
        getattr(blis, f"bli_{ch}gemm")(TR[ta], TR[tb], m, n, k, C.byref(T(b["alpha"][i])),
                                       b["A"].data_ptr() + int(b["oa"][i]) * es, 1,
                                       int(b["lda"][i]),
                                       b["B"].data_ptr() + int(b["ob"][i]) * es, 1,
                                       int(b["ldb"][i]), C.byref(T(b["beta"][i])),
                                       single.data_ptr(), 1, int(b["ldc"][i]))
        torch.cuda.synchronize()
        batch_piece = out_typed[int(b["oc"][i]):int(b["oc"][i + 1])]
        ldc = int(b["ldc"][i])
        s2 = single[:ldc * n].view(n, ldc)[:, :m]
        b2 = batch_piece[:ldc * n].view(n, ldc)[:, :m]
        assert torch.equal(s2, b2), (i, m, n, k)
    del b, C0, out_typed
    torch.cuda.empty_cache()
