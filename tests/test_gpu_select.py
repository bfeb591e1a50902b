"""Tile selector on the GPU: MLP kernels against the CPU model, the
configuration space's bit-identity, the autotuner and the selector modes.

Bars:
  * loss and gradient from the HIP kernels match oracle/mlp.py (float64) to
    1e-4 relative (FP32 arithmetic over at most 64-term sums);
  * forward scores match to 1e-5 relative; predictions are the argmin;
  * training is deterministic (same seed -> bit-identical parameters) and
    lowers the loss;
  * every configuration gives bit-identical GEMM output;
  * the learned and exhaustive selectors give output bit-identical to the
    heuristic's;
  * the learned selector is the default mode, and its host-side inference
    picks what the HIP inference kernels pick (with the FP64 large-tile
    band kept on the heuristic's tile).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import mlp as M
from tests.conftest import Mat, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from blis_amd import lib, tune
    return tune._setup(lib())


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _data(seed, n):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 3, (n, M.IN)).astype(np.float32)
    t = rng.uniform(0, 3, (n, M.OUT)).astype(np.float32)
    p = rng.normal(0, 0.25, M.NPARAM).astype(np.float32)
    musd = np.concatenate([rng.normal(0, 1, M.IN), rng.uniform(0.5, 2, M.IN)]).astype(np.float32)
    return f, t, p, musd


def test_param_count(L):
    assert L.bli_amd_mlp_param_count() == M.NPARAM


@pytest.mark.parametrize("n", [1, 37, 256])
def test_gradient_matches_oracle(L, n):
    say(f"mlp gradient n={n}")
    f, t, p, musd = _data(n, n)
    net = L.bli_amd_mlp_create(1)
    L.bli_amd_mlp_set_params(net, fp(p), fp(musd))
    g = np.zeros(M.NPARAM, dtype=np.float32)
    loss = L.bli_amd_mlp_loss_and_grad(net, fp(f), fp(t), n, fp(g))
    L.bli_amd_mlp_destroy(net)
    lo, go = M.loss_and_grad(p, f, t, musd[:M.IN].astype(np.float64), musd[M.IN:].astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    scale = np.abs(go).max()
    assert np.max(np.abs(g - go)) <= 1e-4 * scale, np.max(np.abs(g - go)) / scale


def test_forward_and_predict_match_oracle(L):
    say("mlp forward/predict")
    n = 500
    f, t, p, musd = _data(9, n)
    net = L.bli_amd_mlp_create(1)
    L.bli_amd_mlp_set_params(net, fp(p), fp(musd))
    s = np.zeros((n, M.OUT), dtype=np.float32)
    best = np.zeros(n, dtype=np.int32)
    L.bli_amd_mlp_predict(net, fp(f), n, fp(s), best.ctypes.data_as(C.POINTER(C.c_int)))
    L.bli_amd_mlp_destroy(net)
    y = M.forward(p, f, musd[:M.IN].astype(np.float64), musd[M.IN:].astype(np.float64))[3]
    assert np.allclose(s, y, rtol=1e-5, atol=1e-5 * np.abs(y).max())
    assert np.array_equal(best, np.argmin(s, axis=1))


def test_training_is_deterministic_and_learns(L):
    say("mlp training")
    n = 1000
    rng = np.random.default_rng(4)
    f = rng.normal(0, 1, (n, M.IN)).astype(np.float32)
    # a learnable target: smooth functions of the features
    w = rng.normal(0, 1, (M.IN, M.OUT))
    t = np.abs(np.tanh(f @ w / 4)).astype(np.float32)
    musd = np.concatenate([np.zeros(M.IN), np.ones(M.IN)]).astype(np.float32)
    params = []
    for _ in range(2):
        net = L.bli_amd_mlp_create(11)
        p0 = np.zeros(M.NPARAM, dtype=np.float32)
        L.bli_amd_mlp_get_params(net, fp(p0), None)
        L.bli_amd_mlp_set_params(net, fp(p0), fp(musd))
        l0 = L.bli_amd_mlp_loss_and_grad(net, fp(f), fp(t), n, None)
        l1 = L.bli_amd_mlp_train(net, fp(f), fp(t), n, 30, 64, 3e-3, 5)
        p1 = np.zeros(M.NPARAM, dtype=np.float32)
        L.bli_amd_mlp_get_params(net, fp(p1), None)
        params.append(p1)
        L.bli_amd_mlp_destroy(net)
        assert l1 < 0.25 * l0, (l0, l1)
    assert np.array_equal(params[0], params[1])


def test_save_load_round_trip(L, tmp_path):
    say("mlp save/load")
    f, t, p, musd = _data(2, 8)
    net = L.bli_amd_mlp_create(1)
    L.bli_amd_mlp_set_params(net, fp(p), fp(musd))
    path = str(tmp_path / "w.bin").encode()
    assert L.bli_amd_mlp_save(net, path) == 0
    net2 = L.bli_amd_mlp_load(path)
    assert net2
    p2 = np.zeros(M.NPARAM, dtype=np.float32)
    ms2 = np.zeros(2 * M.IN, dtype=np.float32)
    L.bli_amd_mlp_get_params(net2, fp(p2), fp(ms2))
    assert np.array_equal(p2, p) and np.array_equal(ms2, musd)
    po, muo, sdo = M.load_weights(path.decode())
    assert np.array_equal(po.astype(np.float32), p)
    L.bli_amd_mlp_destroy(net)
    L.bli_amd_mlp_destroy(net2)


def _gemm(blis, ch, m, n, k, a, b, c, ta=0, tb=0):
    T = C.c_double if ch == "d" else C.c_float
    getattr(blis, f"bli_{ch}gemm")(ta, tb, m, n, k, C.byref(T(0.75)), a.ptr, a.rs, a.cs, b.ptr,
                                   b.rs, b.cs, C.byref(T(-0.5)), c.ptr, c.rs, c.cs)


@pytest.mark.parametrize("ch", ["d", "s"])
def test_every_configuration_bit_identical(blis, L, ch):
    dt = np.float64 if ch == "d" else np.float32
    m, n, k = 203, 150, 97
    a = Mat(k, m, dtype=dt, seed=1).to_device()   # A^T stored: transa
    b = Mat(k, n, dtype=dt, seed=2).to_device()
    ref = None
    blis.bli_amd_gemm_set_config.argtypes = [C.c_int] * 4
    try:
        for cfg in range(L.bli_amd_gemm_config_count(1)):
            tile, grp, xcd = C.c_int(), C.c_int(), C.c_int()
            L.bli_amd_gemm_config_describe.argtypes = [C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 6
            L.bli_amd_gemm_config_describe(int(ch == "d"), cfg, C.byref(tile), C.byref(grp),
                                           C.byref(xcd), None, None, None)
            say(f"config {cfg} tile {tile.value} group {grp.value} xcd {xcd.value} {ch}")
            blis.bli_amd_gemm_set_config(tile.value, grp.value, xcd.value, 1)
            c = Mat(m, n, dtype=dt, seed=3).to_device()
            _gemm(blis, ch, m, n, k, a, b, c, ta=8)
            out = c.device_view().copy()
            if ref is None:
                ref = out
            else:
                assert np.array_equal(out, ref), cfg
    finally:
        blis.bli_amd_gemm_set_config(-1, 8, 1, 1)


def test_autotune_returns_positive_times(L):
    say("autotune 300x200x100")
    ms = np.zeros(L.bli_amd_gemm_config_count(1), dtype=np.float32)
    assert L.bli_amd_gemm_autotune(1, 0, 8, 300, 200, 100, 304, 201, 300, fp(ms)) == 0
    assert np.all(ms > 0) and np.all(np.isfinite(ms))
    assert L.bli_amd_gemm_autotune(1, 0, 0, 300, 200, 100, 299, 100, 300, fp(ms)) == -1


@pytest.mark.parametrize("mode", [1, 2])
def test_selector_modes_bit_identical_to_heuristic(blis, L, mode, tmp_path):
    say(f"selector mode {mode}")
    L.bli_amd_gemm_set_selector.argtypes = [C.c_int, C.c_void_p]
    L.bli_amd_gemm_selected_config.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_int64] * 6
    m, n, k = 130, 77, 300
    a = Mat(m, k, seed=5).to_device()
    b = Mat(k, n, seed=6).to_device()
    c0 = Mat(m, n, seed=7).to_device()
    _gemm(blis, "d", m, n, k, a, b, c0)
    ref = c0.device_view().copy()
    net = None
    if mode == 1:
        # a freshly initialised network (the shipped weights are tested
        # separately): any prediction is a valid configuration
        net = L.bli_amd_mlp_create(3)
    prev = L.bli_amd_gemm_get_selector()
    try:
        L.bli_amd_gemm_set_selector(mode, net)
        cfg = L.bli_amd_gemm_selected_config(1, 0, 0, m, n, k, m, k, m)
        assert 0 <= cfg < L.bli_amd_gemm_config_count(1)
        c1 = Mat(m, n, seed=7).to_device()
        _gemm(blis, "d", m, n, k, a, b, c1)
        assert np.array_equal(c1.device_view(), ref)
    finally:
        L.bli_amd_gemm_set_selector(prev, None)
        if net:
            L.bli_amd_mlp_destroy(net)


def test_shipped_weights_predict_like_the_model(L):
    """The HIP inference kernels and the CPU model agree on the held-out
    shapes of the tuning database (FP32 vs FP64 arithmetic may flip a pick
    only between configurations with near-equal scores)."""
    from pathlib import Path
    say("shipped selector predictions")
    data = Path(__file__).resolve().parent.parent / "blis_amd" / "data"
    db = np.load(data / "gemm_tuning_gfx950.npz")
    feats = np.ascontiguousarray(db["feats"][20000:], dtype=np.float32)
    ms = db["ms"][20000:].astype(np.float64)
    net = L.bli_amd_mlp_load(str(data / "selector_gfx950.bin").encode())
    assert net
    n = len(feats)
    best = np.zeros(n, dtype=np.int32)
    L.bli_amd_mlp_predict(net, fp(feats), n, None, best.ctypes.data_as(C.POINTER(C.c_int)))
    L.bli_amd_mlp_destroy(net)
    p, mu, sd = M.load_weights(data / "selector_gfx950.bin")
    ref = M.predict(p, feats.astype(np.float64), mu, sd)
    assert np.mean(best == ref) >= 0.995
    rows = np.arange(n)
    tot_gpu = ms[rows, best].sum() / ms.min(axis=1).sum() - 1
    tot_ref = ms[rows, ref].sum() / ms.min(axis=1).sum() - 1
    assert abs(tot_gpu - tot_ref) <= 0.002 and tot_gpu <= 0.05


def test_learned_mode_with_shipped_weights(blis, L):
    say("selector mode 1 with shipped weights")
    L.bli_amd_gemm_set_selector.argtypes = [C.c_int, C.c_void_p]
    L.bli_amd_gemm_selected_config.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_int64] * 6
    m, n, k = 1000, 333, 517
    a = Mat(m, k, seed=15).to_device()
    b = Mat(k, n, seed=16).to_device()
    c0 = Mat(m, n, seed=17).to_device()
    _gemm(blis, "d", m, n, k, a, b, c0)
    ref = c0.device_view().copy()
    prev = L.bli_amd_gemm_get_selector()
    try:
        L.bli_amd_gemm_set_selector(1, None)
        assert L.bli_amd_gemm_get_selector() == 1
        cfg = L.bli_amd_gemm_selected_config(1, 0, 0, m, n, k, m, k, m)
        assert 0 <= cfg < 42
        c1 = Mat(m, n, seed=17).to_device()
        _gemm(blis, "d", m, n, k, a, b, c1)
        assert np.array_equal(c1.device_view(), ref)
    finally:
        L.bli_amd_gemm_set_selector(prev, None)


def test_exhaustive_mode_general_stride_b(blis, L):
    """Exhaustive selection for a product whose B has general strides, at a
    size where the autotuner's large-tile configurations transpose their own
    operands through the library's scratch: the configuration is chosen
    before the caller's B is packed, so the packed copy is not overwritten
    and the result is bit-identical to the heuristic's."""
    say("selector mode 2, general-stride B, 4096x3000x3000")
    L.bli_amd_gemm_set_selector.argtypes = [C.c_int, C.c_void_p]
    m, n, k = 4096, 3000, 3000
    a = Mat(m, k, seed=25).to_device()
    b = Mat(k, n, rs=2, cs=2 * k, seed=26).to_device()
    c0 = Mat(m, n, seed=27).to_device()
    al, be = C.c_double(1.0), C.c_double(0.5)

    def run(c):
        blis.bli_dgemm(0, 0, m, n, k, C.byref(al), a.ptr, a.rs, a.cs, b.ptr, b.rs, b.cs,
                       C.byref(be), c.ptr, c.rs, c.cs)
    run(c0)
    ref = c0.device_view().copy()
    prev = L.bli_amd_gemm_get_selector()
    try:
        L.bli_amd_gemm_set_selector(2, None)
        c1 = Mat(m, n, seed=27).to_device()
        run(c1)
        assert np.array_equal(c1.device_view(), ref)
    finally:
        L.bli_amd_gemm_set_selector(prev, None)


def test_learned_selector_is_default_and_host_inference_matches(L):
    """The default mode is the learned selector (BLIS_AMD_SELECTOR unset).
    Its host-side inference (bli_amd_gemm_selected_config) picks, on 400
    held-out shapes of both tuning databases, the configuration that the
    HIP inference kernels pick with the shipped weights, after the FP64
    large-tile band rule (tune.guarded)."""
    import os
    from pathlib import Path
    from blis_amd import tune
    say("learned selector default; host vs HIP inference")
    if os.environ.get("BLIS_AMD_SELECTOR") or os.environ.get("BLIS_AMD_SELECTOR_WEIGHTS"):
        pytest.skip("selector chosen through the environment")
    L.bli_amd_gemm_get_selector.restype = C.c_int
    L.bli_amd_gemm_selected_config.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_int64] * 6
    L.bli_amd_gemm_selected_config.restype = C.c_int
    assert L.bli_amd_gemm_get_selector() == 1
    data = Path(__file__).resolve().parent.parent / "blis_amd" / "data"
    db = tune.load_db([data / "gemm_tuning_gfx950.npz", data / "gemm_tuning_round_gfx950.npz"], 0.1)
    rows = np.where(db["held"])[0][::6][:400]
    sub = {k: v[rows] for k, v in db.items()}
    feats = np.ascontiguousarray(sub["feats"], dtype=np.float32)
    net = L.bli_amd_mlp_load(str(data / "selector_gfx950.bin").encode())
    assert net
    dev = np.zeros(len(rows), dtype=np.int32)
    L.bli_amd_mlp_predict(net, fp(feats), len(rows), None, dev.ctypes.data_as(C.POINTER(C.c_int)))
    L.bli_amd_mlp_destroy(net)
    want = tune.guarded(sub, dev, tune.heuristic_picks(L, sub))
    got = np.array([L.bli_amd_gemm_selected_config(int(sub["dbl"][i]), 8 * int(sub["ta"][i]),
                                                   8 * int(sub["tb"][i]), *map(int, sub["dims"][i]),
                                                   int(sub["lda"][i]), int(sub["ldb"][i]),
                                                   int(sub["ldc"][i]))
                    for i in range(len(rows))])
    assert np.array_equal(got, want), np.where(got != want)[0][:10]
