"""The shipped tile selector against exhaustive autotuning (CPU).

Two gfx950 tuning databases hold measured milliseconds of all 42
configurations per shape: blis_amd/data/gemm_tuning_gfx950.npz, 22,000
shapes with m, n, k log-uniform in [16, 16384] (FP32/FP64, random
transposes and leading dimensions), and gemm_tuning_round_gfx950.npz, 3,000
shapes whose dimensions are multiples of 64 (a quarter powers of two). The
shipped network (blis_amd/data/selector_gfx950.bin) was trained on the GPU
without the last 10% of each database, which are held out here.

The library's learned-mode picks on the held-out shapes (the network's
argmin, recomputed with the CPU model oracle/mlp.py, and the FP64
large-tile band kept on the size heuristic's tile, tune.guarded) are priced
with the measured times:
  * time regret against exhaustive autotuning <= 5% (total time and
    geomean);
  * the gain over the built-in size heuristic (gemm_default_tile): at least
    1.5x lower total-time regret and 2x lower geomean regret.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp as M

DATA = Path(__file__).resolve().parent.parent / "blis_amd" / "data"


@pytest.fixture(scope="module")
def db():
    from blis_amd import tune
    return tune.load_db([DATA / "gemm_tuning_gfx950.npz", DATA / "gemm_tuning_round_gfx950.npz"],
                        0.1)


def _regret(ms, pick):
    best = ms.min(axis=1)
    got = ms[np.arange(len(ms)), pick]
    return got.sum() / best.sum() - 1.0, float(np.exp(np.mean(np.log(got / best))) - 1.0)


def test_database_shape(db):
    assert db["ms"].shape == (25000, M.OUT)
    assert np.all(db["ms"] > 0)
    rand, rnd = db["src"] == 0, db["src"] == 1
    assert np.array_equal(db["index"][rand], np.arange(22000))
    assert np.array_equal(db["index"][rnd], np.arange(3000))
    assert np.all(db["dims"][rnd] % 64 == 0)
    assert db["held"][rand].sum() == 2200 and db["held"][rnd].sum() == 300


def test_features_match_library(db):
    import blis_amd
    from blis_amd import tune
    L = tune._setup(blis_amd.lib())
    for i in range(0, 25000, 997):
        f = tune.features(L, db["dbl"][i], db["ta"][i], db["tb"][i], *db["dims"][i],
                          db["lda"][i], db["ldb"][i], db["ldc"][i])
        assert np.array_equal(f, db["feats"][i]), i


def test_learned_selector_regret_and_gain(db):
    import blis_amd
    from blis_amd import tune
    p, mu, sd = M.load_weights(DATA / "selector_gfx950.bin")
    ho = db["held"]
    sub = {k: v[ho] for k, v in db.items()}
    ms = sub["ms"].astype(np.float64)
    L = tune._setup(blis_amd.lib())
    heur = tune.heuristic_picks(L, sub)
    pick = tune.guarded(sub, M.predict(p, sub["feats"].astype(np.float64), mu, sd), heur)
    tot, geo = _regret(ms, pick)
    htot, hgeo = _regret(ms, heur)
    print(f"learned: total {tot:.4f} geomean {geo:.4f}; heuristic: total {htot:.4f} geomean {hgeo:.4f}")
    assert tot <= 0.05 and geo <= 0.05
    assert htot >= 1.5 * tot and hgeo >= 2 * geo
