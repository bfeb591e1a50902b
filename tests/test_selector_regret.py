"""The shipped tile selector against exhaustive autotuning (CPU).

The gfx950 tuning database (blis_amd/data/gemm_tuning_gfx950.npz) holds
measured milliseconds of all 42 configurations for 22,000 shapes (m, n, k
log-uniform in [16, 16384], FP32/FP64, random transposes and leading
dimensions). The shipped network (blis_amd/data/selector_gfx950.bin) was
trained on the GPU on the first 20,000; the last 2,000 are held out.

Here the network's picks on the held-out shapes are recomputed with the CPU
model of the network (oracle/mlp.py) and priced with the measured times:
  * time regret against exhaustive autotuning <= 5% (total time and geomean);
  * the gain over the method without the network (the built-in size
    heuristic of the library, gemm_default_tile): at least 2x lower regret.
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
    d = np.load(DATA / "gemm_tuning_gfx950.npz")
    return {k: d[k] for k in d.files}


def _regret(ms, pick):
    best = ms.min(axis=1)
    got = ms[np.arange(len(ms)), pick]
    return got.sum() / best.sum() - 1.0, float(np.exp(np.mean(np.log(got / best))) - 1.0)


def test_database_shape(db):
    assert db["ms"].shape == (22000, M.OUT)
    assert np.all(db["ms"] > 0)
    assert np.array_equal(db["index"], np.arange(22000))


def test_features_match_library(db):
    import blis_amd
    from blis_amd import tune
    L = tune._setup(blis_amd.lib())
    for i in range(0, 22000, 997):
        f = tune.features(L, db["dbl"][i], db["ta"][i], db["tb"][i], *db["dims"][i],
                          db["lda"][i], db["ldb"][i], db["ldc"][i])
        assert np.array_equal(f, db["feats"][i]), i


def test_learned_selector_regret_and_gain(db):
    import blis_amd
    from blis_amd import tune
    p, mu, sd = M.load_weights(DATA / "selector_gfx950.bin")
    ho = slice(20000, 22000)
    ms = db["ms"][ho].astype(np.float64)
    pick = M.predict(p, db["feats"][ho].astype(np.float64), mu, sd)
    tot, geo = _regret(ms, pick)
    L = tune._setup(blis_amd.lib())
    heur = tune.heuristic_picks(L, {k: v[ho] for k, v in db.items() if np.ndim(v)})
    htot, hgeo = _regret(ms, heur)
    print(f"learned: total {tot:.4f} geomean {geo:.4f}; heuristic: total {htot:.4f} geomean {hgeo:.4f}")
    assert tot <= 0.05 and geo <= 0.05
    assert htot >= 2 * tot and hgeo >= 2 * geo
