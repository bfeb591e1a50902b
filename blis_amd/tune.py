"""Autotuning database and learned tile selector for the MI355X GEMM.

    python -m blis_amd.tune sweep --out DB.npz [--shapes N] [--start S] [--seed X]
        Time every tile/schedule configuration (bli_amd_gemm_autotune) on
        random GEMM shapes: m, n, k log-uniform in [16, 16384], FP32 or
        FP64, random transA/transB, leading dimensions natural, padded to 64
        or misaligned by one element. Writes shapes, features and per-config
        milliseconds to an .npz file.

    python -m blis_amd.tune sweep --kind round ...
        The same on "round" shapes: every dimension a multiple of 64, a
        quarter of them powers of two.

    python -m blis_amd.tune train --db DB.npz [...] --out WEIGHTS.bin
        Train the selector MLP on the GPU (bli_amd_mlp_train) on all but the
        last --holdout-frac of each database's shapes, then report its time
        regret against the exhaustive choice on the held-out shapes next to
        the built-in heuristic's.

Targets are y_c = log2(t_c / t_best) per configuration c. The selector picks
argmin of the predicted y.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import lib

NFEAT = 16
DATA = Path(__file__).resolve().parent / "data"


def _setup(L):
    I64 = C.c_int64
    PF = C.POINTER(C.c_float)
    L.bli_amd_gemm_config_count.argtypes = [C.c_int]
    L.bli_amd_gemm_config_count.restype = C.c_int
    L.bli_amd_gemm_autotune.argtypes = [C.c_int, C.c_int, C.c_int, I64, I64, I64, I64, I64, I64, PF]
    L.bli_amd_gemm_autotune.restype = C.c_int
    L.bli_amd_gemm_features.argtypes = [C.c_int, C.c_int, C.c_int, I64, I64, I64, I64, I64, I64, PF]
    L.bli_amd_gemm_features.restype = None
    L.bli_amd_gemm_default_tile.argtypes = [C.c_int, C.c_longlong, C.c_longlong, C.c_longlong]
    L.bli_amd_gemm_default_tile.restype = C.c_int
    L.bli_amd_mlp_create.argtypes = [C.c_uint]
    L.bli_amd_mlp_create.restype = C.c_void_p
    L.bli_amd_mlp_destroy.argtypes = [C.c_void_p]
    L.bli_amd_mlp_train.argtypes = [C.c_void_p, PF, PF, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint]
    L.bli_amd_mlp_train.restype = C.c_double
    L.bli_amd_mlp_predict.argtypes = [C.c_void_p, PF, C.c_int, PF, C.POINTER(C.c_int)]
    L.bli_amd_mlp_predict.restype = None
    L.bli_amd_mlp_set_params.argtypes = [C.c_void_p, PF, PF]
    L.bli_amd_mlp_get_params.argtypes = [C.c_void_p, PF, PF]
    L.bli_amd_mlp_save.argtypes = [C.c_void_p, C.c_char_p]
    L.bli_amd_mlp_save.restype = C.c_int
    L.bli_amd_mlp_load.argtypes = [C.c_char_p]
    L.bli_amd_mlp_load.restype = C.c_void_p
    L.bli_amd_mlp_param_count.restype = C.c_int
    L.bli_amd_mlp_loss_and_grad.argtypes = [C.c_void_p, PF, PF, C.c_int, PF]
    L.bli_amd_mlp_loss_and_grad.restype = C.c_double
    return L


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def random_shapes(n: int, seed: int, kind: str = "random"):
    """Shapes of the selector workload (deterministic for a seed).

    kind "random": m, n, k log-uniform in [16, 16384]. kind "round":
    multiples of 64, 64 * round(e^U(0, ln 256)) in [64, 16384], a quarter of
    them snapped to the nearest power of two (the sizes callers most often
    use, which random integers almost never hit: a random shape has all
    three dimensions divisible by 64 about once in 2^17 draws)."""
    rng = np.random.default_rng(seed)
    if kind == "round":
        q = np.rint(np.exp(rng.uniform(0.0, np.log(256), (n, 3)))).astype(np.int64)
        p2 = 2 ** np.rint(np.log2(q)).astype(np.int64)
        q = np.where(rng.uniform(size=(n, 3)) < 0.25, p2, q)
        dims = 64 * np.clip(q, 1, 256)
    else:
        dims = np.rint(np.exp(rng.uniform(np.log(16), np.log(16384), (n, 3)))).astype(np.int64)
    dbl = rng.integers(0, 2, n)
    ta = rng.integers(0, 2, n)
    tb = rng.integers(0, 2, n)
    ldkind = rng.choice(3, size=(n, 3), p=[0.5, 0.3, 0.2])   # natural, pad 64, +1

    def ld(rows, kd):
        return np.where(kd == 0, rows, np.where(kd == 1, (rows + 63) // 64 * 64, rows + 1))
    m, nn, k = dims[:, 0], dims[:, 1], dims[:, 2]
    lda = ld(np.where(ta == 1, k, m), ldkind[:, 0])
    ldb = ld(np.where(tb == 1, nn, k), ldkind[:, 1])
    ldc = ld(m, ldkind[:, 2])
    return {"dims": dims, "dbl": dbl, "ta": ta, "tb": tb, "lda": lda, "ldb": ldb, "ldc": ldc}


def features(L, dbl, ta, tb, m, n, k, lda, ldb, ldc) -> np.ndarray:
    out = np.zeros(NFEAT, dtype=np.float32)
    L.bli_amd_gemm_features(int(dbl), 8 * int(ta), 8 * int(tb), int(m), int(n), int(k),
                            int(lda), int(ldb), int(ldc), _fp(out))
    return out


def sweep(out: Path, shapes: int, start: int, count: int, seed: int, max_seconds: float,
          kind: str = "random"):
    L = _setup(lib())
    ncfg = L.bli_amd_gemm_config_count(1)
    S = random_shapes(shapes, seed, kind)
    stop = min(shapes, start + count)
    idx, feats, ms = [], [], []
    t0 = time.time()
    for i in range(start, stop):
        m, n, k = (int(x) for x in S["dims"][i])
        args = (int(S["dbl"][i]), 8 * int(S["ta"][i]), 8 * int(S["tb"][i]), m, n, k,
                int(S["lda"][i]), int(S["ldb"][i]), int(S["ldc"][i]))
        t = np.zeros(ncfg, dtype=np.float32)
        rc = L.bli_amd_gemm_autotune(*args, _fp(t))
        if rc != 0:
            raise RuntimeError(f"autotune rejected shape {i}: {args}")
        f = np.zeros(NFEAT, dtype=np.float32)
        L.bli_amd_gemm_features(*args, _fp(f))
        idx.append(i)
        feats.append(f)
        ms.append(t)
        done = i - start + 1
        if done % 50 == 0 or i == stop - 1:
            print(f"sweep: {done}/{stop - start} shapes, {time.time() - t0:.0f} s", flush=True)
        if done % 2000 == 0:
            _save(out, S, idx, feats, ms, seed)   # keep partial results
        if time.time() - t0 > max_seconds:
            print(f"sweep: time budget reached after {done} shapes", flush=True)
            break
    _save(out, S, idx, feats, ms, seed)
    print(f"sweep: wrote {len(idx)} shapes to {out}", flush=True)


def _save(out, S, idx, feats, ms, seed):
    idx = np.array(idx, dtype=np.int64)
    np.savez_compressed(out, index=idx, dims=S["dims"][idx], dbl=S["dbl"][idx], ta=S["ta"][idx],
                        tb=S["tb"][idx], lda=S["lda"][idx], ldb=S["ldb"][idx], ldc=S["ldc"][idx],
                        feats=np.array(feats), ms=np.array(ms), seed=seed)


def load_db(paths, holdout_frac: float = 0.0):
    """Concatenate databases, each in shape order. Returns the merged arrays
    plus "src" (database number) and "held" (True for the last holdout_frac
    of each database's shapes)."""
    keys = ["index", "dims", "dbl", "ta", "tb", "lda", "ldb", "ldc", "feats", "ms"]
    parts = []
    for i, path in enumerate(paths):
        p = np.load(path)
        order = np.argsort(p["index"], kind="stable")
        part = {k: p[k][order] for k in keys}
        n = len(order)
        part["src"] = np.full(n, i, dtype=np.int64)
        part["held"] = np.arange(n) >= n - int(round(n * holdout_frac))
        parts.append(part)
    return {k: np.concatenate([q[k] for q in parts]) for k in keys + ["src", "held"]}


def targets(ms: np.ndarray) -> np.ndarray:
    best = ms.min(axis=1, keepdims=True)
    return np.log2(ms / best).astype(np.float32)


def regret(ms: np.ndarray, pick: np.ndarray) -> dict:
    """Time of the picked configurations against the exhaustive best."""
    best = ms.min(axis=1)
    got = ms[np.arange(len(ms)), pick]
    ratio = got / best
    return {"total_time_regret": float(got.sum() / best.sum() - 1.0),
            "geomean_regret": float(np.exp(np.mean(np.log(ratio))) - 1.0),
            "p95_ratio": float(np.quantile(ratio, 0.95)),
            "exact_pick_rate": float(np.mean(ratio == 1.0))}


def heuristic_picks(L, db) -> np.ndarray:
    nsched = L.bli_amd_gemm_config_count(1) // 14
    return np.array([L.bli_amd_gemm_default_tile(int(d), int(m), int(n), int(k)) * nsched
                     for d, (m, n, k) in zip(db["dbl"], db["dims"])], dtype=np.int64)


def guarded(db, pick: np.ndarray, heur: np.ndarray) -> np.ndarray:
    """The library's learned-mode choice (select.hip, selector_pick): the
    network's pick, except for FP64 products of at least 1024 output tiles
    of 128 x 128 whose operands all allow 16-byte access, which keep the
    size heuristic's large tile (the band the heuristic was tuned on)."""
    m, n = db["dims"][:, 0].astype(np.int64), db["dims"][:, 1].astype(np.int64)
    t128 = ((m + 127) // 128) * ((n + 127) // 128)
    f = db["feats"]
    band = ((db["dbl"] == 1) & (t128 >= 1024) & (f[:, 6] == 1) & (f[:, 7] == 1) &
            (f[:, 8] == 1))
    return np.where(band, heur, pick)


def train(db_paths, out: Path, holdout_frac: float, epochs: int, batch: int, lr: float,
          seed: int, seeds: int = 1):
    """Train `seeds` networks (seeds seed, seed+1, ...) on the training
    shapes minus a validation tenth of each database, keep the one with the
    lowest time regret on the validation shapes, and report it on the
    held-out shapes."""
    L = _setup(lib())
    db = load_db(db_paths, holdout_frac)
    n = len(db["ms"])
    tr = np.where(~db["held"])[0]
    ho = np.where(db["held"])[0]
    # validation: the last tenth of each database's training shapes
    val_mask = np.zeros(n, dtype=bool)
    for i in range(len(db_paths)):
        t = tr[db["src"][tr] == i]
        val_mask[t[len(t) - len(t) // 10:]] = True
    fit = tr[~val_mask[tr]]
    val = tr[val_mask[tr]]
    F = np.ascontiguousarray(db["feats"], dtype=np.float32)
    Y = np.ascontiguousarray(targets(db["ms"]))
    Ffit = np.ascontiguousarray(F[fit])
    Yfit = np.ascontiguousarray(Y[fit])
    mu = Ffit.mean(axis=0)
    sd = Ffit.std(axis=0)
    sd[sd < 1e-6] = 1.0
    musd = np.concatenate([mu, sd]).astype(np.float32)
    heur = heuristic_picks(L, db)
    P = np.zeros(L.bli_amd_mlp_param_count(), dtype=np.float32)
    best = None
    tried = []
    t0 = time.time()
    for sd_i in range(seed, seed + max(1, seeds)):
        net = L.bli_amd_mlp_create(sd_i)
        L.bli_amd_mlp_get_params(net, _fp(P), None)
        L.bli_amd_mlp_set_params(net, _fp(P), _fp(musd))
        loss = L.bli_amd_mlp_train(net, _fp(Ffit), _fp(Yfit), len(fit), epochs, batch, lr, sd_i)
        pick = np.zeros(n, dtype=np.int32)
        L.bli_amd_mlp_predict(net, _fp(F), n, None, pick.ctypes.data_as(C.POINTER(C.c_int)))
        pick = guarded(db, pick, heur)
        r = regret(db["ms"][val], pick[val])["total_time_regret"]
        tried.append({"seed": sd_i, "train_loss": loss, "validation_regret": r})
        if best is None or r < best[0]:
            if best is not None:
                L.bli_amd_mlp_destroy(best[1])
            best = (r, net, pick, sd_i, loss)
        else:
            L.bli_amd_mlp_destroy(net)
    t_train = time.time() - t0
    _, net, pick, chosen, loss = best
    res = {"shapes": n, "fit": len(fit), "validation": len(val), "holdout": len(ho),
           "holdout_frac": holdout_frac, "dbs": [str(Path(p).name) for p in db_paths],
           "epochs": epochs, "batch": batch, "lr": lr, "seeds": tried, "chosen_seed": chosen,
           "final_train_loss": loss, "train_seconds": round(t_train, 2),
           "learned_holdout": regret(db["ms"][ho], pick[ho]),
           "heuristic_holdout": regret(db["ms"][ho], heur[ho]),
           "learned_train": regret(db["ms"][fit], pick[fit])}
    for i, path in enumerate(db_paths):
        h = ho[db["src"][ho] == i]
        res[f"holdout_{Path(path).stem}"] = {"shapes": len(h),
                                             "learned": regret(db["ms"][h], pick[h]),
                                             "heuristic": regret(db["ms"][h], heur[h])}
    if L.bli_amd_mlp_save(net, str(out).encode()) != 0:
        raise RuntimeError(f"cannot write {out}")
    L.bli_amd_mlp_destroy(net)
    print(json.dumps(res, indent=1), flush=True)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m blis_amd.tune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--out", type=Path, required=True)
    s.add_argument("--shapes", type=int, default=22000)
    s.add_argument("--start", type=int, default=0)
    s.add_argument("--count", type=int, default=22000)
    s.add_argument("--seed", type=int, default=2024)
    s.add_argument("--max-seconds", type=float, default=1e9)
    s.add_argument("--kind", choices=("random", "round"), default="random")
    t = sub.add_parser("train")
    t.add_argument("--db", type=Path, nargs="+", required=True)
    t.add_argument("--out", type=Path, default=DATA / "selector_gfx950.bin")
    t.add_argument("--holdout-frac", type=float, default=0.1,
                   help="fraction of each database's shapes (its last ones) held out")
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--seeds", type=int, default=1,
                   help="networks trained (seeds seed..seed+N-1); the best on validation is kept")
    a = ap.parse_args(argv)
    if a.cmd == "sweep":
        sweep(a.out, a.shapes, a.start, a.count, a.seed, a.max_seconds, a.kind)
    else:
        train(a.db, a.out, a.holdout_frac, a.epochs, a.batch, a.lr, a.seed, a.seeds)


if __name__ == "__main__":
    sys.exit(main())
