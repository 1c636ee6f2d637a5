"""Autotuner and learned variant selector for single systems (AUTO of sptrsv_solve).

python -m sptrsv_amd.tune_single --train 240 --test 80 --seed 6 --out sptrsv_amd/selector_single.json

* collect(): factors drawn from several families (3D and 2D grid factors with entries missing,
  banded Cholesky factors, Zipf factors, random supernodal factors with their supernodes built,
  banded factors), sizes log-uniform; for each and each m in (1, 4, 32) the wall time of one
  solve (data on the device, warm) under every variant valid there -- the autotuner's table.
* a 24-64-64-K ReLU network (sptrsv_mlp_*, trained on the GPU) maps the plan's 24 features
  (sptrsv_plan_features) to the fastest variant, invalid variants masked;
* evaluated on held-out factors drawn from another seed against exhaustive search (the
  autotuner, the judge), the hand rule AUTO uses without a network, and the plain sync-free
  kernel (CSC push, the reference's algorithm) as the fallback; regret = chosen time / best
  time - 1, summed over the systems.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from . import capi, gen

CLASSES = [capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL, capi.VAR_CHAIN, capi.VAR_BLOCK, capi.VAR_PANEL, capi.VAR_TILE,
           capi.VAR_SUPERNODAL, capi.VAR_LEVELSET, capi.VAR_CSC_PUSH]
K = len(CLASSES)
MS = (1, 4, 32)
SLOW_MS = 150.0  # a variant whose warm-up solve takes longer is timed once


def grid_factor(dims, seed: int, holes: float = 0.0):
    """natural-order 2D/3D grid factor: entries at +1, +nx (and +nx*ny) inside the grid, a
    fraction `holes` of them missing; diagonal in [2, 3], off-diagonals in [-0.3, 0.3]"""
    rng = np.random.default_rng(seed)
    nx, ny = dims[0], dims[1]
    nz = dims[2] if len(dims) > 2 else 1
    n = nx * ny * nz
    r = np.arange(n)
    i, j, k = r % nx, (r // nx) % ny, r // (nx * ny)
    rows, cols = [r], [r]
    for off, ok in ((1, i + 1 < nx), (nx, j + 1 < ny), (nx * ny, k + 1 < nz)):
        keep = ok & (rng.random(n) >= holes)
        rows.append(r[keep])
        cols.append(r[keep] + off)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    vals = np.where(rows == cols, rng.uniform(2, 3, len(rows)), rng.uniform(-0.3, 0.3, len(rows)))
    o = np.lexsort((cols, rows))
    return gen._from_rows(n, rows[o], cols[o], vals[o], "grid")


def draw(rng: np.random.Generator):
    """one factor of a random family and size"""
    fam = rng.choice(["grid3d", "grid2d", "chol2d", "zipf", "supernodal", "banded"])
    s = int(rng.integers(0, 1 << 30))
    if fam == "grid3d":
        d = [int(v) for v in np.exp(rng.uniform(np.log(8), np.log(90), 3))]
        return fam, grid_factor(d, s, holes=float(rng.choice([0.0, 0.0, 0.1])))
    if fam == "grid2d":
        d = [int(v) for v in np.exp(rng.uniform(np.log(16), np.log(900), 2))]
        return fam, grid_factor(d, s, holes=float(rng.choice([0.0, 0.1])))
    if fam == "chol2d":
        return fam, gen.laplace2d_cholesky(int(rng.integers(8, 110)))
    if fam == "zipf":
        n = int(np.exp(rng.uniform(np.log(2000), np.log(100000))))
        return fam, gen.random_zipf(n, seed=s, lam=float(rng.choice([2.0, 4.0, 8.0, 16.0])),
                                    s=float(rng.choice([1.2, 1.5, 2.0, 3.0])))
    if fam == "supernodal":
        S = int(np.exp(rng.uniform(np.log(16), np.log(600))))
        return fam, gen.random_supernodal(S, seed=s, wmin=8, wmax=int(rng.choice([32, 64, 128])))
    n = int(np.exp(rng.uniform(np.log(2000), np.log(200000))))
    return fam, gen.banded(n, int(rng.integers(2, 65)), seed=s)


def valid(P: capi.Plan, v: int, m: int, grid: bool, has_sn: bool) -> bool:
    """the variants the library's AUTO may choose for this plan (select.hip: variant_valid)"""
    if v in (capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL, capi.VAR_LEVELSET, capi.VAR_CSC_PUSH):
        return True
    if v in (capi.VAR_CHAIN, capi.VAR_TILE):
        return m == 1
    if v == capi.VAR_BLOCK:
        return m == 1 or (grid and m <= 3)
    if v == capi.VAR_PANEL:
        return grid
    if v == capi.VAR_SUPERNODAL:
        return has_sn
    return False


def time_variant(P, d_b, d_x, m, v, reps=3):
    prm = capi.params(variant=v)
    t0 = time.perf_counter()
    P.solve(d_b, d_x, m, prm)  # builds layouts on first use: not timed
    P.sync()
    first = (time.perf_counter() - t0) * 1e3
    if first > SLOW_MS:
        return first
    best = np.inf
    for _ in range(reps):
        t0 = time.perf_counter()
        P.solve(d_b, d_x, m, prm)
        P.sync()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def collect(count: int, seed: int, log=print):
    """features (count x |MS| rows x 24), timing table (rows x K, ms, inf = invalid), the hand
    rule's choice per row, family names"""
    rng = np.random.default_rng(seed)
    F, T, H, fams = [], [], [], []
    for it in range(count):
        fam, U = draw(rng)
        P = capi.Plan.from_upper(U)
        try:
            capi.selector_clear()  # the hand rule decides below (and may build relaxed supernodes)
            if fam == "supernodal":
                P.build_supernodes(U.snodes)
            has_sn = fam == "supernodal" or P.auto_variant(1) == capi.VAR_SUPERNODAL
            grid = P.auto_variant(1) == capi.VAR_BLOCK or P.auto_variant(40) == capi.VAR_PANEL
            for m in MS:
                xt, b = gen.rhs_for(U, m, seed=it, layout="row")
                d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
                try:
                    row = np.full(K, np.inf)
                    for k, v in enumerate(CLASSES):
                        if valid(P, v, m, grid, has_sn):
                            row[k] = time_variant(P, d_b, d_x, m, v)
                    hand = P.auto_variant(m)
                finally:
                    d_b.free()
                    d_x.free()
                F.append(P.features(m))
                T.append(row)
                H.append(CLASSES.index(hand) if hand in CLASSES else -1)
                fams.append(fam)
        finally:
            P.close()
        if (it + 1) % 10 == 0:
            log(f"[tune_single] {it + 1}/{count} factors timed")
    return np.array(F, dtype=np.float32), np.array(T), np.array(H, dtype=np.int32), fams


def regret(T: np.ndarray, choice: np.ndarray) -> float:
    chosen = T[np.arange(len(choice)), choice]
    return float(chosen.sum() / T.min(axis=1).sum() - 1.0)


def forward(doc: dict, F: np.ndarray) -> np.ndarray:
    """the network's scores in numpy (the library's forward pass, for checking it)"""
    i, h1, h2, o = doc["sizes"]
    p = np.asarray(doc["params"], dtype=np.float64)
    at = 0

    def take(cnt, shape):
        nonlocal at
        out = p[at:at + cnt].reshape(shape)
        at += cnt
        return out
    W1, b1 = take(i * h1, (i, h1)), take(h1, (h1,))
    W2, b2 = take(h1 * h2, (h1, h2)), take(h2, (h2,))
    W3, b3 = take(h2 * o, (h2, o)), take(o, (o,))
    x = (np.asarray(F, dtype=np.float64) - np.asarray(doc["mu"])) / np.asarray(doc["sd"])
    return np.maximum(np.maximum(x @ W1 + b1, 0) @ W2 + b2, 0) @ W3 + b3


def predict(doc: dict, F: np.ndarray, T: np.ndarray) -> np.ndarray:
    s = forward(doc, F)
    s[~np.isfinite(T)] = -np.inf
    return np.argmax(s, axis=1)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=240)
    ap.add_argument("--test", type=int, default=80)
    ap.add_argument("--seed", type=int, default=6)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_single.json"))
    ap.add_argument("--heldout", default="", help="also write the held-out table (features, times, hand rule)")
    a = ap.parse_args(argv)
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    t0 = time.time()
    Ftr, Ttr, Htr, _ = collect(a.train, a.seed, log)
    Fte, Tte, Hte, fte = collect(a.test, a.seed + 1000, log)
    mu, sd = Ftr.mean(axis=0), Ftr.std(axis=0) + 1e-6
    mlp = capi.MLP(24, 64, 64, K, seed=0)
    try:
        mask = np.isfinite(Ttr).astype(np.float32)
        y = np.argmin(Ttr, axis=1).astype(np.int32)
        loss = mlp.train(((Ftr - mu) / sd).astype(np.float32), y, mask, epochs=a.epochs, lr=5e-3)
        params = mlp.get_params()
    finally:
        mlp.close()
    doc = {"sizes": [24, 64, 64, K], "mu": mu.tolist(), "sd": sd.tolist(), "params": params.tolist(),
           "classes": CLASSES}
    pick_tr, pick_te = predict(doc, Ftr, Ttr), predict(doc, Fte, Tte)
    push = CLASSES.index(capi.VAR_CSC_PUSH)
    meta = {"train_factors": a.train, "test_factors": a.test, "seed": a.seed, "ms": list(MS), "epochs": a.epochs,
            "final_loss": float(loss), "wall_s": round(time.time() - t0, 1),
            "classes": [capi.VARIANT_NAMES[v] for v in CLASSES],
            "train": {"accuracy": float(np.mean(pick_tr == np.argmin(Ttr, axis=1))), "regret_mlp": regret(Ttr, pick_tr),
                      "regret_hand_rule": regret(Ttr, np.maximum(Htr, 0))},
            "test": {"accuracy": float(np.mean(pick_te == np.argmin(Tte, axis=1))), "regret_mlp": regret(Tte, pick_te),
                     "regret_hand_rule": regret(Tte, np.maximum(Hte, 0)),
                     "regret_syncfree_fallback": regret(Tte, np.full(len(Tte), push)),
                     "best_total_ms": float(Tte.min(axis=1).sum()),
                     "mlp_total_ms": float(Tte[np.arange(len(Tte)), pick_te].sum()),
                     "hand_rule_total_ms": float(Tte[np.arange(len(Tte)), np.maximum(Hte, 0)].sum()),
                     "label_counts": np.bincount(np.argmin(Tte, axis=1), minlength=K).tolist()}}
    doc["meta"] = meta
    with open(a.out, "w") as f:
        json.dump(doc, f)
    if a.heldout:
        with open(a.heldout, "w") as f:
            json.dump({"features": Fte.tolist(), "times_ms": np.where(np.isfinite(Tte), Tte, -1.0).tolist(),
                       "hand_rule": Hte.tolist(), "families": fte, "classes": CLASSES}, f)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
