"""Autotuner and learned kernel selector for the batched solver.

* exhaustive(): one timed launch per configuration (every system on it), device time of every
  system -> an (nsys x K) table; configurations invalid for m are +inf.
* Selector: the 24 features of a system -> the configuration, by a 24-64-64-K ReLU network
  trained on the GPU (sptrsv_mlp_*) on the autotuner's own labels (argmin time).
* lpt_order(): longest-predicted-time-first queue order for the persistent launch.
* guard_longest(): re-decides the few systems that bound the launch's makespan (most levels,
  most nonzeros) by timing them alone under every configuration: a persistent launch cannot end
  before its longest system, so one wrong pick there costs the whole batch.
"""
from __future__ import annotations

import numpy as np

from . import capi

K = capi.BCFG_COUNT
CHAIN_MAX_M = 32  # the chain configuration serves up to this many right-hand sides


def valid_mask(nsys: int, m: int) -> np.ndarray:
    mask = np.ones((nsys, K), dtype=np.float32)
    if m > CHAIN_MAX_M:
        mask[:, capi.BCFG_CHAIN] = 0.0
    return mask


def exhaustive(batch: capi.Batch, d_b, d_x, m: int = 1, precision: int = capi.FP64, reps: int = 2) -> np.ndarray:
    """Device microseconds of every system under every configuration (min over reps)."""
    table = np.full((batch.nsys, K), np.inf)
    mask = valid_mask(batch.nsys, m)
    for c in range(K):
        if mask[0, c] == 0:
            continue
        best = None
        for _ in range(reps):
            t = batch.time_config(c, d_b, d_x, m, precision)
            best = t if best is None else np.minimum(best, t)
        table[:, c] = best
    return table


def lpt_order(cost: np.ndarray) -> np.ndarray:
    """Largest predicted cost first (longest-processing-time list scheduling)."""
    return np.argsort(-np.asarray(cost), kind="stable").astype(np.int32)


class Selector:
    """Feature -> configuration network, trained and evaluated on the GPU."""

    def __init__(self, seed: int = 0, hidden: int = 64):
        self.mlp = capi.MLP(24, hidden, hidden, K, seed=seed)
        self.mu = np.zeros(24, dtype=np.float32)
        self.sd = np.ones(24, dtype=np.float32)

    def _norm(self, feats):
        return ((np.asarray(feats, dtype=np.float32) - self.mu) / self.sd).astype(np.float32)

    def fit(self, feats: np.ndarray, times: np.ndarray, m: int = 1, epochs: int = 600, lr: float = 5e-3) -> float:
        feats = np.asarray(feats, dtype=np.float32)
        self.mu = feats.mean(axis=0)
        self.sd = feats.std(axis=0) + 1e-6
        y = np.argmin(times, axis=1).astype(np.int32)
        return self.mlp.train(self._norm(feats), y, valid_mask(len(y), m), epochs=epochs, lr=lr)

    def fit_multi(self, feats: list, times: list, ms: list, epochs: int = 600, lr: float = 5e-3) -> float:
        """One network over several RHS counts (the m feature tells them apart): features and
        autotuner tables per m, concatenated; configurations invalid at an m are masked."""
        F = np.concatenate([np.asarray(f, dtype=np.float32) for f in feats])
        y = np.concatenate([np.argmin(t, axis=1) for t in times]).astype(np.int32)
        mask = np.concatenate([valid_mask(len(t), m) for t, m in zip(times, ms)])
        self.mu = F.mean(axis=0)
        self.sd = F.std(axis=0) + 1e-6
        return self.mlp.train(self._norm(F), y, mask, epochs=epochs, lr=lr)

    def predict(self, feats: np.ndarray, m: int = 1) -> np.ndarray:
        logits = self.mlp.forward(self._norm(feats))
        logits[valid_mask(len(logits), m) == 0] = -np.inf
        return np.argmax(logits, axis=1).astype(np.int32)

    def close(self):
        self.mlp.close()

    def save(self, path: str, meta: dict | None = None):
        """Network sizes, input normalisation and parameters as JSON text."""
        import json
        doc = {"sizes": list(self.mlp.sizes), "mu": self.mu.tolist(), "sd": self.sd.tolist(),
               "params": self.mlp.get_params().tolist(), "meta": meta or {}}
        with open(path, "w") as f:
            json.dump(doc, f)

    @classmethod
    def load(cls, path: str) -> "Selector":
        import json
        with open(path) as f:
            doc = json.load(f)
        sel = cls(hidden=int(doc["sizes"][1]))
        sel.mu = np.asarray(doc["mu"], dtype=np.float32)
        sel.sd = np.asarray(doc["sd"], dtype=np.float32)
        sel.mlp.set_params(np.asarray(doc["params"], dtype=np.float32))
        sel.meta = doc.get("meta", {})
        return sel


def sub_batch(U, sys_ptr, idx):
    """systems `idx` of a block-diagonal batch as a batch of their own: (CSR, sys_ptr)"""
    from .gen import UpperCSR
    rp, col, val = np.asarray(U.rowptr, dtype=np.int64), np.asarray(U.col), np.asarray(U.val)
    sp = np.asarray(sys_ptr, dtype=np.int64)
    rps, cols, vals, ptr = [np.zeros(1, dtype=np.int64)], [], [], [0]
    for s in idx:
        r0, r1 = int(sp[s]), int(sp[s + 1])
        e0, e1 = int(rp[r0]), int(rp[r1])
        rps.append(rp[r0 + 1:r1 + 1] - e0 + rps[-1][-1])
        cols.append(col[e0:e1].astype(np.int64) - r0 + ptr[-1])
        vals.append(val[e0:e1])
        ptr.append(ptr[-1] + r1 - r0)
    n = ptr[-1]
    sub = UpperCSR(n, np.concatenate(rps), np.concatenate(cols).astype(np.int32), np.concatenate(vals), None, "sub")
    return sub, np.asarray(ptr, dtype=np.int32)


def guard_longest(U, sys_ptr, feats: np.ndarray, choice: np.ndarray, m: int, k: int = 8) -> np.ndarray:
    """The selector's choices with the k systems of most levels and the k of most nonzeros
    (features 6 and 1: log2 levels, log2 nnz) re-decided by timing them alone, in a batch of
    their own, under every configuration valid at m."""
    feats = np.asarray(feats)
    idx = np.unique(np.concatenate([np.argsort(-feats[:, 6], kind="stable")[:k],
                                    np.argsort(-feats[:, 1], kind="stable")[:k]]))
    S, sp = sub_batch(U, sys_ptr, idx)
    B = capi.Batch(S, sp)
    b = np.ones((S.n, m)) if m > 1 else np.ones(S.n)
    d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
    try:
        T = exhaustive(B, d_b, d_x, m=m, reps=2)
    finally:
        d_b.free()
        d_x.free()
        B.close()
    out = np.array(choice, dtype=np.int32, copy=True)
    out[idx] = np.argmin(T, axis=1)
    return out


def regret(times: np.ndarray, choice: np.ndarray) -> float:
    """Sum of chosen device times over the sum of the best ones, minus one."""
    chosen = times[np.arange(len(choice)), choice]
    return float(chosen.sum() / times.min(axis=1).sum() - 1.0)


def collect(total: int, chunk: int, seed: int, nmin: int, nmax: int, ms=(1, 4, 16), reps: int = 2, log=print):
    """Autotuner labels for `total` systems of the native generator (seed, drawn in chunks that
    reproduce one draw): per m, the features (total x 24) and the exhaustive timing table
    (total x K).  Right-hand sides are random: the timings do not depend on their values."""
    from . import gen
    feats = {m: [] for m in ms}
    times = {m: [] for m in ms}
    rng = np.random.default_rng(seed)
    for first in range(0, total, chunk):
        cnt = min(chunk, total - first)
        U, sp, _ = gen.batch_systems_native(cnt, seed=seed, nmin=nmin, nmax=nmax, first=first)
        B = capi.Batch(U, sp)
        n, nnz = U.n, U.nnz
        del U
        for m in ms:
            b = rng.uniform(-1.0, 1.0, (n, m) if m > 1 else n)
            d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
            del b
            try:
                times[m].append(exhaustive(B, d_b, d_x, m=m, reps=reps))
            finally:
                d_b.free()
                d_x.free()
            feats[m].append(B.features(m))
        B.close()
        log(f"[tune] systems {first}..{first + cnt - 1}: n={n} nnz={nnz} timed")
    return {m: (np.concatenate(feats[m]), np.concatenate(times[m])) for m in ms}


def train_main(argv=None):
    """python -m sptrsv_amd.tune --count 20000 --chunk 2000 --seed 4 --out sptrsv_amd/selector.json

    Collects the autotuner's labels on the native generator's systems and trains the selector
    network on the GPU; writes its parameters and the training summary as JSON."""
    import argparse
    import json
    import os
    import sys
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--chunk", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--nmin", type=int, default=256)
    ap.add_argument("--nmax", type=int, default=65536)
    ap.add_argument("--epochs", type=int, default=1500)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector.json"))
    a = ap.parse_args(argv)
    t0 = time.time()
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    ms = (1, 4, 16)
    data = collect(a.count, a.chunk, a.seed, a.nmin, a.nmax, ms, a.reps, log)
    sel = Selector(seed=0)
    loss = sel.fit_multi([data[m][0] for m in ms], [data[m][1] for m in ms], list(ms), epochs=a.epochs)
    summary = {"count": a.count, "seed": a.seed, "nmin": a.nmin, "nmax": a.nmax, "ms": list(ms),
               "epochs": a.epochs, "final_loss": loss, "wall_s": round(time.time() - t0, 1)}
    for m in ms:
        F, T = data[m]
        pick = sel.predict(F, m)
        best = np.argmin(T, axis=1)
        summary[f"m{m}"] = {"train_accuracy": float(np.mean(pick == best)), "train_regret": regret(T, pick),
                            "label_counts": np.bincount(best, minlength=K).tolist()}
    sel.save(a.out, summary)
    print(json.dumps(summary))
    sel.close()


if __name__ == "__main__":
    train_main()
