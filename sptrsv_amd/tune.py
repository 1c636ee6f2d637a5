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
