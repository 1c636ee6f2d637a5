"""GPU: batched persistent solves, graph replays, selector features, the MLP kernels and the
learned selector's gain over the fallback configuration."""
import numpy as np
import pytest

from oracle import features as F
from oracle import mlp_ref
from oracle import oracle as O
from sptrsv_amd import capi, gen, tune

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-12


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def small_batch():
    U, sp, kinds = gen.batch_systems(48, seed=3, nmin=64, nmax=3000)
    banner(f"batch n={U.n} nnz={U.nnz}")
    B = capi.Batch(U, sp)
    yield U, sp, kinds, B
    B.close()


def oracle_rowmajor(U, b):
    if b.ndim == 1:
        return O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0]
    x, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))
    return np.ascontiguousarray(x.T)


def solve_batch(B, b, m, precision=capi.FP64, graph=False):
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    B.solve(d_b, d_x, m, precision, graph)
    B.sync()
    return d_x.download()


@pytest.mark.parametrize("cfg", range(capi.BCFG_COUNT))
@pytest.mark.parametrize("m", [1, 4, 16, 32])
def test_batch_every_config(small_batch, cfg, m):
    U, sp, kinds, B = small_batch
    banner(f"batch cfg={capi.BCFG_NAMES[cfg]} m={m}")
    B.set_configs(np.full(B.nsys, cfg), np.arange(B.nsys)[::-1])
    xt, b = gen.rhs_for(U, m, seed=4, layout="row")
    x = solve_batch(B, b, m)
    ref = oracle_rowmajor(U, b)
    exact = cfg != capi.BCFG_CHAIN and (cfg == capi.BCFG_LS1 or m > 1)
    if exact:
        assert np.array_equal(x, ref), O.rel_error(ref, x)
    for s in range(B.nsys):  # every system on its own
        lo, hi = sp[s], sp[s + 1]
        assert O.rel_error(ref[lo:hi], x[lo:hi]) <= FP64_TOL, (s, kinds[s])


def test_batch_mixed_configs_order_and_fp32(small_batch):
    U, sp, kinds, B = small_batch
    rng = np.random.default_rng(0)
    cfg = rng.integers(0, capi.BCFG_COUNT, B.nsys)
    B.set_configs(cfg, rng.permutation(B.nsys))
    xt, b = gen.rhs_for(U, 1, seed=5)
    ref = oracle_rowmajor(U, b)
    assert O.rel_error(ref, solve_batch(B, b, 1)) <= FP64_TOL
    assert O.rel_error(ref, solve_batch(B, b, 1, capi.FP32)) <= 2e-5
    with pytest.raises(capi.SptrsvError):
        B.set_configs(order=np.zeros(B.nsys, dtype=np.int32))


def test_batch_graph_replays(small_batch):
    """The captured graph is replayed several times with new right-hand sides uploaded by
    ordinary (eager) copies in between; every replay is checked."""
    U, sp, kinds, B = small_batch
    B.set_configs(np.full(B.nsys, capi.BCFG_LS4), np.arange(B.nsys))
    d_b = capi.DeviceArray((U.n,))
    d_x = capi.DeviceArray((U.n,))
    for rep in range(4):
        xt, b = gen.rhs_for(U, 1, seed=100 + rep)
        d_b.upload(b)
        B.solve(d_b, d_x, 1, capi.FP64, True)
        B.sync()
        assert O.rel_error(oracle_rowmajor(U, b), d_x.download()) <= FP64_TOL, rep


@pytest.mark.parametrize("m", [1, 4, 16])
def test_batch_refined_reaches_fp64(small_batch, m):
    """FP32 batched solves + FP64 refinement: every system within 1e-12 of the oracle."""
    U, sp, kinds, B = small_batch
    B.set_configs(np.full(B.nsys, capi.BCFG_LS4), np.arange(B.nsys))
    xt, b = gen.rhs_for(U, m, seed=8, layout="row")
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    it, est = B.solve_refined(d_b, d_x, m)
    x = d_x.download()
    ref = oracle_rowmajor(U, b)
    assert 1 <= it <= 4 and est <= 1e-13, (it, est)
    for s in range(B.nsys):
        lo, hi = sp[s], sp[s + 1]
        assert O.rel_error(ref[lo:hi], x[lo:hi]) <= FP64_TOL, (s, kinds[s])


def test_batch_rejects_cross_system_columns():
    U = gen.random_zipf(100, seed=1)
    with pytest.raises(capi.SptrsvError) as e:
        capi.Batch(U, np.array([0, 50, 100], dtype=np.int32))
    assert e.value.code == capi.ERR_BAD_INPUT


def test_features_match_restatement(small_batch):
    U, sp, kinds, B = small_batch
    f = B.features(m=4)
    lv, _ = O.levels(U.rowptr, U.col)
    for s in range(B.nsys):
        ref = F.features(U.rowptr, U.col, U.val, lv, sp[s], sp[s + 1], m=4)
        np.testing.assert_allclose(f[s], ref, rtol=2e-6, atol=1e-6, err_msg=f"system {s} ({kinds[s]})")


def test_time_config_and_exhaustive(small_batch):
    U, sp, kinds, B = small_batch
    xt, b = gen.rhs_for(U, 4, seed=6, layout="row")
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    T = tune.exhaustive(B, d_b, d_x, m=4, reps=1)
    assert np.all(np.isfinite(T)) and np.all(T > 0)
    xt, b = gen.rhs_for(U, 40, seed=6, layout="row")
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    T = tune.exhaustive(B, d_b, d_x, m=40, reps=1)  # beyond the chain configuration's m
    assert np.all(np.isinf(T[:, capi.BCFG_CHAIN]))
    assert np.all(np.isfinite(T[:, :capi.BCFG_CHAIN])) and np.all(T[:, :capi.BCFG_CHAIN] > 0)


def test_guard_longest_redecides_only_the_longest(small_batch):
    """the makespan guard changes at most the systems of most levels / most nonzeros, and its
    picks are configurations valid at m"""
    U, sp, kinds, B = small_batch
    feats = B.features(4)
    base = np.zeros(B.nsys, dtype=np.int32)
    g = tune.guard_longest(U, sp, feats, base, 4, k=3)
    guarded = set(np.argsort(-feats[:, 6], kind="stable")[:3]) | set(np.argsort(-feats[:, 1], kind="stable")[:3])
    changed = set(np.nonzero(g != base)[0])
    assert changed <= guarded
    assert np.all((g >= 0) & (g < capi.BCFG_COUNT))
    xt, b = gen.rhs_for(U, 4, seed=8, layout="row")
    B.set_configs(g, tune.lpt_order(np.ones(B.nsys)))
    assert O.rel_error(oracle_rowmajor(U, b), solve_batch(B, b, 4)) <= FP64_TOL


# ---- the selector network ---------------------------------------------------------------------

def test_mlp_gradient_matches_fp64_reference():
    rng = np.random.default_rng(1)
    net = capi.MLP(24, 64, 64, 5, seed=3)
    X = rng.normal(0, 1, (300, 24)).astype(np.float32)
    y = rng.integers(0, 5, 300).astype(np.int32)
    mask = np.ones((300, 5), dtype=np.float32)
    mask[::2, 4] = 0
    y[::2] = np.where(y[::2] == 4, 0, y[::2])
    p = net.get_params()
    losses, g = net.loss_grad(X, y, mask)
    rl, rg = mlp_ref.loss_grad(p, net.sizes, X, y, mask)
    np.testing.assert_allclose(losses, rl, rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(g, rg, rtol=2e-4, atol=2e-6)


def test_mlp_gradient_finite_differences_on_device_loss():
    rng = np.random.default_rng(2)
    net = capi.MLP(24, 64, 64, 5, seed=4)
    X = rng.normal(0, 1, (64, 24)).astype(np.float32)
    y = rng.integers(0, 5, 64).astype(np.int32)
    p = net.get_params().astype(np.float64)
    _, g = net.loss_grad(X, y)
    eps = 2e-3
    idx = rng.choice(len(p), 20, replace=False)
    for k in idx:
        pp = p.copy(); pp[k] += eps
        net.set_params(pp)
        lp = net.loss_grad(X, y)[0].astype(np.float64).mean()
        pm = p.copy(); pm[k] -= eps
        net.set_params(pm)
        lm = net.loss_grad(X, y)[0].astype(np.float64).mean()
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 3e-3 + 3e-2 * abs(g[k]), (k, fd, g[k])


def test_mlp_training_fits_a_separable_task():
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (2000, 24)).astype(np.float32)
    y = (np.argmax(X[:, :5], axis=1)).astype(np.int32)
    net = capi.MLP(24, 64, 64, 5, seed=5)
    l0 = net.loss_grad(X, y)[0].mean()
    final = net.train(X, y, epochs=400, lr=1e-2)
    pred = np.argmax(net.forward(X), axis=1)
    assert final < 0.5 * l0 and (pred == y).mean() > 0.9


def test_selector_gains_over_fallback():
    """Train on autotuner labels of one batch (seed 4), apply to another (seed 3): the learned
    choice must cost less device time than the fixed fallback configuration and approach the
    exhaustive optimum."""
    Ut, spt, _ = gen.batch_systems(160, seed=4, nmin=64, nmax=3000)
    Ue, spe, _ = gen.batch_systems(96, seed=3, nmin=64, nmax=3000)
    Bt, Be = capi.Batch(Ut, spt), capi.Batch(Ue, spe)
    res = {}
    for name, (U, B) in {"train": (Ut, Bt), "eval": (Ue, Be)}.items():
        xt, b = gen.rhs_for(U, 1, seed=7)
        d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray((U.n,))
        res[name] = (B.features(1), tune.exhaustive(B, d_b, d_x, m=1, reps=2))
    sel = tune.Selector(seed=0)
    sel.fit(*res["train"], m=1, epochs=600)
    fe, Te = res["eval"]
    choice = sel.predict(fe, m=1)
    fallback = np.full(len(choice), capi.BCFG_LS1)
    r_sel, r_fb = tune.regret(Te, choice), tune.regret(Te, fallback)
    banner(f"selector regret {r_sel:.3f}  fallback regret {r_fb:.3f}")
    assert r_sel < r_fb
    assert r_sel < 0.5
    Bt.close(); Be.close(); sel.close()


def test_batch_residual_matches_host(small_batch):
    U, sp, kinds, B = small_batch
    for m in (1, 4):
        xt, b = gen.rhs_for(U, m, seed=6, layout="row")
        d_b, d_x, d_r = capi.DeviceArray.from_host(b), capi.DeviceArray.from_host(xt), capi.DeviceArray(b.shape)
        B.residual(d_b, d_x, d_r, m)
        B.sync()
        r = d_r.download()
        assert np.abs(r).max() <= 1e-13 * np.abs(b).max()  # b = U x_true exactly up to rounding
        d_x.upload(np.zeros_like(xt))
        B.residual(d_b, d_x, d_r, m)
        B.sync()
        assert np.array_equal(d_r.download(), b)


def _system(U, sp, s):
    """system s of a block-diagonal batch as its own CSR (columns local)"""
    lo, hi = int(sp[s]), int(sp[s + 1])
    e0, e1 = int(U.rowptr[lo]), int(U.rowptr[hi])
    return gen.UpperCSR(hi - lo, U.rowptr[lo:hi + 1] - e0, (U.col[e0:e1] - lo).astype(np.int32), U.val[e0:e1])


def test_c5_full_size_batch():
    """configs: 4,096 independent systems, n ~ log-uniform[256, 65536], banded / Zipf /
    supernodal (widths U[16,128], 1-6 ancestors), one persistent graph-replayed launch at
    m in {1, 16}, then FP32 solves with FP64 refinement.  Every system: normwise backward error
    of x <= 1e-12; 24 sampled systems (the 3 largest of each kind among them): x against the
    serial oracle <= 1e-12."""
    U, sp, kinds = gen.batch_systems_native(4096, seed=3, nmin=256, nmax=65536)
    banner(f"C5 batch n={U.n} nnz={U.nnz}")
    B = capi.Batch(U, sp)
    try:
        ns = np.diff(sp)
        k = np.array([gen.BATCH_KINDS.index(x) for x in kinds])
        cfg = np.where(k == 2, capi.BCFG_LS64, capi.BCFG_CHAIN).astype(np.int32)
        rng = np.random.default_rng(9)
        sample = set(rng.choice(4096, 15, replace=False).tolist())
        for kk in range(3):
            idx = np.flatnonzero(k == kk)
            sample |= set(idx[np.argsort(-ns[idx])[:3]].tolist())
        # |U| row sums per system, for the backward error
        absrow = np.empty(U.n)
        for s in range(4096):
            lo, hi = int(sp[s]), int(sp[s + 1])
            e0 = int(U.rowptr[lo])
            absrow[lo:hi] = np.add.reduceat(np.abs(U.val[e0:int(U.rowptr[hi])]), U.rowptr[lo:hi] - e0)
        for m in (1, 16):
            B.set_configs(cfg, np.argsort(-ns).astype(np.int32))
            shape = (U.n, m) if m > 1 else (U.n,)
            xt = rng.uniform(-1.0, 1.0, shape)
            d_xt, d_0 = capi.DeviceArray.from_host(xt), capi.DeviceArray.from_host(np.zeros(shape))
            d_b, d_x, d_r = capi.DeviceArray(shape), capi.DeviceArray(shape), capi.DeviceArray(shape)
            B.residual(d_0, d_xt, d_b, m)  # b = -U x_true
            banner(f"C5 solve m={m}")
            B.solve(d_b, d_x, m, capi.FP64, True)
            B.sync()
            for label in ("fp64", "refined"):
                if label == "refined":
                    info = B.solve_refined_ex(d_b, d_x, m, fallback=True)
                    banner(f"C5 refined m={m}: {info}")
                    assert info["outcome"] == "converged" or info["fell_back"], info
                B.residual(d_b, d_x, d_r, m)
                B.sync()
                x, r, b = d_x.download(), d_r.download(), d_b.download()
                x2, r2, b2 = (a.reshape(U.n, -1) for a in (x, r, b))
                for s in range(4096):
                    lo, hi = int(sp[s]), int(sp[s + 1])
                    den = np.abs(absrow[lo:hi]).max() * np.abs(x2[lo:hi]).max() + np.abs(b2[lo:hi]).max()
                    assert np.abs(r2[lo:hi]).max() <= 1e-12 * den, (label, m, s, kinds[s])
                for s in sorted(sample):
                    lo, hi = int(sp[s]), int(sp[s + 1])
                    S = _system(U, sp, s)
                    bs = b2[lo:hi]
                    ref = O.base_gauss_upper(S.rowptr, S.col, S.val, bs[:, 0] if m == 1 else np.ascontiguousarray(bs.T))[0]
                    got = x2[lo:hi, 0] if m == 1 else np.ascontiguousarray(x2[lo:hi].T)
                    assert O.rel_error(ref, got) <= FP64_TOL, (label, m, s, kinds[s], ns[s])
            for d in (d_xt, d_0, d_b, d_x, d_r):
                d.free()
    finally:
        B.close()
