"""CPU pins of the selector oracles: the FP64 network gradient against finite differences and
the feature restatement against closed forms."""
import numpy as np

from oracle import features as F
from oracle import mlp_ref
from sptrsv_amd import gen


def test_mlp_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    sizes = (5, 7, 6, 4)
    n = 5 * 7 + 7 + 7 * 6 + 6 + 6 * 4 + 4
    p = rng.normal(0, 0.5, n)
    X = rng.normal(0, 1, (9, 5))
    y = rng.integers(0, 4, 9)
    mask = np.ones((9, 4))
    mask[::3, 1] = 0
    y[::3] = np.where(y[::3] == 1, 0, y[::3])
    losses, g = mlp_ref.loss_grad(p, sizes, X, y, mask)
    eps = 1e-6
    for k in rng.choice(n, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mlp_ref.loss_grad(pp, sizes, X, y, mask)[0].mean() - mlp_ref.loss_grad(pm, sizes, X, y, mask)[0].mean()) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd)


def test_mlp_reference_softmax_closed_form():
    # zero weights: logits = biases, loss = logsumexp(b3) - b3[y]
    sizes = (3, 2, 2, 3)
    n = 3 * 2 + 2 + 2 * 2 + 2 + 2 * 3 + 3
    p = np.zeros(n)
    p[-3:] = [0.0, 1.0, 2.0]
    losses, _ = mlp_ref.loss_grad(p, sizes, np.ones((1, 3)), np.array([2]))
    assert abs(losses[0] - (np.log(1 + np.e + np.e ** 2) - 2.0)) < 1e-14


def test_features_closed_forms():
    n = 300
    rows = np.concatenate([np.arange(n), np.arange(n - 1)])
    cols = np.concatenate([np.arange(n), np.arange(1, n)])
    vals = np.concatenate([np.full(n, 2.0), np.full(n - 1, -1.0)])
    o = np.lexsort((cols, rows))
    U = gen._from_rows(n, rows[o], cols[o], vals[o], "chain")
    lev = n - np.arange(n)  # closed form of a bidiagonal chain
    f = F.features(U.rowptr, U.col, U.val, lev, 0, n, m=4)
    assert f[0] == np.log2(n) and f[5] == 1.0 and f[13] == 1.0 and f[14] == 1.0 and f[15] == 0.0
    assert abs(f[16] - (n - 1) / n) < 1e-15 and f[10] == 1.0 / n and f[22] == 2.0
    assert abs(f[23] - 1.0) < 1e-15  # mean |diag| 2 over mean |offdiag| 1
    # bidiagonal: pattern(r) = {r, r+1} is not {r} U pattern(r+1) = {r, r+1, r+2}, except for
    # the last two rows (row n-1 holds only its diagonal): n-1 supernodes, widest 2
    assert f[17] == (n - 1) / n and f[19] == 2
    # a dense upper triangle is one fundamental supernode
    m = 40
    r_, c_ = np.triu_indices(m)
    T = gen._from_rows(m, r_, c_, np.where(r_ == c_, 2.0, 0.1), "dense")
    h = F.features(T.rowptr, T.col, T.val, m - np.arange(m), 0, m)
    assert h[17] == 1.0 / m and h[19] == m and h[20] == 1.0
    # diagonal only: every row its own level-1 row
    D = gen.UpperCSR(50, np.arange(51), np.arange(50, dtype=np.int32), np.ones(50))
    g = F.features(D.rowptr, D.col, D.val, np.ones(50, dtype=np.int64), 0, 50)
    assert g[5] == 1 / 50 and g[8] == 1.0 and g[10] == 1.0
