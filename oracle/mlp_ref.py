"""FP64 numpy restatement of the selector network's forward and backward pass (checker for
sptrsv_amd/csrc/mlp.hip).  TEST INFRASTRUCTURE ONLY.  Layout: W1[in][h1], b1, W2[h1][h2], b2,
W3[h2][out], b3, flattened in that order."""
import numpy as np


def unpack(p, sizes):
    nin, h1, h2, out = sizes
    p = np.asarray(p, dtype=np.float64)
    o = 0
    shapes = [(nin, h1), (h1,), (h1, h2), (h2,), (h2, out), (out,)]
    res = []
    for s in shapes:
        k = int(np.prod(s))
        res.append(p[o:o + k].reshape(s))
        o += k
    return res


def loss_grad(p, sizes, X, y, mask=None):
    W1, b1, W2, b2, W3, b3 = unpack(p, sizes)
    X = np.asarray(X, dtype=np.float64)
    B = len(y)
    a1 = np.maximum(X @ W1 + b1, 0.0)
    a2 = np.maximum(a1 @ W2 + b2, 0.0)
    z = a2 @ W3 + b3
    mk = np.ones_like(z) if mask is None else np.asarray(mask, dtype=np.float64)
    zm = np.where(mk > 0, z, -np.inf)
    mx = zm.max(axis=1, keepdims=True)
    lse = mx[:, 0] + np.log(np.exp(zm - mx).sum(axis=1))
    losses = lse - z[np.arange(B), y]
    p_ = np.where(mk > 0, np.exp(z - lse[:, None]), 0.0)
    d3 = p_.copy()
    d3[np.arange(B), y] -= 1.0
    d3 /= B
    gW3 = a2.T @ d3
    gb3 = d3.sum(axis=0)
    d2 = (d3 @ W3.T) * (a2 > 0)
    gW2 = a1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return losses, grad
