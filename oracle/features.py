"""numpy restatement of the 24 batched-selector features (checker for the library's
system_features in sptrsv_amd/csrc/batch.hip).  TEST INFRASTRUCTURE ONLY."""
import numpy as np


def features(rowptr, col, val, levels, lo, hi, m=1):
    rp = np.asarray(rowptr, dtype=np.int64)
    c = np.asarray(col, dtype=np.int64)
    v = np.asarray(val, dtype=np.float64)
    lev = np.asarray(levels, dtype=np.int64)[lo:hi]
    n = hi - lo
    e0, e1 = rp[lo], rp[hi]
    nnz = float(e1 - e0)
    L = int(lev.max())
    width = np.bincount(lev, minlength=L + 1)[1:].astype(np.float64)
    lens = np.diff(rp[lo:hi + 1]).astype(np.float64)
    rows = np.repeat(np.arange(lo, hi), lens.astype(np.int64))
    cc = c[e0:e1]
    vv = v[e0:e1]
    first = rp[lo:hi] - e0
    off = np.ones(len(cc), bool)
    off[first] = False
    o = (cc - rows)[off].astype(np.float64)
    nd = max(len(o), 1)
    rowmax = np.zeros(n)
    if len(o):
        np.maximum.at(rowmax, rows[off] - lo, o)
    # fundamental supernodes
    widths, cur = [], 1
    for r in range(lo, hi - 1):
        li, lj = rp[r + 1] - rp[r], rp[r + 2] - rp[r + 1]
        cont = li == lj + 1 and li >= 2 and c[rp[r] + 1] == r + 1 and np.array_equal(c[rp[r] + 1:rp[r + 1]], c[rp[r + 1]:rp[r + 2]])
        if cont:
            cur += 1
        else:
            widths.append(cur)
            cur = 1
    widths.append(cur)
    widths = np.array(widths)
    rmean = nnz / n
    wmean = width.mean()
    f = np.zeros(24)
    f[0] = np.log2(n)
    f[1] = np.log2(nnz)
    f[2] = rmean
    f[3] = lens.max()
    f[4] = np.sqrt(max(0.0, (lens ** 2).mean() - rmean ** 2))
    f[5] = L / n
    f[6] = np.log2(L)
    f[7] = np.log2(wmean)
    f[8] = width.max() / n
    f[9] = np.sqrt(max(0.0, (width ** 2).mean() - wmean ** 2)) / wmean
    f[10] = (lens == 1).sum() / n
    f[11] = o.sum() / nd / n
    f[12] = (o.max() if len(o) else 0.0) / n
    f[13] = (o == 1).sum() / nd
    f[14] = (o <= 64).sum() / nd
    f[15] = (o > 4096).sum() / nd
    f[16] = (o == 1).sum() / n
    f[17] = len(widths) / n
    f[18] = n / len(widths)
    f[19] = widths.max()
    f[20] = widths[widths >= 16].sum() / n
    f[21] = rowmax.sum() / n / n
    f[22] = np.log2(m)
    dsum = np.abs(vv[first]).sum()
    osum = np.abs(vv[off]).sum()
    f[23] = np.log2((dsum / n) / max(osum / nd, 1e-300))
    return f
