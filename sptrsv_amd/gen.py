"""Synthetic upper-triangular factors for the solver's workloads.

Every generator returns an :class:`UpperCSR`: CSR of an upper-triangular U with columns
ascending in each row and the diagonal FIRST -- the layout the reference's CPU solvers read
(src/core.cpp:18-32, src/utils.cpp:9-82).  Right-hand sides follow b = U @ x_true.

Workloads (sizes are parameters so tests can use small instances of the same families):
  * laplace2d_cholesky(N=100)  exact banded Cholesky factor of the 5-point Laplacian on an
    N x N grid, natural ordering, full band of width N (n = N^2, 1,004,950 nnz at N=100).
  * laplace3d_ic0(N=128)       IC(0) factor of the 7-point Laplacian on an N^3 grid with the
    pattern of triu(A) (n = 2,097,152, nnz = 8,339,456 at N=128).
  * random_zipf(n=1e6)         per-row off-diagonal count ~ Poisson(8), column offsets
    j - i ~ Zipf(1.5) capped at n-1-i, duplicates dropped, values U(-1,1), diagonal
    1 + sum |offdiag| of its row.
  * random_supernodal(S=4096)  random elimination tree of S supernodes, widths U[16,128],
    dense upper-triangular diagonal blocks and dense blocks to 1-6 random ancestors.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class UpperCSR:
    n: int
    rowptr: np.ndarray  # int64[n+1]
    col: np.ndarray     # int32[nnz]
    val: np.ndarray     # float64[nnz]
    snodes: np.ndarray | None = None  # int32[sn+1] supernode boundaries, when known
    name: str = ""

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """U @ x for x of shape (n,) or (n, m) (row-major RHS)."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.rowptr))
        if x.ndim == 1:
            return np.bincount(rows, weights=self.val * x[self.col], minlength=self.n)
        out = np.zeros((self.n, x.shape[1]), dtype=np.float64)
        np.add.at(out, rows, self.val[:, None] * x[self.col])
        return out

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rowptr)


def _from_rows(n, row_ids, cols, vals, name, snodes=None) -> UpperCSR:
    """Assemble CSR from (row, col, val) triples already sorted by (row, col)."""
    counts = np.bincount(row_ids, minlength=n)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return UpperCSR(n, rowptr, np.ascontiguousarray(cols, dtype=np.int32),
                    np.ascontiguousarray(vals, dtype=np.float64), snodes, name)


def laplace2d_cholesky(N: int = 100) -> UpperCSR:
    """Exact Cholesky factor U (A = U^T U) of the 5-point Laplacian on an N x N grid.

    Banded right-looking factorisation: row i of U needs the N rows above it restricted to
    columns i..i+N.  All N+1 band positions of every row are stored (the factor fills the
    band), truncated at the matrix edge.
    """
    n = N * N
    bw = N
    # band storage W[k, d] = U[k, k+d], padded to width 2*bw+1 so shifted reads stay in range
    W = np.zeros((n, 2 * bw + 1), dtype=np.float64)
    arow = np.zeros(bw + 1, dtype=np.float64)
    offs = np.arange(bw + 1)
    for i in range(n):
        arow[:] = 0.0
        arow[0] = 4.0
        if (i + 1) % N != 0 and i + 1 < n:
            arow[1] = -1.0
        if i + N < n:
            arow[N] = -1.0
        lo = max(0, i - bw)
        if lo < i:
            ks = np.arange(lo, i)
            blk = W[ks[:, None], (i + offs)[None, :] - ks[:, None]]
            t = arow - blk[:, 0] @ blk
        else:
            t = arow.copy()
        d = np.sqrt(t[0])
        W[i, 0] = d
        W[i, 1:bw + 1] = t[1:] / d
    lens = np.minimum(bw + 1, n - np.arange(n))
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=rowptr[1:])
    rows = np.repeat(np.arange(n), lens)
    d = np.arange(rowptr[-1]) - np.repeat(rowptr[:-1], lens)
    col = (rows + d).astype(np.int32)
    val = W[rows, d]
    return UpperCSR(n, rowptr, col, val, None, f"laplace2d_cholesky_{N}")


def laplace3d_ic0(N: int = 128) -> UpperCSR:
    """IC(0) factor of the 7-point Laplacian (diag 6, off-diag -1) on an N^3 grid.

    Natural ordering i = x + N*y + N^2*z.  With the pattern of triu(A) no two upper
    neighbours of a row share a lower neighbour, so IC(0) reduces to
        U_ii^2 = 6 - sum_{k lower neighbour of i} 1 / U_kk^2,   U_ij = -1 / U_ii.
    The recurrence is evaluated one hyperplane x+y+z = h at a time (vectorised).
    """
    n = N ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % N
    y = (idx // N) % N
    z = idx // (N * N)
    s = np.zeros(n, dtype=np.float64)
    h = x + y + z
    order = np.argsort(h, kind="stable")
    bounds = np.searchsorted(h[order], np.arange(3 * N - 1))
    for lvl in range(3 * N - 2):
        ids = order[bounds[lvl]:bounds[lvl + 1]]
        acc = np.full(len(ids), 6.0)
        m = x[ids] > 0
        acc[m] -= 1.0 / s[ids[m] - 1]
        m = y[ids] > 0
        acc[m] -= 1.0 / s[ids[m] - N]
        m = z[ids] > 0
        acc[m] -= 1.0 / s[ids[m] - N * N]
        s[ids] = acc
    diag = np.sqrt(s)
    has = np.stack([np.ones(n, bool), x < N - 1, y < N - 1, z < N - 1], axis=1)
    cols = np.stack([idx, idx + 1, idx + N, idx + N * N], axis=1)
    vals = np.stack([diag, -1.0 / diag, -1.0 / diag, -1.0 / diag], axis=1)
    lens = has.sum(axis=1)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=rowptr[1:])
    return UpperCSR(n, rowptr, cols[has].astype(np.int32), vals[has], None, f"laplace3d_ic0_{N}")


def random_zipf(n: int = 1_000_000, seed: int = 1, lam: float = 8.0, s: float = 1.5,
                diag_scale: float = 1.0) -> UpperCSR:
    """Random sparse upper-triangular U (Poisson row counts, Zipf column offsets).

    Draw order from numpy.random.default_rng(seed): counts ~ Poisson(lam) per row; one
    Zipf(s) offset per draw, capped at n-1-i (offset 0 is the diagonal and is dropped);
    duplicates inside a row are dropped; one U(-1,1) value per kept entry in (row, col)
    order.  The diagonal is diag_scale * (1 + sum |offdiag|) (diag_scale < 1 weakens the
    dominance to raise the condition number).
    """
    rng = np.random.default_rng(seed)
    counts = rng.poisson(lam, n)
    rows = np.repeat(np.arange(n, dtype=np.int64), counts)
    offs = rng.zipf(s, len(rows)).astype(np.int64)
    offs = np.minimum(offs, n - 1 - rows)
    keep = offs > 0
    rows, cols = rows[keep], rows[keep] + offs[keep]
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    vals = rng.uniform(-1.0, 1.0, len(rows))
    absum = np.bincount(rows, weights=np.abs(vals), minlength=n)
    diag = diag_scale * (1.0 + absum)
    all_rows = np.concatenate([np.arange(n, dtype=np.int64), rows])
    all_cols = np.concatenate([np.arange(n, dtype=np.int64), cols])
    all_vals = np.concatenate([diag, vals])
    order = np.lexsort((all_cols, all_rows))
    return _from_rows(n, all_rows[order], all_cols[order], all_vals[order], f"random_zipf_{n}")


def random_supernodal(S: int = 4096, seed: int = 2, wmin: int = 16, wmax: int = 128,
                      max_anc: int = 6, sigma: float = 0.1, diag_lo: float = 1.0,
                      diag_hi: float = 2.0) -> UpperCSR:
    """Random supernodal upper factor.

    Draw order from numpy.random.default_rng(seed): widths w_s ~ U{wmin..wmax}; parent of s
    ~ U{s+1..S-1} (S-1 is the root); then per supernode s = 0..S-1: k ~ U{1..max_anc},
    min(k, #ancestors) distinct ancestors chosen uniformly from the path to the root,
    the diagonal block's diagonal ~ U[diag_lo, diag_hi] and strict upper part ~ N(0, sigma),
    and the rectangular block (w_s x sum of chosen ancestor widths) ~ N(0, sigma).
    Every row of supernode s holds its triangle part followed by the same ancestor columns,
    the structure supernodal solvers assume (src/core.cpp:53-78).
    """
    rng = np.random.default_rng(seed)
    w = rng.integers(wmin, wmax + 1, S)
    start = np.zeros(S + 1, dtype=np.int64)
    np.cumsum(w, out=start[1:])
    n = int(start[-1])
    parent = np.full(S, -1, dtype=np.int64)
    for sidx in range(S - 1):
        parent[sidx] = rng.integers(sidx + 1, S)
    row_cols, row_vals, row_len = [], [], np.zeros(n, dtype=np.int64)
    for sidx in range(S):
        path = []
        p = parent[sidx]
        while p >= 0:
            path.append(p)
            p = parent[p]
        k = int(rng.integers(1, max_anc + 1))
        chosen = np.sort(rng.choice(np.array(path, dtype=np.int64), size=min(k, len(path)),
                                    replace=False)) if path else np.zeros(0, dtype=np.int64)
        rect = (np.concatenate([np.arange(start[a], start[a + 1]) for a in chosen])
                if len(chosen) else np.zeros(0, dtype=np.int64))
        ws = int(w[sidx])
        s0 = int(start[sidx])
        dvals = rng.uniform(diag_lo, diag_hi, ws)
        tri = rng.normal(0.0, sigma, (ws, ws))
        rv = rng.normal(0.0, sigma, (ws, len(rect)))
        tri[np.arange(ws), np.arange(ws)] = dvals
        iu = np.triu_indices(ws)
        # per row d: triangle cols s0+d..s0+ws-1, then rect cols
        tcols = np.arange(s0, s0 + ws)
        for d in range(ws):
            row_cols.append(tcols[d:])
            row_cols.append(rect)
            row_vals.append(tri[d, d:])
            row_vals.append(rv[d])
            row_len[s0 + d] = (ws - d) + len(rect)
        del iu
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(row_len, out=rowptr[1:])
    col = np.concatenate(row_cols).astype(np.int32)
    val = np.concatenate(row_vals)
    return UpperCSR(n, rowptr, col, val, start.astype(np.int32), f"random_supernodal_{S}")


def scale_diagonal(U: UpperCSR, s: float) -> UpperCSR:
    """A copy of U with every diagonal entry multiplied by s (the diagonal is first in each row):
    s < 1 weakens the diagonal dominance and raises the condition number, s > 1 lowers it."""
    val = U.val.copy()
    val[U.rowptr[:-1]] *= s
    return UpperCSR(U.n, U.rowptr, U.col, val, U.snodes, f"{U.name}_d{s:.6g}")


# Diagonal scales that put kappa_1(U) at 1e2 ... 1e8 (one per decade) for the two families of the
# mixed-precision sweep, the random Zipf factor (n = 1e6, seed 1) and the random supernodal factor
# (4,096 supernodes, seed 2).  Found by tests/golden/calibrate_kappa.py (Hager/Higham estimate
# on the CPU); the GPU run re-estimates kappa for every scaled factor and reports that value.
KAPPA_SCALES: dict = {
    "zipf_1e6": {1e2: 346.196, 1e3: 99.8524, 1e4: 32.0, 1e5: 9.92624, 1e6: 3.3663, 1e7: 1.29564,
                 1e8: 0.608024},
    # estimated kappa_1: 98.8, 996, 9.32e3, 9.96e4, 9.88e5, 9.6e6, 9.89e7
    "supernodal_4096": {1e2: 1135.85, 1e3: 322.128, 1e4: 80.017, 1e5: 26.5364, 1e6: 9.01103, 1e7: 3.26189,
                        1e8: 1.45479},
    # estimated kappa_1: 99.2, 995, 9.98e3, 1.09e5, 9.97e5, 9.12e6, 9.46e7
}


def kappa_scaled(family: str, target: float) -> UpperCSR:
    """The family's factor with the diagonal scale calibrated for condition number `target`."""
    base = random_zipf(1_000_000, seed=1) if family == "zipf_1e6" else random_supernodal(4096, seed=2)
    return scale_diagonal(base, KAPPA_SCALES[family][float(target)])


def banded(n: int, bw: int, seed: int = 0) -> UpperCSR:
    """Random banded upper factor: every row holds all columns i..min(i+bw, n-1); off-diagonal
    values U(-1,1), diagonal 1 + sum |offdiag| of its row (diagonally dominant)."""
    rng = np.random.default_rng(seed)
    lens = np.minimum(bw + 1, n - np.arange(n))
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=rowptr[1:])
    rows = np.repeat(np.arange(n), lens)
    d = np.arange(rowptr[-1]) - np.repeat(rowptr[:-1], lens)
    col = (rows + d).astype(np.int32)
    val = rng.uniform(-1.0, 1.0, len(col))
    off = d > 0
    absum = np.bincount(rows[off], weights=np.abs(val[off]), minlength=n)
    val[~off] = 1.0 + absum
    return UpperCSR(n, rowptr, col, val, None, f"banded_{n}_{bw}")


BATCH_KINDS = ("banded", "zipf", "supernodal")


def batch_systems(count: int = 4096, seed: int = 3, nmin: int = 256, nmax: int = 65536):
    """Independent systems for the batched solver, as one block-diagonal UpperCSR.

    Draw order from numpy.random.default_rng(seed): per system, n ~ log-uniform[nmin, nmax]
    (rounded down), kind uniform over banded / random Zipf / random supernodal, then a
    per-system seed; banded bandwidth ~ U{4..64}; Zipf rows as random_zipf; supernodal with
    widths U[8,32] and 1-3 ancestors (sized so the system has about n rows).
    Returns (U, sys_ptr int32[count+1], kinds list).
    """
    rng = np.random.default_rng(seed)
    parts, kinds, sys_ptr = [], [], [0]
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        kind = BATCH_KINDS[int(rng.integers(0, 3))]
        s = int(rng.integers(0, 2**31 - 1))
        if kind == "banded":
            U = banded(n, int(rng.integers(4, 65)), seed=s)
        elif kind == "zipf":
            U = random_zipf(n, seed=s)
        else:
            U = random_supernodal(max(1, n // 20), seed=s, wmin=8, wmax=32, max_anc=3)
        parts.append(U)
        kinds.append(kind)
        sys_ptr.append(sys_ptr[-1] + U.n)
    sys_ptr = np.array(sys_ptr, dtype=np.int64)
    nnz_ptr = np.concatenate([[0], np.cumsum([p.nnz for p in parts])])
    rowptr = np.concatenate([[0]] + [p.rowptr[1:] + nnz_ptr[i] for i, p in enumerate(parts)]).astype(np.int64)
    col = np.concatenate([p.col.astype(np.int64) + sys_ptr[i] for i, p in enumerate(parts)]).astype(np.int32)
    val = np.concatenate([p.val for p in parts])
    total = UpperCSR(int(sys_ptr[-1]), rowptr, col, val, None, f"batch_{count}_{seed}")
    return total, sys_ptr.astype(np.int32), kinds


def batch_systems_native(count: int = 4096, seed: int = 3, nmin: int = 256, nmax: int = 65536, first: int = 0):
    """The batched-solve workload at full size, drawn by the library's native generator
    (include/sptrsv_gen.h): banded (bandwidth U{4..64}), random Zipf (Poisson(8) rows, Zipf(1.5)
    offsets) and random supernodal systems shaped like random_supernodal's defaults (widths
    U[16,128], 1-6 ancestors), n ~ log-uniform[nmin, nmax].  Returns (U, sys_ptr int32[count+1],
    kinds list) like batch_systems; U.rowptr is int64 (the batch may pass 2^31 nonzeros).
    first: index of the first system, so that chunks reproduce pieces of one larger draw."""
    from . import capi
    L = capi.lib()
    ns = np.zeros(count, dtype=np.int32)
    nz = np.zeros(count, dtype=np.int64)
    kd = np.zeros(count, dtype=np.int32)
    capi.check(L.sptrsv_gen_batch_sizes(count, first, seed, nmin, nmax, capi._ptr(ns), capi._ptr(nz), capi._ptr(kd)),
               "gen_batch_sizes")
    n = int(ns.sum())
    nnz = int(nz.sum())
    rowptr = np.empty(n + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    capi.check(L.sptrsv_gen_batch_fill(count, first, seed, nmin, nmax, capi._ptr(ns), capi._ptr(nz), capi._ptr(rowptr),
                                       capi._ptr(col), capi._ptr(val)), "gen_batch_fill")
    sys_ptr = np.zeros(count + 1, dtype=np.int32)
    np.cumsum(ns, out=sys_ptr[1:])
    kinds = [BATCH_KINDS[k] for k in kd]
    return UpperCSR(n, rowptr, col, val, None, f"batch_native_{first}_{count}_{seed}"), sys_ptr, kinds


def rhs_for(U: UpperCSR, m: int = 1, seed: int = 0, layout: str = "col"):
    """x_true ~ U(-1,1) and b = U x_true.

    layout "col": arrays of shape (m, n) -- column-major n x m (x[i + rh*n]) like the
    reference's CPU solvers; layout "row": shape (n, m) -- row-major (x[i*m + k]) like its
    GPU SpTRSM (src/kernel_full.cl:163).  For m == 1 both are returned as shape (n,).
    """
    rng = np.random.default_rng(seed)
    xt = rng.uniform(-1.0, 1.0, (U.n, m))
    b = U.matvec(xt if m > 1 else xt[:, 0])
    if m == 1:
        return xt[:, 0].copy(), np.ascontiguousarray(b)
    if layout == "row":
        return np.ascontiguousarray(xt), np.ascontiguousarray(b)
    return np.ascontiguousarray(xt.T), np.ascontiguousarray(b.T)


def to_csc(U: UpperCSR):
    """CSC of U (rows ascending per column, diagonal last): (colptr int32, rowidx int32, val)."""
    rows = np.repeat(np.arange(U.n, dtype=np.int64), np.diff(U.rowptr))
    order = np.lexsort((rows, U.col))
    counts = np.bincount(U.col, minlength=U.n)
    colptr = np.zeros(U.n + 1, dtype=np.int64)
    np.cumsum(counts, out=colptr[1:])
    return colptr, rows[order].astype(np.int32), U.val[order]
