"""Synthetic workload generators: library cross-checks and structural invariants."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import oracle as O
from sptrsv_amd import gen


def laplacian2d(N):
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(N, N))
    A = (sp.kron(sp.identity(N), T) + sp.kron(T, sp.identity(N))).tocsr()
    A.eliminate_zeros()
    return A


def laplacian3d(N):
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(N, N))
    I = sp.identity(N)
    A = (sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)).tocsr()
    A.eliminate_zeros()
    return A


def is_upper_diag_first(U):
    rows = np.repeat(np.arange(U.n), np.diff(U.rowptr))
    first = U.rowptr[:-1]
    ok = np.all(U.col[first] == np.arange(U.n))
    off = np.ones(U.nnz, bool)
    off[first] = False
    ok &= np.all(U.col[off] > rows[off])
    # strictly ascending within rows
    d = np.diff(U.col.astype(np.int64))
    same_row = rows[1:] == rows[:-1]
    return bool(ok and np.all(d[same_row] > 0))


@pytest.mark.parametrize("N", [10, 100])
def test_laplace2d_cholesky_matches_scipy(N):
    U = gen.laplace2d_cholesky(N)
    n = N * N
    assert U.n == n and U.nnz == sum(min(N + 1, n - i) for i in range(n))
    assert is_upper_diag_first(U)
    A = laplacian2d(N)
    ab = np.zeros((N + 1, n))
    for d in range(N + 1):
        ab[N - d, d:] = A.diagonal(d)
    cb = sla.cholesky_banded(ab, lower=False)
    rows = np.repeat(np.arange(n), np.diff(U.rowptr))
    np.testing.assert_allclose(U.val, cb[N + rows - U.col, U.col], rtol=1e-13, atol=1e-14)


def test_laplace2d_full_size_counts():
    U = gen.laplace2d_cholesky(100)
    assert (U.n, U.nnz) == (10_000, 1_004_950)


@pytest.mark.parametrize("N", [5, 12])
def test_ic0_reproduces_A_on_its_pattern(N):
    U = gen.laplace3d_ic0(N)
    assert is_upper_diag_first(U)
    A = laplacian3d(N)
    S = U.to_scipy()
    P = (S.T @ S).tocsr()
    Au = sp.triu(A).tocsr()
    Au.sort_indices()
    rows = np.repeat(np.arange(U.n), np.diff(Au.indptr))
    np.testing.assert_allclose(np.asarray(P[rows, Au.indices]).ravel(), Au.data, rtol=1e-13, atol=1e-13)
    # same pattern as triu(A)
    assert np.array_equal(S.indptr, Au.indptr) and np.array_equal(S.indices, Au.indices)


def test_ic0_full_size_counts():
    U = gen.laplace3d_ic0(128)
    assert (U.n, U.nnz) == (2_097_152, 8_339_456)
    lv, L = O.levels(U.rowptr, U.col)
    assert L == 382


def test_random_zipf_structure_and_determinism():
    U = gen.random_zipf(20_000, seed=1)
    assert is_upper_diag_first(U)
    V = gen.random_zipf(20_000, seed=1)
    assert np.array_equal(U.col, V.col) and np.array_equal(U.val, V.val)
    first = U.rowptr[:-1]
    rows = np.repeat(np.arange(U.n), np.diff(U.rowptr))
    off = np.ones(U.nnz, bool)
    off[first] = False
    absum = np.bincount(rows[off], weights=np.abs(U.val[off]), minlength=U.n)
    np.testing.assert_allclose(U.val[first], 1.0 + absum, rtol=1e-15)
    assert np.all(np.abs(U.val[off]) < 1.0)
    # mostly a chain: row i depends on i+1 for most rows (Zipf(1.5) puts ~38% mass on offset 1)
    has_next = np.zeros(U.n, bool)
    has_next[rows[off][U.col[off] == rows[off] + 1]] = True
    assert has_next.mean() > 0.9


def test_random_supernodal_structure():
    U = gen.random_supernodal(40, seed=2)
    assert is_upper_diag_first(U)
    sn = U.snodes
    assert sn[0] == 0 and sn[-1] == U.n and np.all(np.diff(sn) >= 16) and np.all(np.diff(sn) <= 128)
    for k in range(len(sn) - 1):
        s, e = sn[k], sn[k + 1]
        w = e - s
        tails = []
        for d in range(w):
            r = s + d
            cols = U.col[U.rowptr[r]:U.rowptr[r + 1]]
            assert np.array_equal(cols[:w - d], np.arange(r, e))  # dense triangle row
            tails.append(cols[w - d:])
        assert all(np.array_equal(t, tails[0]) for t in tails)  # identical rectangular pattern
        assert np.all(tails[0] >= e)


def test_rhs_layouts():
    U = gen.random_zipf(500, seed=3)
    xt, b = gen.rhs_for(U, 3, seed=0, layout="row")
    assert xt.shape == (500, 3) and b.shape == (500, 3)
    xc, bc = gen.rhs_for(U, 3, seed=0, layout="col")
    assert np.array_equal(xc, xt.T) and np.allclose(bc, b.T)
    x, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, bc)
    assert O.rel_error(xc, x) < 1e-12


def test_to_csc_matches_oracle_transpose():
    U = gen.random_zipf(800, seed=4)
    cp, ri, cv = gen.to_csc(U)
    ocp, ori, ocv = O.transpose(U.rowptr, U.col, U.val)
    assert np.array_equal(cp, ocp) and np.array_equal(ri, ori) and np.array_equal(cv, ocv)


def test_sub_batch_extracts_systems():
    """tune.sub_batch: the chosen systems of a block-diagonal batch, renumbered, same entries"""
    from sptrsv_amd import tune
    U, sp, kinds = gen.batch_systems(40, seed=3, nmin=64, nmax=800)
    idx = [0, 7, 39]
    S, p = tune.sub_batch(U, sp, idx)
    assert p[0] == 0 and p[-1] == S.n and S.rowptr[-1] == S.nnz
    rp = np.asarray(U.rowptr)
    for j, s in enumerate(idx):
        r0, r1, q0, q1 = sp[s], sp[s + 1], p[j], p[j + 1]
        assert r1 - r0 == q1 - q0
        e0, e1, f0, f1 = rp[r0], rp[r1], S.rowptr[q0], S.rowptr[q1]
        assert np.array_equal(np.asarray(U.col)[e0:e1] - r0, np.asarray(S.col)[f0:f1] - q0)
        assert np.array_equal(np.asarray(U.val)[e0:e1], np.asarray(S.val)[f0:f1])
        assert np.array_equal(np.diff(rp[r0:r1 + 1]), np.diff(np.asarray(S.rowptr)[q0:q1 + 1]))


# ---- native batch generator (include/sptrsv_gen.h) -----------------------------------------------

def _native(count=60, seed=7, nmin=256, nmax=4000):
    return gen.batch_systems_native(count, seed=seed, nmin=nmin, nmax=nmax)


def test_native_batch_is_valid_block_diagonal_and_reproducible():
    from sptrsv_amd import capi
    U, sp, kinds = _native()
    assert capi.validate_csr(U.n, U.rowptr, U.col, U.val) == capi.OK
    rows = np.repeat(np.arange(U.n), np.diff(U.rowptr))
    sys_of = np.searchsorted(sp, np.arange(U.n), side="right") - 1
    assert np.array_equal(sys_of[U.col], sys_of[rows])  # no entry leaves its system
    U2, sp2, k2 = _native()
    assert np.array_equal(U.val, U2.val) and np.array_equal(U.col, U2.col) and kinds == k2
    U3, _, _ = _native(seed=8)
    assert U3.nnz != U.nnz or not np.array_equal(U3.val, U.val)
    # chunks with `first` reproduce pieces of the whole draw
    Ua, spa, ka = _native(count=25)
    Ub, spb, kb = gen.batch_systems_native(35, seed=7, nmin=256, nmax=4000, first=25)
    assert ka + kb == kinds
    assert np.array_equal(np.concatenate([Ua.val, Ub.val]), U.val)
    assert np.array_equal(np.concatenate([Ua.col, Ub.col + Ua.n]), U.col)
    ns = np.diff(sp)
    assert set(kinds) == {"banded", "zipf", "supernodal"}
    for s, k in enumerate(kinds):
        if k != "supernodal":
            assert 256 <= ns[s] <= 4000


def test_native_banded_and_zipf_systems():
    U, sp, kinds = _native(count=90)
    rp, col, val = U.rowptr, U.col, U.val
    offs_all = []
    for s, k in enumerate(kinds):
        if k == "supernodal":
            continue
        lo, hi = sp[s], sp[s + 1]
        seg = slice(rp[lo], rp[hi])
        rows = np.repeat(np.arange(lo, hi), np.diff(rp[lo:hi + 1]))
        c, v = col[seg], val[seg]
        diag = c == rows
        absum = np.bincount(rows[~diag] - lo, weights=np.abs(v[~diag]), minlength=hi - lo)
        assert np.allclose(v[diag], 1.0 + absum, rtol=0, atol=1e-12), k  # diagonal 1 + sum |off|
        assert np.all(np.abs(v[~diag]) <= 1.0)
        if k == "banded":
            bw = int(np.diff(rp[lo:hi + 1]).max()) - 1
            assert 4 <= bw <= 64
            n = hi - lo
            assert np.array_equal(np.diff(rp[lo:hi + 1]), np.minimum(bw + 1, n - np.arange(n)))
        elif k == "zipf":
            offs_all.append(c[~diag] - rows[~diag])
    offs = np.concatenate(offs_all)
    # the same rows drawn with numpy's Poisson and Zipf samplers (random_zipf): after the cap and
    # the removal of duplicates, the shares of small offsets and the row length must agree
    R = gen.random_zipf(200_000, seed=5)
    rrows = np.repeat(np.arange(R.n), np.diff(R.rowptr))
    roffs = (R.col - rrows)[R.col != rrows]
    for k in (1, 2, 3, 5, 10):
        assert abs(np.mean(offs == k) - np.mean(roffs == k)) < 0.01, k
    band = lambda o: np.mean((o > 20) & (o <= 100))  # tails beyond ~n are capped differently
    assert abs(band(offs) - band(roffs)) < 0.01
    zn = sum(sp[s + 1] - sp[s] for s, k in enumerate(kinds) if k == "zipf")
    assert abs(len(offs) / zn - len(roffs) / R.n) < 0.1


def test_native_supernodal_systems():
    U, sp, kinds = _native(count=60, nmin=2000, nmax=6000)
    rp, col, val = U.rowptr, U.col, U.val
    seen = 0
    offd = []
    for s, k in enumerate(kinds):
        if k != "supernodal":
            continue
        seen += 1
        lo, hi = sp[s], sp[s + 1]
        # rows of one supernode: dense triangle then the same rectangular columns
        r = lo
        while r < hi:
            c0 = col[rp[r]:rp[r + 1]]
            w = 1
            while r + w < hi and np.array_equal(col[rp[r + w]:rp[r + w + 1]], c0[w:]) and c0[w] == r + w:
                w += 1
            # a supernode whose columns continue exactly into its neighbour reads as one wider
            # run here, so only the lower bound is a property of every run
            assert w >= 16 or r + w == hi
            assert np.array_equal(c0[:w], np.arange(r, r + w))
            d = val[rp[r]]
            assert 1.0 <= d < 2.0
            offd.append(val[rp[r] + 1:rp[r + 1]])
            r += w
    assert seen > 5
    o = np.concatenate(offd)
    assert abs(o.mean()) < 0.01 and abs(o.std() - 0.1) < 0.01
