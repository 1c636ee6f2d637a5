"""Pins of the CPU oracle (oracle/oracle.c) against the reference's fixtures, exact rational
arithmetic, closed forms, library routines and brute force -- never against a retyped copy of
the oracle's own formula."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from conftest import GOLDEN
from oracle import oracle as O
from sptrsv_amd import gen

EXPECTED = json.load(open(os.path.join(GOLDEN, "expected.json")))


def exact_back_substitution(n, rp, col, val, b):
    """Rational arithmetic, dense dictionary rows: an independent restatement of U x = b."""
    x = [None] * n
    for i in range(n - 1, -1, -1):
        acc = Fraction(b[i])
        diag = None
        for p in range(rp[i], rp[i + 1]):
            j = int(col[p])
            v = Fraction(val[p]).limit_denominator(10**6)
            if j == i:
                diag = v
            else:
                acc -= v * x[j]
        x[i] = acc / diag
    return x


def longest_chain_levels(n, rp, col):
    """Levels by fixed-point relaxation over all rows until nothing changes (a different
    algorithm from the reference's single reverse sweep; same mathematical definition)."""
    lv = [1] * n
    changed = True
    while changed:
        changed = False
        for i in range(n):
            deps = [int(col[p]) for p in range(rp[i], rp[i + 1]) if int(col[p]) != i]
            want = 1 + max((lv[j] for j in deps), default=0)
            if want != lv[i]:
                lv[i] = want
                changed = True
    return lv


def random_upper(n, density, seed, dominant=True):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="csr")
    A = sp.triu(A, k=1).tocsr()
    A.data = rng.uniform(-1, 1, A.nnz)
    d = 1.0 + np.abs(A).sum(axis=1).A1 if dominant else rng.uniform(1, 2, n)
    U = (A + sp.diags(d)).tocsr()
    U.sort_indices()
    return U


@pytest.mark.parametrize("name", ["csr_5_upper", "csr_6_upper"])
def test_fixture_solution_is_exact(name, csr5, csr6):
    n, rp, col, val = csr5 if name == "csr_5_upper" else csr6
    exp = EXPECTED[name]
    b = np.ones(n)
    exact = exact_back_substitution(n, rp, col, val, [1] * n)
    assert [str(f) for f in exact] == exp["x"]
    x, _ = O.base_gauss_upper(rp, col, val, b)
    np.testing.assert_allclose(x, [float(Fraction(s)) for s in exp["x"]], rtol=0, atol=1e-15)
    xn, _ = O.base_gauss_upper(rp, col, val, b, new=True)
    assert np.array_equal(x, xn)


@pytest.mark.parametrize("name", ["csr_5_upper", "csr_6_upper"])
def test_fixture_levels_sets_indegree(name, csr5, csr6):
    n, rp, col, val = csr5 if name == "csr_5_upper" else csr6
    exp = EXPECTED[name]
    lv, L = O.levels(rp, col)
    assert lv.tolist() == exp["levels"] == longest_chain_levels(n, rp, col)
    assert L == exp["max_level"]
    ptr, rows = O.level_sets(lv, L)
    assert ptr.tolist() == exp["level_ptr"]
    assert rows.tolist() == exp["level_rows"]
    cp, ri, _ = O.transpose(rp, col, val)
    assert O.indegree_csc(n, ri).tolist() == exp["indegree"] == np.diff(rp).tolist()


def test_supernodal_fixture(csr6, snodes6):
    n, rp, col, val = csr6
    assert snodes6.tolist() == EXPECTED["csr_6_upper"]["snodes"]
    want = np.array([float(Fraction(s)) for s in EXPECTED["csr_6_upper"]["x"]])
    xs, _ = O.supernodal_upper(rp, col, val, snodes6, np.ones(n))
    xs2, _ = O.supernodal_upper(rp, col, val, snodes6, np.ones(n), new=True)
    np.testing.assert_allclose(xs, want, rtol=0, atol=1e-15)
    assert np.array_equal(xs, xs2)
    vp, cpad, rpp = O.ccs2ccs_pad(rp, col, val, snodes6)
    assert len(vp) == len(val) + int(O.lib().orc_pad_extra(len(snodes6) - 1, snodes6))
    xb, _ = O.supernodal_blas_upper(vp, cpad, rpp, snodes6, np.ones(n))
    np.testing.assert_allclose(xb, want, rtol=0, atol=1e-15)


def test_closed_form_levels_banded_and_grid():
    # banded exact Cholesky: every row depends on the next one -> level[i] = n - i
    U = gen.laplace2d_cholesky(12)
    lv, L = O.levels(U.rowptr, U.col)
    assert L == U.n and np.array_equal(lv, U.n - np.arange(U.n))
    cp, ri, _ = O.transpose(U.rowptr, U.col, U.val)
    assert np.array_equal(O.indegree_csc(U.n, ri), np.minimum(13, U.n - np.arange(U.n)))
    # 7-point grid, natural order: level = 1 + (N-1-x) + (N-1-y) + (N-1-z)
    N = 9
    V = gen.laplace3d_ic0(N)
    idx = np.arange(V.n)
    x, y, z = idx % N, (idx // N) % N, idx // (N * N)
    lv, L = O.levels(V.rowptr, V.col)
    assert L == 3 * N - 2
    assert np.array_equal(lv, 1 + (N - 1 - x) + (N - 1 - y) + (N - 1 - z))
    cp, ri, _ = O.transpose(V.rowptr, V.col, V.val)
    indeg = 1 + (x < N - 1) + (y < N - 1) + (z < N - 1)
    assert np.array_equal(O.indegree_csc(V.n, ri), indeg)
    ptr, rows = O.level_sets(lv, L)
    # level sizes = number of grid points with the given coordinate sum
    s = (N - 1 - x) + (N - 1 - y) + (N - 1 - z)
    assert np.array_equal(np.diff(ptr), np.bincount(s, minlength=L))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_levels_random_against_relaxation(seed):
    U = random_upper(60, 0.08, seed)
    lv, L = O.levels(U.indptr, U.indices)
    assert lv.tolist() == longest_chain_levels(60, U.indptr, U.indices)
    ptr, rows = O.level_sets(lv, L)
    assert ptr[0] == 0 and ptr[-1] == 60
    for l in range(1, L + 1):
        seg = rows[ptr[l - 1]:ptr[l]]
        assert np.all(np.diff(seg) > 0) and np.all(lv[seg] == l)


@pytest.mark.parametrize("seed,n,density", [(0, 200, 0.05), (1, 300, 0.02), (2, 150, 0.2)])
def test_serial_against_dense_and_scipy(seed, n, density):
    U = random_upper(n, density, seed)
    rng = np.random.default_rng(seed + 10)
    xt = rng.uniform(-1, 1, n)
    b = U @ xt
    x, _ = O.base_gauss_upper(U.indptr, U.indices, U.data, b)
    dense = sla.solve_triangular(U.toarray(), b, lower=False)
    spl = sp.linalg.spsolve_triangular(U, b, lower=False)
    assert O.rel_error(dense, x) < 1e-13
    assert O.rel_error(spl, x) < 1e-13
    assert O.rel_error(xt, x) < 1e-12


def test_serial_multi_rhs_column_major():
    U = random_upper(120, 0.05, 7)
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (4, 120))  # (m, n) == column-major n x m
    B = (U @ X.T).T.copy()
    x, _ = O.base_gauss_upper(U.indptr, U.indices, U.data, B)
    xn, _ = O.base_gauss_upper(U.indptr, U.indices, U.data, B, new=True)
    assert np.array_equal(x, xn)
    for k in range(4):
        xk, _ = O.base_gauss_upper(U.indptr, U.indices, U.data, B[k])
        assert np.array_equal(xk, x[k])
    assert O.rel_error(X, x) < 1e-12


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_barrier_is_bitwise_serial(threads):
    U = gen.random_zipf(3000, seed=5)
    xt, b = gen.rhs_for(U, 3, seed=1)
    x, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, b)
    xb, ta, ts = O.barrier_upper(U.rowptr, U.col, U.val, b, threads)
    xb2, _, _ = O.barrier_upper(U.rowptr, U.col, U.val, b, threads, new=True)
    assert np.array_equal(x, xb) and np.array_equal(x, xb2)
    assert ta >= 0 and ts >= 0


def test_tasks_by_level_partition():
    U = gen.random_zipf(2000, seed=9)
    lv, L = O.levels(U.rowptr, U.col)
    ptr, rows = O.level_sets(lv, L)
    for T in (1, 4, 7):
        tasks = O.tasks_by_level(U.rowptr, ptr, rows, L, T)
        for l in range(1, L + 1):
            tk = tasks[l]
            assert tk[0] == 0 and np.all(np.diff(tk) >= 0)
            assert tk[-1] == ptr[l] - ptr[l - 1]


def test_tasks_by_level_hand_case():
    # one level of 5 rows with lengths 1,3,1,1,2 (level nnz 8) split in 2 tasks: limit 4;
    # rows 0,1 fill task 1 (1+3), row 2 overflows -> task 2 takes the rest
    rp = np.array([0, 1, 4, 5, 6, 8], dtype=np.uint64)
    level_ptr = np.array([0, 5], dtype=np.int32)
    level_rows = np.arange(5, dtype=np.int32)
    tasks = O.tasks_by_level(rp, level_ptr, level_rows, 1, 2)
    assert tasks[1].tolist() == [0, 2, 5]


def test_lower_forward_substitution():
    U = random_upper(80, 0.1, 4)
    Lw = U.T.tocsr()
    Lw.sort_indices()  # diagonal is the last entry of each row
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, 80)
    x, _ = O.base_gauss_lower(Lw.indptr, Lw.indices, Lw.data, b)
    assert O.rel_error(sla.solve_triangular(Lw.toarray(), b, lower=True), x) < 1e-13


def test_transpose_matches_scipy():
    U = random_upper(90, 0.1, 6)
    cp, ri, cv = O.transpose(U.indptr, U.indices, U.data)
    C = U.tocsc()
    C.sort_indices()
    assert np.array_equal(cp, C.indptr) and np.array_equal(ri, C.indices) and np.array_equal(cv, C.data)
    # diagonal last in every column
    assert np.all(ri[cp[1:] - 1] == np.arange(90))


def test_supernodal_random_against_dense():
    U = gen.random_supernodal(12, seed=3, wmin=3, wmax=9)
    xt, b = gen.rhs_for(U, 2, seed=4)
    D = U.to_scipy().toarray()
    want = np.stack([sla.solve_triangular(D, b[k], lower=False) for k in range(2)])
    xs, _ = O.supernodal_upper(U.rowptr, U.col, U.val, U.snodes, b)
    xs2, _ = O.supernodal_upper(U.rowptr, U.col, U.val, U.snodes, b, new=True)
    assert np.array_equal(xs, xs2)
    assert O.rel_error(want, xs) < 1e-13
    vp, cpad, rpp = O.ccs2ccs_pad(U.rowptr, U.col, U.val, U.snodes)
    # padded format: each supernode is a dense row-major w x (w + r) block
    for k in range(len(U.snodes) - 1):
        s, e = U.snodes[k], U.snodes[k + 1]
        w = e - s
        blk = vp[rpp[s]:rpp[e]].reshape(w, -1)
        cols = cpad[rpp[s]:rpp[s + 1]].astype(np.int64)
        assert np.array_equal(blk, D[s:e][:, cols])
    xb, _ = O.supernodal_blas_upper(vp, cpad, rpp, U.snodes, b)
    assert O.rel_error(want, xb) < 1e-13


def test_rel_error_closed_form():
    assert O.rel_error(np.array([3.0, 4.0]), np.array([3.0, 4.0])) == 0.0
    assert abs(O.rel_error(np.array([3.0, 4.0]), np.array([0.0, 4.0])) - 3.0 / 5.0) < 1e-16


def test_residual_dense():
    U = random_upper(70, 0.1, 8)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, 70)
    b = rng.uniform(-1, 1, 70)
    r = O.spmv_residual(U.indptr, U.indices, U.data, x, b)
    np.testing.assert_allclose(r, b - U.toarray() @ x, rtol=0, atol=1e-13)


def test_planted_mistake_is_caught():
    """A plausible oracle slip (summing the diagonal into the off-diagonal sum) must fail the
    fixture pin: this is what the exact-fraction check above protects against."""
    n, rp, col, val = EXPECTED and (6, np.array([0, 2, 6, 9, 10, 12, 13]),
                                    np.array([0, 3, 1, 2, 4, 5, 2, 4, 5, 3, 4, 5, 5]),
                                    np.array([1, 1, 1, 2, -1, 1, 1, 2, 3, 2, 4, 1, 6.]))
    x = np.zeros(n)
    for i in range(n - 1, -1, -1):  # slip: starts at rp[i] instead of rp[i] + 1
        acc = sum(val[p] * (x[col[p]] if col[p] != i else 0.5) for p in range(rp[i], rp[i + 1]))
        x[i] = (1.0 - acc) / val[rp[i]]
    want = [float(Fraction(s)) for s in EXPECTED["csr_6_upper"]["x"]]
    assert not np.allclose(x, want)


def test_supernodal_blas_cblas_matches_loops():
    """The CBLAS-calling restatement of supernodal_blas_upper (the reference's two BLAS calls
    through a run-time-loaded CBLAS) against the plain-loop restatement and the serial solve."""
    path = O.find_cblas()
    if path is None:
        pytest.skip("no CBLAS library on this host")
    U = gen.random_supernodal(40, seed=9)
    vp, cp, rpp = O.ccs2ccs_pad(U.rowptr, U.col, U.val, U.snodes)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, (3, U.n))
    xb, t = O.supernodal_blas_upper_cblas(path, vp, cp, rpp, U.snodes, b)
    xl, _ = O.supernodal_blas_upper(vp, cp, rpp, U.snodes, b)
    xs, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, b)
    assert t >= 0
    assert O.rel_error(xl, xb) <= 1e-13 and O.rel_error(xs, xb) <= 1e-12
    with pytest.raises(OSError):
        O.supernodal_blas_upper_cblas("/nonexistent/libcblas.so", vp, cp, rpp, U.snodes, b)
