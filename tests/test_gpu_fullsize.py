"""The benchmark configurations at their full sizes on the GPU, against the oracle.

C1  exact Cholesky factor of the 5-point Laplacian on a 100 x 100 grid (n = 10,000, bandwidth 100)
C2a IC(0) factor of the 7-point Laplacian on a 128^3 grid, m in {32, 128} (level pull, brick panels, AUTO)
C3  random supernodal factor, 4,096 supernodes of widths U[16,128], m in {32, 128}
Bars: analysis arrays bit-exact (closed forms where they exist), x bitwise equal to the serial
oracle for the exact-order kernels, relative error <= 1e-12 otherwise."""
import numpy as np
import pytest

from oracle import oracle as O
from sptrsv_amd import capi, gen

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def oracle_rows(U, b):
    """serial oracle for row-major b (n,) or (n, m), columns on host threads (bitwise the same)"""
    if b.ndim == 1:
        return O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0]
    x = O.base_gauss_upper_columns(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))
    return np.ascontiguousarray(x.T)


@pytest.fixture(scope="module")
def c1():
    U = gen.laplace2d_cholesky(100)
    banner(f"plan C1 chol2d_100 n={U.n} nnz={U.nnz}")
    P = capi.Plan.from_upper(U)
    yield U, P
    P.close()


def test_c1_analysis_closed_form(c1):
    U, P = c1
    n = U.n
    assert n == 10_000 and U.nnz == 1_004_950
    i = np.arange(n)
    assert P.n_levels == n
    assert np.array_equal(P.levels(), (n - i).astype(np.int32))  # row i waits on row i + 1
    assert np.array_equal(P.indegree(), np.minimum(101, n - i).astype(np.int32))
    # U^T U reproduces the Laplacian on its pattern (the factor is exact)
    A = U.to_scipy().T @ U.to_scipy()
    assert abs(A[0, 0] - 4.0) < 1e-12 and abs(A[0, 1] + 1.0) < 1e-12 and abs(A[0, 100] + 1.0) < 1e-12


@pytest.mark.parametrize("variant", [capi.VAR_AUTO, capi.VAR_CHAIN, capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL])
def test_c1_solve(c1, variant):
    U, P = c1
    banner(f"C1 solve {capi.VARIANT_NAMES.get(variant, 'auto')}")
    xt, b = gen.rhs_for(U, 1, seed=0)
    ref = oracle_rows(U, b)
    x = P.solve_host(b, capi.params(variant=variant))
    if variant in (capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL):
        assert np.array_equal(x, ref)
    else:
        assert O.rel_error(ref, x) <= FP64_TOL
    assert O.rel_error(xt, ref) < 1e-9


def test_c1_relaxed_supernodes(c1):
    """AUTO turns the band into relaxed supernodes of 128 rows: 79 dense blocks on MFMA."""
    U, P = c1
    assert P.auto_variant(1) == capi.VAR_SUPERNODAL
    sb = P.supernodes()
    assert len(sb) - 1 == (U.n + 127) // 128 and np.all(np.diff(sb)[:-1] == 128)
    for m in (1, 8, 32):
        xt, b = gen.rhs_for(U, m, seed=0, layout="row")
        x = P.solve_host(b, capi.params())
        assert O.rel_error(oracle_rows(U, b), x) <= FP64_TOL, m


def test_c1_exact_auto_is_bitwise(c1):
    U, P = c1
    xt, b = gen.rhs_for(U, 1, seed=0)
    x = P.solve_host(b, capi.params(exact=True))
    assert np.array_equal(x, oracle_rows(U, b))


def test_c1_fp32_auto(c1):
    """AUTO at FP32 on a chain-like factor runs the FP32 chain kernel."""
    U, P = c1
    assert P.auto_variant(1, capi.FP32) == capi.VAR_CHAIN
    xt, b = gen.rhs_for(U, 1, seed=0)
    x = P.solve_host(b, capi.params(precision=capi.FP32))
    assert O.rel_error(oracle_rows(U, b), x) <= 1e-4


@pytest.fixture(scope="module")
def c2a():
    U = gen.laplace3d_ic0(128)
    banner(f"plan C2a ic0_128 n={U.n} nnz={U.nnz}")
    P = capi.Plan.from_upper(U)
    yield U, P
    P.close()


@pytest.mark.parametrize("m", [32, 128])
def test_c2a_multi_rhs(c2a, m):
    U, P = c2a
    banner(f"C2a m={m}")
    xt, b = gen.rhs_for(U, m, seed=1, layout="row")
    ref = oracle_rows(U, b)
    x = P.solve_host(b, capi.params(variant=capi.VAR_LEVEL_PULL))
    assert np.array_equal(x, ref)
    x = P.solve_host(b, capi.params(variant=capi.VAR_PANEL))
    assert np.array_equal(x, ref)
    x = P.solve_host(b, capi.params())  # AUTO (brick panels)
    assert np.array_equal(x, ref)
    assert P.auto_variant(m) == capi.VAR_PANEL
    assert O.rel_error(xt, ref) < 1e-10


@pytest.fixture(scope="module")
def c3():
    U = gen.random_supernodal(4096, seed=2)
    banner(f"plan C3 supernodal_4096 n={U.n} nnz={U.nnz}")
    P = capi.Plan.from_upper(U)
    P.build_supernodes(U.snodes)
    yield U, P
    P.close()


@pytest.mark.parametrize("m", [32, 128])
def test_c3_multi_rhs(c3, m):
    U, P = c3
    banner(f"C3 m={m}")
    xt, b = gen.rhs_for(U, m, seed=2, layout="row")
    ref = oracle_rows(U, b)
    x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
    assert O.rel_error(ref, x) <= FP64_TOL
    if m == 32:
        x = P.solve_host(b, capi.params(variant=capi.VAR_LEVEL_PULL))
        assert np.array_equal(x, ref)
