"""The no-allocation entry point sptrsv_solve_push_noalloc against sptrsv_solve with the CSC push
variant (the entry point it is a variant of) and the oracle, through the C ABI."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from sptrsv_amd import capi, gen

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _solve_both(U, m, seed):
    """(no-allocation x, sptrsv_solve CSC_PUSH x, oracle x), row-major n x m"""
    xt, b = gen.rhs_for(U, m, seed=seed, layout="row")
    with capi.Plan.from_upper(U) as P:
        d_b = capi.DeviceArray.from_host(b)
        d_x = capi.DeviceArray(b.shape)
        rc = capi.lib().sptrsv_solve_push_noalloc(P.handle, d_b.ptr, d_x.ptr, int(m))
        assert rc == capi.OK, capi.last_error()
        P.sync()
        x_na = d_x.download()
        x_push = P.solve_host(b, capi.params(variant=capi.VAR_CSC_PUSH))
    if m == 1:
        ref = O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0]
    else:
        ref = np.ascontiguousarray(O.base_gauss_upper(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))[0].T)
    return x_na, x_push, ref


def test_noalloc_matches_push_twice_then_smaller():
    """Two calls in a row, the second on a smaller input: the fixed array keeps the first call's
    contents, so the second only passes if every call clears what must start at zero."""
    cases = [(gen.random_zipf(3000, seed=11), 4, 1), (gen.laplace3d_ic0(6), 1, 2)]
    for U, m, seed in cases:
        print(f"[gpu] noalloc n={U.n} m={m}", flush=True)
        x_na, x_push, ref = _solve_both(U, m, seed)
        assert O.rel_error(x_push, x_na) <= FP64_TOL
        assert O.rel_error(ref, x_na) <= FP64_TOL


def test_noalloc_refuses_sizes_beyond_the_array():
    U = gen.random_zipf(2000, seed=3)
    m = capi.NOALLOC_MAX_ELEMS // U.n + 1  # n*m just above the limit: refused on the host
    with capi.Plan.from_upper(U) as P:
        d = capi.DeviceArray((8,))
        rc = capi.lib().sptrsv_solve_push_noalloc(P.handle, d.ptr, d.ptr, int(m))
        assert rc == capi.ERR_UNSUPPORTED
        assert "exceed the fixed work array" in capi.last_error()
        assert capi.lib().sptrsv_solve_push_noalloc(P.handle, d.ptr, d.ptr, 0) == capi.ERR_BAD_INPUT
