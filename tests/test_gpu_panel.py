"""PANEL variant: one workgroup per (brick of a natural-order grid factor, panel of 16 columns of
X), bricks in dependency order, in-brick dependencies through LDS, the upstream bricks' face
cells read after their epoch flags.  FP64 x bitwise equal to the serial oracle column by column
(the kernel keeps the reference's summation order), for full and partial bricks, 2D and 3D
grids, entries missing inside the grid, m not a multiple of 16, repeated solves (epochs) with
changing m, and x at an address that is only 8-byte aligned."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from sptrsv_amd import capi, gen
from tests.test_gpu_block import grid2d, grid3d, offsets_factor

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def oracle_rows(U, b):
    return np.ascontiguousarray(O.base_gauss_upper(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))[0].T)


def holes3d(nx, ny, nz, seed):
    """grid factor with about a tenth of its off-diagonal entries missing inside the grid"""
    rng = np.random.default_rng(seed + 100)
    drop = {o: rng.random(nx * ny * nz) < 0.1 for o in (1, nx, nx * ny)}

    def keep(r, o):
        i, j = r % nx, (r // nx) % ny
        edge = ((o == 1) & (i == nx - 1)) | ((o == nx) & (j == ny - 1))
        return ~edge & ~drop[o][r]
    return offsets_factor(nx * ny * nz, (1, nx, nx * ny), seed, keep)


CASES = {
    "ic0_16": lambda: gen.laplace3d_ic0(16),          # 8 full bricks
    "ic0_10": lambda: gen.laplace3d_ic0(10),          # partial bricks at the far faces
    "grid3d_9x6x5": lambda: grid3d(9, 6, 5),
    "holes3d_12x11x10": lambda: holes3d(12, 11, 10, 5),
    "grid2d_20x13": lambda: grid2d(20, 13),           # 16x16 bricks
    "grid2d_40x40": lambda: grid2d(40, 40, seed=6),
}


@pytest.mark.parametrize("m", [1, 5, 16, 20, 48, 80])  # up to 4 panels: face stores flagged first
@pytest.mark.parametrize("name", list(CASES))
def test_panel_parity(name, m):
    U = CASES[name]()
    banner(f"panel m={m} {name}")
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, m, seed=3, layout="row")
        x = P.solve_host(b, capi.params(variant=capi.VAR_PANEL))
        ref = oracle_rows(U, b)
        assert np.array_equal(x, ref), (name, m, O.rel_error(ref, x))


def test_panel_epochs_and_changing_m():
    """solves with m = 16, 40, 16, 7, 40 on one plan: the flags of each solve hold its own
    epoch, so earlier solves' flags (at any panel count) never release a wait"""
    U = gen.laplace3d_ic0(20)
    with capi.Plan.from_upper(U) as P:
        for it, m in enumerate((16, 40, 16, 7, 40)):
            banner(f"panel epochs solve {it} m={m}")
            xt, b = gen.rhs_for(U, m, seed=10 + it, layout="row")
            x = P.solve_host(b, capi.params(variant=capi.VAR_PANEL))
            assert np.array_equal(x, oracle_rows(U, b)), (it, m)


def test_panel_unaligned_x():
    U = grid3d(9, 6, 5)
    m = 16
    banner("panel m=16 unaligned x")
    xt, b = gen.rhs_for(U, m, seed=8, layout="row")
    ref = oracle_rows(U, b)
    out = np.empty((U.n, m))
    with capi.Plan.from_upper(U) as P:
        d_b = capi.DeviceArray.from_host(np.ascontiguousarray(b))
        big = capi.DeviceArray((U.n * m + 2,))
        try:
            xp = ctypes.c_void_p(big.ptr.value + 8)
            prm = capi.params(variant=capi.VAR_PANEL)
            capi.check(capi.lib().sptrsv_solve(P.handle, ctypes.byref(prm), d_b.ptr, xp, m), "solve")
            P.sync()
            capi.check(capi.lib().sptrsv_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), xp, out.nbytes), "d2h")
        finally:
            d_b.free()
            big.free()
    assert np.array_equal(out, ref)


def test_panel_rejects():
    U = gen.random_zipf(3000, seed=11)  # not a grid factor
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 16, seed=1, layout="row")
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_PANEL))
        assert e.value.code == capi.ERR_UNSUPPORTED
    U = gen.laplace3d_ic0(12)
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 16, seed=1, layout="row")
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_PANEL, precision=capi.FP32))
        assert e.value.code == capi.ERR_UNSUPPORTED


def test_panel_watchdog_reports_and_recovers():
    """An 8x8x1600 grid: 200 bricks in one chain along z, each work item waiting on the one
    before.  A 1 us per-wait limit makes the flag waits give up (every workgroup exits
    normally); the host sees ERR_DEADLOCK, and the next solve is bitwise right."""
    U = grid3d(8, 8, 1600, seed=13)
    banner("panel watchdog grid3d_8x8x1600")
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 16, seed=9, layout="row")
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_PANEL, watchdog_us=1))
        assert e.value.code == capi.ERR_DEADLOCK
        x = P.solve_host(b, capi.params(variant=capi.VAR_PANEL))
        assert np.array_equal(x, oracle_rows(U, b))


def test_division_with_reused_reciprocal_is_ieee():
    """The kernels divide by a diagonal whose refined reciprocal they compute once per row and
    reuse over the right-hand sides; the quotient must be bitwise the compiled IEEE division for
    every input: exponents over the whole range (the guard's edges at 200 / 1800 and a
    difference of 600 included), zeros, denormals, infinities and NaNs.  Checked against the
    device's own division and, where the result is not a NaN, against numpy's."""
    rng = np.random.default_rng(5)
    N = 1 << 20
    mant = rng.uniform(1.0, 2.0, (2, N)) * rng.choice([-1.0, 1.0], (2, N))
    ex = rng.integers(-1074, 1024, (2, N))
    n, d = np.ldexp(mant[0], ex[0]), np.ldexp(mant[1], ex[1])
    # operands around the guard's exponent limits (biased 200 and 1800, difference 600)
    k = N // 4
    e_d = rng.integers(-830, 780, k)
    with np.errstate(over="ignore"):  # a few numerators overflow to infinity: kept as cases
        n[:k] = np.ldexp(mant[0, :k], e_d + rng.integers(-605, 606, k))
    d[:k] = np.ldexp(mant[1, :k], e_d)
    n[k:2 * k] = mant[0, k:2 * k] * 3.0          # the solver's ordinary range
    d[k:2 * k] = mant[1, k:2 * k] + 1.5
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.2250738585072014e-308,
                        1.7976931348623157e308, 1.0, -1.0, 3.0])
    sn, sd = np.meshgrid(special, special)
    n = np.concatenate([n, sn.ravel()])
    d = np.concatenate([d, sd.ravel()])
    fast, ieee = capi.division_check(n, d)
    assert np.array_equal(fast.view(np.uint64), ieee.view(np.uint64))
    with np.errstate(all="ignore"):
        host = n / d
    ok = ~np.isnan(host)
    assert np.array_equal(fast[ok].view(np.uint64), host[ok].view(np.uint64))
    assert np.isnan(fast[~ok]).all()
