"""BLOCK variant: blocks of rows solved inside one workgroup with the block's internal
dependencies passed through LDS.  Checks the block order the library builds (a partition into
blocks whose every dependency lies in the same or an earlier block; 8x8x8 / 32x16 bricks for
natural-order grid factors, 64-row chunks of level bands otherwise) on the CPU from the arrays
it returns, and
FP64 x bitwise equal to the serial oracle (the kernel keeps the reference's summation order)."""
import numpy as np
import pytest

from oracle import oracle as O
from sptrsv_amd import capi, gen

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def offsets_factor(n, offsets, seed, keep=None):
    """upper factor with entries (r, r + o) for o in offsets (kept where keep(r, o) is true),
    diagonal in [2, 3], off-diagonals in [-0.3, 0.3]"""
    rng = np.random.default_rng(seed)
    rows, cols = [np.arange(n)], [np.arange(n)]
    for o in offsets:
        r = np.arange(n - o)
        if keep is not None:
            r = r[keep(r, o)]
        rows.append(r)
        cols.append(r + o)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    vals = np.where(rows == cols, rng.uniform(2, 3, len(rows)), rng.uniform(-0.3, 0.3, len(rows)))
    o = np.lexsort((cols, rows))
    return gen._from_rows(n, rows[o], cols[o], vals[o], "offsets")


def grid2d(nx, ny, seed=1):
    return offsets_factor(nx * ny, (1, nx), seed, keep=lambda r, o: (o != 1) | (r % nx != nx - 1))


def grid3d(nx, ny, nz, seed=2):
    def keep(r, o):
        i, j = r % nx, (r // nx) % ny
        return ~(((o == 1) & (i == nx - 1)) | ((o == nx) & (j == ny - 1)))
    return offsets_factor(nx * ny * nz, (1, nx, nx * ny), seed, keep)


CASES = {
    # name: (factor, expected kind)
    "ic0_12": (gen.laplace3d_ic0(12), 1),
    "ic0_10": (gen.laplace3d_ic0(10), 1),        # partial bricks at the far faces
    "grid3d_9x6x5": (grid3d(9, 6, 5), 1),
    "grid2d_20x13": (grid2d(20, 13), 1),
    "grid2d_8x8": (grid2d(8, 8), 1),             # one brick
    # offsets {1, 64} with the +1 entries wrapping from line to line across two 32-wide bricks:
    # the brick order is not topological, so the level-band order is used
    "wrap_64x20": (offsets_factor(64 * 20, (1, 64), 3), 2),
    "zipf_3000": (gen.random_zipf(3000, seed=11), 2),
    "chol2d_15": (gen.laplace2d_cholesky(15), 2),
    "chain_300": (offsets_factor(300, (1,), 4), 2),
    "diag_only": (gen.UpperCSR(100, np.arange(101), np.arange(100, dtype=np.int32), np.linspace(1, 2, 100)), 2),
    # row 0 depends on all 1,999 others: more external values than a block's 512 LDS slots,
    # the rest read from x directly
    "long_row_2000": (offsets_factor(2000, (1,) + tuple(range(2, 2000)), 6,
                                     keep=lambda r, o: (o == 1) | (r == 0)), 2),
}


def check_order(U, kind, ptr, perm):
    n = U.n
    assert ptr[0] == 0 and ptr[-1] == n
    sizes = np.diff(ptr)
    assert sizes.min() >= 1 and sizes.max() <= 512  # 512 for grid factors, 64 otherwise
    assert np.array_equal(np.sort(perm), np.arange(n))
    blk = np.empty(n, dtype=np.int64)
    blk[perm] = np.repeat(np.arange(len(sizes)), sizes)
    rp = np.asarray(U.rowptr)
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = np.asarray(U.col) != rows
    assert np.all(blk[np.asarray(U.col)[off]] <= blk[rows[off]]), "a dependency lies in a later block"
    return blk


@pytest.mark.parametrize("name", list(CASES))
def test_block_order_and_parity(name):
    U, kind = CASES[name]
    banner(f"block {name} n={U.n}")
    with capi.Plan.from_upper(U) as P:
        got_kind, ptr, perm = P.blocks()
        assert got_kind == kind
        check_order(U, got_kind, ptr, perm)
        xt, b = gen.rhs_for(U, 1, seed=5)
        ref = O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0]
        for occ in (0, 1, 4, 32):
            x = P.solve_host(b, capi.params(variant=capi.VAR_BLOCK, occupancy=occ))
            assert np.array_equal(x, ref), (name, occ, O.rel_error(ref, x))


def test_bricks_of_ic0_12():
    """12^3 grid: 8 bricks of up to 8x8x8 rows (8 or 4 along each axis), each block one brick"""
    U = CASES["ic0_12"][0]
    with capi.Plan.from_upper(U) as P:
        kind, ptr, perm = P.blocks()
    assert kind == 1 and len(ptr) == 9
    N = 12
    sizes = []
    for b in range(8):
        r = perm[ptr[b]:ptr[b + 1]]
        cells = {(int(i) // 8, int(j) // 8, int(k) // 8) for i, j, k in zip(r % N, (r // N) % N, r // (N * N))}
        assert len(cells) == 1
        sizes.append(len(r))
    assert sorted(sizes) == [64, 128, 128, 128, 256, 256, 256, 512]
    # bricks by descending anti-diagonal: the first block is the far corner brick (4x4x4)
    r = perm[:64]
    assert ptr[1] == 64 and np.all(r % N >= 8) and np.all((r // N) % N >= 8) and np.all(r // (N * N) >= 8)


def test_chunks_of_a_general_factor():
    """a factor without grid structure: bands of levels cut into 64-row chunks"""
    U = CASES["zipf_3000"][0]
    with capi.Plan.from_upper(U) as P:
        kind, ptr, perm = P.blocks()
    sizes = np.diff(ptr)
    assert kind == 2 and sizes.max() <= 64 and np.median(sizes) == 64


GRID_CASES = [k for k, (U, kind) in CASES.items() if k.startswith(("ic0", "grid", "wrap"))]


@pytest.mark.parametrize("m", [2, 3, 4])
@pytest.mark.parametrize("name", GRID_CASES)
def test_block_several_rhs(name, m):
    """2..4 right-hand sides per lane on grid factors (8x8x8 / 32x16 bricks, or the band order
    of a grid whose bricks are not topological): x bitwise equal to the oracle, column by
    column"""
    U, kind = CASES[name]
    banner(f"block m={m} {name}")
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, m, seed=7, layout="row")
        x = P.solve_host(b, capi.params(variant=capi.VAR_BLOCK))
        ref = np.ascontiguousarray(O.base_gauss_upper(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))[0].T)
        assert np.array_equal(x, ref), (name, m, O.rel_error(ref, x))


@pytest.mark.parametrize("m", [2, 3, 4])
def test_block_several_rhs_unaligned_x(m):
    """x at an address 8 bytes past a 16-byte boundary (any 8-byte aligned x is valid): x
    bitwise equal to the oracle"""
    import ctypes
    U = CASES["grid3d_9x6x5"][0]
    banner(f"block m={m} unaligned x")
    xt, b = gen.rhs_for(U, m, seed=8, layout="row")
    ref = np.ascontiguousarray(O.base_gauss_upper(U.rowptr, U.col, U.val, np.ascontiguousarray(b.T))[0].T)
    out = np.empty((U.n, m))
    with capi.Plan.from_upper(U) as P:
        d_b = capi.DeviceArray.from_host(np.ascontiguousarray(b))
        big = capi.DeviceArray((U.n * m + 2,))  # device allocations are 256-byte aligned
        try:
            xp = ctypes.c_void_p(big.ptr.value + 8)
            prm = capi.params(variant=capi.VAR_BLOCK)
            capi.check(capi.lib().sptrsv_solve(P.handle, ctypes.byref(prm), d_b.ptr, xp, m), "solve")
            P.sync()
            capi.check(capi.lib().sptrsv_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), xp, out.nbytes), "d2h")
        finally:
            d_b.free()
            big.free()
    assert np.array_equal(out, ref), (m, O.rel_error(ref, out))


def test_block_rejects():
    U = CASES["ic0_12"][0]
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 5, seed=1, layout="row")
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_BLOCK))
        assert e.value.code == capi.ERR_UNSUPPORTED
        xt, b = gen.rhs_for(U, 1, seed=1)
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_BLOCK, precision=capi.FP32))
        assert e.value.code == capi.ERR_UNSUPPORTED
    U = CASES["zipf_3000"][0]  # not a grid: one right-hand side only
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 2, seed=1, layout="row")
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_BLOCK))
        assert e.value.code == capi.ERR_UNSUPPORTED


def test_block_transposed():
    """the transposed solve U^T y = b runs the same kernel on the companion plan of J U^T J
    (again a grid factor, so again bricks)"""
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve_triangular
    U = CASES["grid3d_9x6x5"][0]
    A = sp.csr_matrix((U.val, U.col, U.rowptr), shape=(U.n, U.n))
    with capi.Plan.from_upper(U) as P:
        for m in (1, 3):
            xt, b = gen.rhs_for(U, m, seed=9, layout="row")
            if m == 1:
                b = b.reshape(-1)
            got = P.solve_transposed_host(b, capi.params(variant=capi.VAR_BLOCK))
            ref = spsolve_triangular(A.T.tocsr(), b, lower=True)
            assert O.rel_error(ref, got) <= 1e-12


@pytest.mark.parametrize("name", ["chain_4000", "grid3d_8x8x1600"])
def test_block_watchdog_reports_and_recovers(name):
    """A 1 us per-wait limit forces the block kernel's own bounded waits (poller and compute
    waves) to give up, every wave exiting normally; the host sees ERR_DEADLOCK, and the next
    solve is bitwise right.  A 4000-row chain (64-row blocks, one after the other) and an
    8x8x1600 grid (200 bricks of 8x8x8 in one chain along z: the later ones wait ~1 ms)."""
    U = offsets_factor(4000, (1,), 12) if name == "chain_4000" else grid3d(8, 8, 1600, seed=13)
    banner(f"block watchdog {name}")
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 1, seed=9)
        with pytest.raises(capi.SptrsvError) as e:
            P.solve_host(b, capi.params(variant=capi.VAR_BLOCK, watchdog_us=1))
        assert e.value.code == capi.ERR_DEADLOCK
        x = P.solve_host(b, capi.params(variant=capi.VAR_BLOCK))
        assert np.array_equal(x, O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0])
