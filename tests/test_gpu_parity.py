"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the same inputs.

Bars: analysis arrays (in-degree, levels, level sets, CSC) bit-exact; FP64 x bitwise equal to
the serial oracle for the exact-order kernels (row/level pull, level-set), relative error
<= 1e-12 for the atomic-order push kernel and the entry-split kernel; FP32 x within
2e-5 relative error on well-conditioned factors."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle import oracle as O
from sptrsv_amd import capi, gen

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5
EXACT = [capi.VAR_ROW_PULL, capi.VAR_LEVEL_PULL, capi.VAR_LEVELSET]
EXACT_M1 = EXACT + [capi.VAR_TILE, capi.VAR_BLOCK]  # the tile and block kernels solve one right-hand side


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def small_matrices():
    rp5 = np.array([0, 2, 3, 4, 5, 6]); c5 = np.array([0, 3, 1, 2, 3, 4]); v5 = np.array([1, 2, 3, 4, 5, 6.0])
    mats = {"csr_5": gen.UpperCSR(5, rp5, c5.astype(np.int32), v5)}
    mats["zipf_3000"] = gen.random_zipf(3000, seed=11)
    mats["ic0_12"] = gen.laplace3d_ic0(12)
    mats["chol2d_15"] = gen.laplace2d_cholesky(15)
    mats["snode_10"] = gen.random_supernodal(10, seed=3)
    mats["diag_only"] = gen.UpperCSR(257, np.arange(258), np.arange(257, dtype=np.int32),
                                     np.linspace(1, 2, 257))
    n = 70  # dense upper triangle
    rows, cols = np.triu_indices(n)
    rng = np.random.default_rng(5)
    vals = rng.uniform(-0.05, 0.05, len(rows))
    vals[rows == cols] = rng.uniform(1, 2, n)
    mats["dense_70"] = gen._from_rows(n, rows, cols, vals, "dense_70")
    n = 4000  # bidiagonal chain
    rows = np.concatenate([np.arange(n), np.arange(n - 1)])
    cols = np.concatenate([np.arange(n), np.arange(1, n)])
    vals = np.concatenate([np.full(n, 2.0), np.full(n - 1, -0.9)])
    o = np.lexsort((cols, rows))
    mats["chain_4000"] = gen._from_rows(n, rows[o], cols[o], vals[o], "chain")
    # grid factors: lines of 130 rows (three 64-row segments of the level sweep, the last one
    # short) with a tenth of the entries missing; and offsets {1, 8, 64} with every entry present,
    # wrap-arounds included, where the sweep's line order is not topological (row order instead)
    mats["holes_130x5x4"] = _offsets(130 * 5 * 4, (1, 130, 650), 7, nx=130, ny=5, drop=0.1)
    mats["holes_100x30"] = _offsets(100 * 30, (1, 100), 9, nx=100, ny=30, drop=0.1)  # a 2D grid
    mats["wrap_8x8x8"] = _offsets(512, (1, 8, 64), 8)
    return mats


def _offsets(n, offsets, seed, nx=0, ny=0, drop=0.0):
    """entries (r, r + o) for o in offsets; with nx, ny none crossing a line or plane edge and
    a fraction `drop` removed; diagonal in [2, 3], off-diagonals in [-0.3, 0.3]"""
    rng = np.random.default_rng(seed)
    rows, cols = [np.arange(n)], [np.arange(n)]
    for o in offsets:
        r = np.arange(n - o)
        if nx:
            i, j = r % nx, (r // nx) % ny
            r = r[~(((o == 1) & (i == nx - 1)) | ((o == nx) & (j == ny - 1))) & (rng.random(len(r)) >= drop)]
        rows.append(r)
        cols.append(r + o)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    vals = np.where(rows == cols, rng.uniform(2, 3, len(rows)), rng.uniform(-0.3, 0.3, len(rows)))
    o = np.lexsort((cols, rows))
    return gen._from_rows(n, rows[o], cols[o], vals[o], "offsets")


MATS = small_matrices()


@pytest.fixture(scope="module")
def plans():
    out = {}
    for k, U in MATS.items():
        banner(f"plan {k} n={U.n} nnz={U.nnz}")
        out[k] = capi.Plan.from_upper(U)
    yield out
    for p in out.values():
        p.close()


def oracle_solve(U, b_rowmajor):
    """oracle serial solve; b row-major (n,) or (n, m) -> x same layout"""
    if b_rowmajor.ndim == 1:
        return O.base_gauss_upper(U.rowptr, U.col, U.val, b_rowmajor)[0]
    x, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, np.ascontiguousarray(b_rowmajor.T))
    return np.ascontiguousarray(x.T)


@pytest.mark.parametrize("name", list(MATS))
def test_analysis_bit_exact(plans, name):
    U, P = MATS[name], plans[name]
    banner(f"analysis {name}")
    lv, L = O.levels(U.rowptr, U.col)
    assert P.n_levels == L
    assert np.array_equal(P.levels(), lv.astype(np.int32))
    ptr, rows = O.level_sets(lv, L)
    gptr, grows = P.level_sets()
    assert np.array_equal(gptr, ptr) and np.array_equal(grows, rows)
    cp, ri, cv = O.transpose(U.rowptr, U.col, U.val)
    gcp, gri, gcv = P.csc()
    assert np.array_equal(gcp, cp) and np.array_equal(gri, ri) and np.array_equal(gcv, cv)
    assert np.array_equal(P.indegree(), O.indegree_csc(U.n, ri))


def test_device_transpose_general_pattern():
    rng = np.random.default_rng(1)
    import scipy.sparse as sp
    A = sp.random(300, 300, density=0.03, random_state=rng, format="csr")
    A.sort_indices()
    cp, ri, cv = capi.csr_to_csc(300, A.indptr, A.indices, A.data)
    C = A.tocsc()
    C.sort_indices()
    assert np.array_equal(cp, C.indptr) and np.array_equal(ri, C.indices) and np.array_equal(cv, C.data)


@pytest.mark.parametrize("n,per_row", [(70000, 7), (4096, 1), (200, 150), (1 << 17, 0)])
def test_device_transpose_sort_sizes(n, per_row):
    """the transpose's own stable radix sort over many tiles and digit passes (n = 70,000: three
    8-bit passes, ~120 tiles of 4,096 entries, a two-level scan of the digit histogram),
    one column per row, dense rows, and columns left empty (per_row 0: a few entries only)"""
    rng = np.random.default_rng(n + per_row)
    import scipy.sparse as sp
    k = per_row if per_row else 0
    rows = np.repeat(np.arange(n), k) if k else rng.integers(0, n, 37)
    cols = rng.integers(0, n, len(rows))
    A = sp.csr_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    cp, ri, cv = capi.csr_to_csc(n, A.indptr, A.indices, A.data)
    C = A.tocsc()
    C.sort_indices()
    assert np.array_equal(cp, C.indptr) and np.array_equal(ri, C.indices) and np.array_equal(cv, C.data)


@pytest.mark.parametrize("variant", EXACT_M1 + [capi.VAR_CSC_PUSH])
@pytest.mark.parametrize("name", list(MATS))
def test_fp64_single_rhs(plans, name, variant):
    U, P = MATS[name], plans[name]
    banner(f"solve m=1 {name} {capi.VARIANT_NAMES[variant]}")
    xt, b = gen.rhs_for(U, 1, seed=2)
    x = P.solve_host(b, capi.params(variant=variant))
    ref = oracle_solve(U, b)
    if variant in EXACT_M1:
        assert np.array_equal(x, ref), O.rel_error(ref, x)
    else:
        assert O.rel_error(ref, x) <= FP64_TOL


@pytest.mark.parametrize("name", list(MATS))
def test_fp64_chain(plans, name):
    U, P = MATS[name], plans[name]
    banner(f"solve chain {name}")
    xt, b = gen.rhs_for(U, 1, seed=13)
    x = P.solve_host(b, capi.params(variant=capi.VAR_CHAIN))
    assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL


@pytest.mark.parametrize("n", [1, 31, 32, 33, 63, 64, 65, 2048 + 130, 4096 + 130, 9000])
def test_chain_window_and_ring_edges(n):
    """window boundaries (32 rows), a partial last window, and entries on both sides of the
    2048-row LDS ring's edge"""
    rng = np.random.default_rng(n)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [rng.uniform(2, 3, n)]
    for off in (1, 2, 31, 32, 33, 63, 64, 65, 2047, 2048, 2049, 4095, 4096, 4097, 5000):
        r = np.arange(max(n - off, 0))
        rows.append(r); cols.append(r + off); vals.append(rng.uniform(-0.2, 0.2, len(r)))
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    o = np.lexsort((cols, rows))
    U = gen._from_rows(n, rows[o], cols[o], vals[o], "edges")
    with capi.Plan.from_upper(U) as P:
        lv, L = O.levels(U.rowptr, U.col)  # n >= 2048: levels from the windowed sweep
        assert P.n_levels == L and np.array_equal(P.levels(), lv.astype(np.int32))
        xt, b = gen.rhs_for(U, 1, seed=1)
        x = P.solve_host(b, capi.params(variant=capi.VAR_CHAIN))
        assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL


def test_chain_direct_path_windows():
    """windows whose entries overflow a record (more than 448 entries within the ring, or more
    than 64 older entries) are solved straight from the CSR, between windows that use records"""
    n = 7000
    rng = np.random.default_rng(5)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [rng.uniform(2, 3, n)]
    r = np.arange(n - 2049)
    for off in (1, 2049):
        rows.append(r); cols.append(r + off); vals.append(rng.uniform(-0.2, 0.2, len(r)))
    band = np.arange(2000, 3000)  # 20 near entries per row: 640 per window
    for off in range(2, 22):
        rows.append(band); cols.append(band + off); vals.append(rng.uniform(-0.04, 0.04, len(band)))
    far = np.arange(4000, 4500)  # 3 old entries per row: 96 per window
    for off in (2100, 2200, 2300):
        rows.append(far); cols.append(far + off); vals.append(rng.uniform(-0.1, 0.1, len(far)))
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    o = np.lexsort((cols, rows))
    U = gen._from_rows(n, rows[o], cols[o], vals[o], "direct")
    with capi.Plan.from_upper(U) as P:
        lv, L = O.levels(U.rowptr, U.col)  # windowed level sweep with direct windows and old columns
        assert P.n_levels == L and np.array_equal(P.levels(), lv.astype(np.int32))
        xt, b = gen.rhs_for(U, 1, seed=3)
        ref = oracle_solve(U, b)
        x = P.solve_host(b, capi.params(variant=capi.VAR_CHAIN))
        assert O.rel_error(ref, x) <= FP64_TOL
        x2 = P.solve_host(b, capi.params(variant=capi.VAR_CHAIN))  # the records are reused
        assert np.array_equal(x, x2)


@pytest.mark.parametrize("group", [2, 8, 64])
@pytest.mark.parametrize("name", ["zipf_3000", "chol2d_15", "dense_70", "chain_4000"])
def test_fp64_entry_split(plans, name, group):
    U, P = MATS[name], plans[name]
    banner(f"solve split {name} G={group}")
    xt, b = gen.rhs_for(U, 1, seed=3)
    for variant in (capi.VAR_ROW_PULL, capi.VAR_LEVEL_PULL):
        x = P.solve_host(b, capi.params(variant=variant, group=group))
        assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL


@pytest.mark.parametrize("unroll", [1, 2, 4])
def test_fp64_unroll(plans, unroll):
    for name in ("zipf_3000", "chol2d_15"):
        U, P = MATS[name], plans[name]
        xt, b = gen.rhs_for(U, 1, seed=4)
        x = P.solve_host(b, capi.params(variant=capi.VAR_ROW_PULL, unroll=unroll))
        assert np.array_equal(x, oracle_solve(U, b))


@pytest.mark.parametrize("m", [2, 3, 8, 32, 64, 100, 128, 300])
@pytest.mark.parametrize("variant", EXACT + [capi.VAR_CSC_PUSH])
def test_fp64_multi_rhs(plans, m, variant):
    for name in ("zipf_3000", "ic0_12", "snode_10"):
        U, P = MATS[name], plans[name]
        banner(f"solve m={m} {name} {capi.VARIANT_NAMES[variant]}")
        xt, b = gen.rhs_for(U, m, seed=5, layout="row")
        x = P.solve_host(b, capi.params(variant=variant))
        ref = oracle_solve(U, b)
        if variant in EXACT:
            assert np.array_equal(x, ref), O.rel_error(ref, x)
        else:
            assert O.rel_error(ref, x) <= FP64_TOL


@pytest.mark.parametrize("block", [64, 128, 1024])
@pytest.mark.parametrize("occupancy", [-1, 1, 2])
def test_pull_launch_shapes(plans, block, occupancy):
    """Workgroup size and residency cap change only the schedule: the level-ordered and
    row-ordered pull kernels stay bitwise equal to the serial oracle."""
    for name in ("zipf_3000", "ic0_12", "chain_4000"):
        U, P = MATS[name], plans[name]
        for m in (1, 8, 128):
            banner(f"shape {name} m={m} block={block} occ={occupancy}")
            xt, b = gen.rhs_for(U, m, seed=21, layout="row")
            ref = oracle_solve(U, b)
            for variant in (capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL):
                x = P.solve_host(b, capi.params(variant=variant, block=block, occupancy=occupancy))
                assert np.array_equal(x, ref), (name, m, variant, O.rel_error(ref, x))


@pytest.mark.parametrize("group", [8, 16, 32])
def test_level_pull_lanes_per_row(plans, group):
    """With several right-hand sides, `group` sets the lanes per row of the level-ordered pull
    kernel (each lane then takes up to 8 columns): only the schedule changes, x stays bitwise
    equal to the serial oracle, for widths that leave a partial last lane block too."""
    for name in ("zipf_3000", "ic0_12", "chain_4000"):
        U, P = MATS[name], plans[name]
        for m in (5, 32, 100, 128, 300):
            banner(f"lanes per row {name} m={m} group={group}")
            xt, b = gen.rhs_for(U, m, seed=23, layout="row")
            ref = oracle_solve(U, b)
            x = P.solve_host(b, capi.params(variant=capi.VAR_LEVEL_PULL, group=group))
            assert np.array_equal(x, ref), (name, m, group, O.rel_error(ref, x))


def lower_of(U):
    """L = U^T as CSR with the diagonal last in each row (the CSC arrays of U)."""
    return O.transpose(U.rowptr, U.col, U.val)


def oracle_lower(U, b_rowmajor):
    """the oracle's forward solve (src/core.cpp:4-16) with L = U^T, column by column"""
    cp, ri, cv = lower_of(U)
    if b_rowmajor.ndim == 1:
        return O.base_gauss_lower(cp, ri, cv, b_rowmajor)[0]
    return np.stack([O.base_gauss_lower(cp, ri, cv, np.ascontiguousarray(b_rowmajor[:, k]))[0]
                     for k in range(b_rowmajor.shape[1])], axis=1)


@pytest.mark.parametrize("variant", [capi.VAR_AUTO, capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL, capi.VAR_LEVELSET,
                                     capi.VAR_CSC_PUSH, capi.VAR_CHAIN, capi.VAR_TILE, capi.VAR_BLOCK])
def test_transposed_solve(plans, variant):
    """U^T x = b through the companion plan of J U^T J matches the oracle's forward solve of
    L = U^T (diagonal last) to 1e-12, for every kernel family."""
    for name in MATS:
        U, P = MATS[name], plans[name]
        for m in ((1,) if variant in (capi.VAR_CHAIN, capi.VAR_TILE, capi.VAR_BLOCK) else (1, 3)):
            banner(f"transposed {name} m={m} {capi.VARIANT_NAMES.get(variant, 'auto')}")
            xt, b = gen.rhs_for(U, m, seed=31, layout="row")
            if m == 1:
                b = b.reshape(-1)
            ref = oracle_lower(U, b)
            x = P.solve_transposed_host(b, capi.params(variant=variant))
            assert O.rel_error(ref, x) <= FP64_TOL, (name, m, O.rel_error(ref, x))


def test_cholesky_solve(plans):
    """U^T U x = b: the transposed solve, then the solve with U, against the oracle's two
    triangular solves (1e-12), and the same x from every combination of launches."""
    for name in MATS:
        U, P = MATS[name], plans[name]
        for m in (1, 4):
            banner(f"cholesky {name} m={m}")
            xt, b = gen.rhs_for(U, m, seed=32, layout="row")
            if m == 1:
                b = b.reshape(-1)
            y = oracle_lower(U, b)
            ref = oracle_solve(U, y)
            x = P.solve_cholesky_host(b)
            assert O.rel_error(ref, x) <= FP64_TOL, (name, m, O.rel_error(ref, x))


def test_transposed_refuses_supernodal_and_times(plans):
    P = plans["snode_10"]
    U = MATS["snode_10"]
    b = np.ones(U.n)
    with pytest.raises(capi.SptrsvError):
        P.solve_transposed_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
    P.timing(True)
    P.solve_transposed_host(b)
    ms, launches = P.timing_read()
    P.timing(False)
    assert launches >= 1 and ms > 0.0


@pytest.mark.parametrize("name", ["csr_6", "snode_10", "sn_random", "one_row_snodes"])
def test_ccs2ccs_pad_device(name):
    """The device-built padded supernodal layout equals the oracle's ccs2ccs_pad (src/core.cpp:
    136-161) entry for entry; the size query agrees."""
    if name == "csr_6":
        n, rp, col, val = capi.read_matrix(os.path.join(GOLDEN, "csr_6_upper.txt"))
        U = gen.UpperCSR(n, rp, col, val)
        snodes = capi.read_snodes(os.path.join(GOLDEN, "csr_6_upper_snodes.txt"))
    elif name == "snode_10":
        U = MATS["snode_10"]
        snodes = U.snodes
    elif name == "sn_random":
        U = gen.random_supernodal(40, seed=9)
        snodes = U.snodes
    else:  # every row its own supernode: no padding
        U = MATS["zipf_3000"]
        snodes = np.arange(U.n + 1)
    banner(f"ccs2ccs_pad {name}")
    vp, cpad, rpp = O.ccs2ccs_pad(U.rowptr, U.col, U.val, snodes)
    with capi.Plan.from_upper(U) as P:
        gv, gc, gr = P.ccs2ccs_pad(snodes)
    assert np.array_equal(gr, rpp) and np.array_equal(gc, cpad) and np.array_equal(gv.view(np.uint64), vp.view(np.uint64))


def test_ccs2ccs_pad_device_rejects_bad_partition(plans):
    P = plans["snode_10"]
    with pytest.raises(capi.SptrsvError):
        P.ccs2ccs_pad(np.array([0, 5]))  # does not reach n
    with pytest.raises(capi.SptrsvError):
        P.ccs2ccs_pad(np.array([0, 0, MATS["snode_10"].n]))  # empty supernode


def test_auto_variant(plans):
    """AUTO sends near-chain factors (levels hold < 4 rows on average) to the chain kernel at
    m == 1, natural-order grid factors to the block kernel at m <= 3 in FP64, a plan with a
    supernodal layout (the caller's, or relaxed supernodes of a factor dense enough) to the
    supernodal kernel in FP64, and everything else to the
    level-ordered pull kernel (block and pull bitwise exact)."""
    assert plans["ic0_12"].auto_variant(1) == capi.VAR_BLOCK
    assert plans["ic0_12"].auto_variant(2) == capi.VAR_BLOCK
    assert plans["ic0_12"].auto_variant(1, capi.FP32) == capi.VAR_LEVEL_PULL
    assert plans["ic0_12"].auto_variant(3) == capi.VAR_BLOCK
    assert plans["ic0_12"].auto_variant(4) == capi.VAR_LEVEL_PULL
    assert plans["zipf_3000"].auto_variant(1) in (capi.VAR_CHAIN, capi.VAR_LEVEL_PULL)
    assert plans["zipf_3000"].auto_variant(2) == capi.VAR_LEVEL_PULL
    assert plans["chol2d_15"].auto_variant(1) != capi.VAR_BLOCK
    # dense chain-like factors run on relaxed supernodes (MFMA) at every m
    assert plans["dense_70"].auto_variant(1) == capi.VAR_SUPERNODAL
    assert plans["dense_70"].auto_variant(4) == capi.VAR_SUPERNODAL
    assert plans["dense_70"].auto_variant(1, capi.FP32) == capi.VAR_CHAIN
    # a supernodal layout built by the caller: the MFMA kernel in FP64, at every m
    S = gen.random_supernodal(12, seed=31)
    with capi.Plan.from_upper(S) as PS:
        PS.build_supernodes(S.snodes)
        assert PS.auto_variant(1) == capi.VAR_SUPERNODAL
        assert PS.auto_variant(8) == capi.VAR_SUPERNODAL
        assert PS.auto_variant(8, capi.FP32) == capi.VAR_LEVEL_PULL
        xt, b = gen.rhs_for(S, 8, seed=3, layout="row")
        assert O.rel_error(oracle_solve(S, b), PS.solve_host(b, capi.params())) <= 1e-12
    # a supernodal factor with many rows per level and no layout from the caller: relaxed
    # supernodes (>= 16 entries per row, >= 32 rows per supernode on average), built by AUTO
    S = gen.random_supernodal(64, seed=32, wmin=32, wmax=96)
    with capi.Plan.from_upper(S) as PS:
        assert PS.n_levels * 4 <= S.n and S.nnz >= 16 * S.n
        assert PS.auto_variant(8) == capi.VAR_SUPERNODAL
        assert PS.auto_variant(8, capi.FP32) == capi.VAR_LEVEL_PULL
        for m in (1, 8):
            xt, b = gen.rhs_for(S, m, seed=4, layout="row")
            assert O.rel_error(oracle_solve(S, b), PS.solve_host(b, capi.params())) <= 1e-12
    for name in MATS:
        U, P = MATS[name], plans[name]
        for m in (1, 2, 3, 4):
            xt, b = gen.rhs_for(U, m, seed=22, layout="row")
            ref = oracle_solve(U, b)
            x = P.solve_host(b, capi.params())
            v = P.auto_variant(m)
            if v in (capi.VAR_CHAIN, capi.VAR_SUPERNODAL):
                assert O.rel_error(ref, x) <= FP64_TOL, name
            else:
                assert np.array_equal(x, ref), (name, m)


def test_column_major_host_layout(plans):
    U, P = MATS["zipf_3000"], plans["zipf_3000"]
    xt, b = gen.rhs_for(U, 4, seed=6, layout="col")  # (m, n)
    x = P.solve_host(b, capi.params(variant=capi.VAR_LEVEL_PULL), col_major=True)
    ref, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, b)
    assert np.array_equal(x, ref)


@pytest.mark.parametrize("variant", EXACT + [capi.VAR_CSC_PUSH])
def test_fp32(plans, variant):
    for name in ("zipf_3000", "ic0_12", "chol2d_15"):
        U, P = MATS[name], plans[name]
        for m in (1, 16):
            xt, b = gen.rhs_for(U, m, seed=7, layout="row")
            x = P.solve_host(b, capi.params(variant=variant, precision=capi.FP32))
            assert O.rel_error(oracle_solve(U, b), x) <= FP32_TOL


def test_device_pointers_and_residual(plans):
    U, P = MATS["zipf_3000"], plans["zipf_3000"]
    xt, b = gen.rhs_for(U, 5, seed=8, layout="row")
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    d_r = capi.DeviceArray(b.shape)
    for _ in range(3):  # repeated solves on the same buffers reuse the plan state
        P.solve(d_b, d_x, 5, capi.params(variant=capi.VAR_LEVEL_PULL))
    P.sync()
    x = d_x.download()
    assert np.array_equal(x, oracle_solve(U, b))
    P.residual(d_b, d_x, d_r, 5)
    P.sync()
    r = d_r.download()
    rref = O.spmv_residual(U.rowptr, U.col, U.val, np.ascontiguousarray(x.T), np.ascontiguousarray(b.T)).T
    np.testing.assert_allclose(r, rref, rtol=0, atol=1e-13)


def test_watchdog_reports_and_recovers(plans):
    """A 1 us per-wait limit on a 4000-long chain forces the kernel's own bounded wait to give
    up (every lane exits normally); the host sees ERR_DEADLOCK, and the next solve is fine."""
    U, P = MATS["chain_4000"], plans["chain_4000"]
    xt, b = gen.rhs_for(U, 1, seed=9)
    with pytest.raises(capi.SptrsvError) as e:
        P.solve_host(b, capi.params(variant=capi.VAR_ROW_PULL, watchdog_us=1))
    assert e.value.code == capi.ERR_DEADLOCK
    x = P.solve_host(b, capi.params(variant=capi.VAR_ROW_PULL))
    assert np.array_equal(x, oracle_solve(U, b))


def test_full_size_ic0_128():
    """configs: 7-point IC(0) on 128^3 (n = 2,097,152): closed-form analysis and FP64 parity."""
    U = gen.laplace3d_ic0(128)
    banner("plan ic0_128")
    with capi.Plan.from_upper(U) as P:
        assert P.n_levels == 382
        N = 128
        idx = np.arange(U.n)
        x, y, z = idx % N, (idx // N) % N, idx // (N * N)
        assert np.array_equal(P.levels(), (1 + (N - 1 - x) + (N - 1 - y) + (N - 1 - z)).astype(np.int32))
        assert np.array_equal(P.indegree(), (1 + (x < N - 1) + (y < N - 1) + (z < N - 1)).astype(np.int32))
        xt, b = gen.rhs_for(U, 1, seed=1)
        ref = oracle_solve(U, b)
        for variant in (capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL, capi.VAR_TILE, capi.VAR_BLOCK):
            banner(f"solve ic0_128 {capi.VARIANT_NAMES[variant]}")
            got = P.solve_host(b, capi.params(variant=variant))
            assert np.array_equal(got, ref)
        got = P.solve_host(b, capi.params(variant=capi.VAR_CSC_PUSH))
        assert O.rel_error(ref, got) <= FP64_TOL
        assert O.rel_error(xt, ref) < 1e-10


def test_full_size_zipf_1e6_level_pull():
    U = gen.random_zipf(1_000_000, seed=1)
    banner("plan zipf_1e6")
    with capi.Plan.from_upper(U) as P:
        lv, L = O.levels(U.rowptr, U.col)
        assert P.n_levels == L
        assert np.array_equal(P.levels(), lv.astype(np.int32))
        xt, b = gen.rhs_for(U, 1, seed=1)
        ref = oracle_solve(U, b)
        for variant in (capi.VAR_ROW_PULL, capi.VAR_TILE, capi.VAR_BLOCK):
            banner(f"solve zipf_1e6 {capi.VARIANT_NAMES[variant]}")
            got = P.solve_host(b, capi.params(variant=variant))
            assert np.array_equal(got, ref)
        got = P.solve_host(b, capi.params(variant=capi.VAR_CHAIN))
        assert O.rel_error(ref, got) <= FP64_TOL


@pytest.mark.parametrize("rows", [1, 64, 1000, 1024, 4096])
def test_tile_sizes(plans, rows):
    """Tiles of `rows` consecutive rows: many tiles (cross-tile polling), ragged last tiles and
    single-row tiles; FP64 x stays bitwise equal to the oracle."""
    for name in ("zipf_3000", "ic0_12", "chain_4000", "chol2d_15", "snode_10"):
        U, P = MATS[name], plans[name]
        banner(f"tile rows={rows} {name}")
        xt, b = gen.rhs_for(U, 1, seed=23)
        x = P.solve_host(b, capi.params(variant=capi.VAR_TILE, block=rows))
        assert np.array_equal(x, oracle_solve(U, b)), (name, rows)
    with pytest.raises(capi.SptrsvError):
        P.solve_host(b, capi.params(variant=capi.VAR_TILE, block=20000))


def test_tile_rejects_multi_rhs(plans):
    U, P = MATS["zipf_3000"], plans["zipf_3000"]
    xt, b = gen.rhs_for(U, 2, seed=1, layout="row")
    with pytest.raises(capi.SptrsvError) as e:
        P.solve_host(b, capi.params(variant=capi.VAR_TILE))
    assert e.value.code == capi.ERR_UNSUPPORTED


def _build(src, out):
    lib_dir = os.path.dirname(capi.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", src), "-o", out, "-L", lib_dir, "-lsptrsv_amd",
                    f"-Wl,-rpath,{lib_dir}"], check=True)
    return out


def _write_txt(path, U):
    with open(path, "w") as f:
        f.write(f"{U.n} {U.nnz}\n")
        f.write(" ".join(repr(float(v)) for v in U.val) + "\n")
        f.write(" ".join(str(int(c)) for c in U.col) + "\n")
        f.write(" ".join(str(int(r)) for r in U.rowptr) + "\n")


@pytest.mark.parametrize("rhs", [1, 3])
@pytest.mark.parametrize("which", ["csr_6", "snode_6"])
def test_cpp_dropin_entry_points(tmp_path, which, rhs):
    """Every reference entry point (include/core.h, include/sptrsv_syncfree_opencl.h) called
    from C++ exactly as the reference's run() calls it, checked against the oracle."""
    exe = _build("dropin_solve.cpp", str(tmp_path / "dropin_solve"))
    if which == "csr_6":
        mpath = os.path.join(GOLDEN, "csr_6_upper.txt")
        spath = os.path.join(GOLDEN, "csr_6_upper_snodes.txt")
        tok = open(mpath).read().split()
        n, nz = int(tok[0]), int(tok[1])
        U = gen.UpperCSR(n, np.array([int(t) for t in tok[2 + 2 * nz:]]),
                         np.array([int(t) for t in tok[2 + nz:2 + 2 * nz]], dtype=np.int32),
                         np.array([float(t) for t in tok[2:2 + nz]]))
        snodes = np.array([0, 1, 3, 4, 6], dtype=np.int32)
    else:
        U = gen.random_supernodal(6, seed=4, wmin=4, wmax=12)
        mpath, spath = str(tmp_path / "m.txt"), str(tmp_path / "s.txt")
        _write_txt(mpath, U)
        with open(spath, "w") as f:
            f.write(f"{len(U.snodes) - 1}\n" + " ".join(str(int(s)) for s in U.snodes) + "\n")
        snodes = U.snodes
    out = subprocess.run([exe, mpath, spath, str(rhs), str(tmp_path)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    n = U.n
    cnt = np.arange(n * rhs)
    b = (1.0 + (cnt * 7) % 11).reshape(rhs, n)  # column-major as written by the C++ test program
    ref, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, b)
    load = lambda k: np.fromfile(str(tmp_path / f"{k}.bin"), dtype=np.float64)
    for k in ("base", "base_new", "barrier"):
        assert np.array_equal(load(k), ref.ravel()), k
    for k in ("custom", "blas"):
        xs, _ = O.supernodal_upper(U.rowptr, U.col, U.val, snodes, b)
        assert O.rel_error(xs.ravel(), load(k)) <= FP64_TOL, k
    lv, L = O.levels(U.rowptr, U.col)
    assert np.array_equal(load("levels"), lv.astype(np.float64))
    assert f"levels {L}" in out.stdout
    ref_row = np.ascontiguousarray(ref.T).ravel()
    assert O.rel_error(ref_row, load("syncfree")) <= FP64_TOL
    if rhs == 1:
        assert O.rel_error(ref_row, load("syncfree2")) <= FP64_TOL
        assert np.array_equal(load("write_first"), ref_row)
        cp, ri, cv = O.transpose(U.rowptr, U.col, U.val)
        lo, _ = O.base_gauss_lower(cp.astype(np.uint64), ri, cv, b[0])
        assert O.rel_error(lo, load("lower")) <= FP64_TOL
    assert "nonsquare -1.0" in out.stdout


# ---- supernodal (MFMA) ------------------------------------------------------------------------

@pytest.mark.parametrize("m", [1, 8, 16, 32, 100, 128])
def test_supernodal_given_partition(m):
    U = gen.random_supernodal(24, seed=6)
    banner(f"supernodal given m={m}")
    with capi.Plan.from_upper(U) as P:
        sn = P.build_supernodes(U.snodes)
        assert np.array_equal(sn, U.snodes)  # widths <= 128 stay unsplit
        xt, b = gen.rhs_for(U, m, seed=2, layout="row")
        x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
        ref = oracle_solve(U, b)
        assert O.rel_error(ref, x) <= FP64_TOL
        bb = b if m == 1 else np.ascontiguousarray(b.T)
        xs, _ = O.supernodal_upper(U.rowptr, U.col, U.val, U.snodes, bb)
        assert O.rel_error(xs if m == 1 else xs.T, x) <= FP64_TOL


def test_supernodal_one_column_kernels():
    """One right-hand side on supernodes of banded factors (the one-column kernel): a single
    band (its supernodes form one chain) and six independent bands side by side (a supernode
    DAG six wide), both against the oracle, the same x on a second solve, and with x over b."""
    one = gen.laplace2d_cholesky(30)
    blocks = [gen.laplace2d_cholesky(12) for _ in range(6)]
    rows, cols, vals, off = [], [], [], 0
    for B in blocks:
        r = np.repeat(np.arange(B.n), np.diff(B.rowptr))
        rows.append(r + off)
        cols.append(B.col.astype(np.int64) + off)
        vals.append(B.val)
        off += B.n
    six = gen._from_rows(off, np.concatenate(rows), np.concatenate(cols).astype(np.int32), np.concatenate(vals), "six")
    for name, U in (("band", one), ("six bands", six)):
        with capi.Plan.from_upper(U) as P:
            sn = P.build_supernodes()
            assert len(sn) > 2, name
            xt, b = gen.rhs_for(U, 1, seed=12)
            x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
            assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL, name
            x2 = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
            assert np.array_equal(x, x2), name  # fixed summation order
            # x over b (outside sptrsv_solve's contract): the one-column kernel, whose fill would
            # destroy b, is not taken; the MFMA kernel solves in place
            d = capi.DeviceArray.from_host(b)
            P.solve(d, d, 1, capi.params(variant=capi.VAR_SUPERNODAL))
            P.sync()
            assert O.rel_error(oracle_solve(U, b), d.download()) <= FP64_TOL, name


def test_supernodal_detected_and_split():
    U = gen.random_supernodal(16, seed=8, wmin=100, wmax=300)  # wider than 128 -> split
    with capi.Plan.from_upper(U) as P:
        sn = P.build_supernodes()
        assert np.all(np.diff(sn) <= 128)
        # detection may merge neighbouring generated supernodes whose patterns nest, but it
        # never cuts a generated supernode except to keep widths <= 128
        gen_b = set(U.snodes.tolist())
        for k in range(len(U.snodes) - 1):
            inside = [s for s in sn.tolist() if U.snodes[k] < s < U.snodes[k + 1]]
            assert all((s - U.snodes[k]) % 128 == 0 or s in gen_b for s in inside) or \
                all(np.diff([U.snodes[k]] + inside + [U.snodes[k + 1]]) <= 128)
        xt, b = gen.rhs_for(U, 24, seed=3, layout="row")
        x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
        assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL


def test_supernodal_on_general_pattern_and_fixture(plans, csr6, snodes6):
    for name in ("zipf_3000", "chain_4000", "chol2d_15"):
        U, P = MATS[name], plans[name]
        P.build_supernodes()
        for m in (1, 4):  # m = 1: the one-column kernel
            xt, b = gen.rhs_for(U, m, seed=4, layout="row")
            x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
            assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL, (name, m)
    n, rp, col, val = csr6
    with capi.Plan.from_csr(rp, col, val) as P:
        P.build_supernodes(snodes6)
        x = P.solve_host(np.ones(n), capi.params(variant=capi.VAR_SUPERNODAL))
        np.testing.assert_allclose(x, [0.5, 7 / 8, 1 / 12, 0.5, 5 / 24, 1 / 6], rtol=0, atol=1e-15)


def test_supernodal_rejects_bad_partition(plans):
    P = plans["zipf_3000"]
    with pytest.raises(capi.SptrsvError) as e:
        P.build_supernodes(np.array([0, 7, 3000], dtype=np.int32))
    assert e.value.code == capi.ERR_BAD_INPUT
    with pytest.raises(capi.SptrsvError):
        P.build_supernodes(np.array([0, 3001], dtype=np.int32))


def test_supernodal_full_size_4096():
    """configs: random supernodal factor, 4096 supernodes, widths U[16,128]."""
    U = gen.random_supernodal(4096, seed=2)
    banner(f"plan supernodal_4096 n={U.n} nnz={U.nnz}")
    with capi.Plan.from_upper(U) as P:
        P.build_supernodes(U.snodes)
        for m in (1, 8):  # m = 1: the one-column kernel, rectangular blocks up to 6 x 128 columns
            xt, b = gen.rhs_for(U, m, seed=2, layout="row")
            banner(f"solve supernodal_4096 m={m}")
            x = P.solve_host(b, capi.params(variant=capi.VAR_SUPERNODAL))
            ref = oracle_solve(U, b)
            assert O.rel_error(ref, x) <= FP64_TOL
            xl = P.solve_host(b, capi.params(variant=capi.VAR_LEVEL_PULL))
            assert np.array_equal(xl, ref)


# ---- mixed precision with FP64 refinement -----------------------------------------------------

def _refined(P, b, m, variant, precision, tol=0.0, max_iters=0):
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    it, ratio = P.solve_refined(d_b, d_x, m, capi.params(variant=variant), precision=precision, tol=tol,
                                max_iters=max_iters)
    return d_x.download(), it, ratio


REFINE_CASES = [(v, m) for m in (1, 8)
                for v in (capi.VAR_LEVEL_PULL, capi.VAR_ROW_PULL, capi.VAR_LEVELSET, capi.VAR_CSC_PUSH,
                          capi.VAR_CHAIN, capi.VAR_AUTO)
                if not (v == capi.VAR_CHAIN and m > 1)]  # the chain kernel solves one right-hand side


@pytest.mark.parametrize("variant,m", REFINE_CASES)
def test_fp32_refined_reaches_fp64(plans, variant, m):
    for name in ("zipf_3000", "ic0_12", "chol2d_15", "snode_10"):
        U, P = MATS[name], plans[name]
        banner(f"refine fp32 {name} m={m} {capi.VARIANT_NAMES.get(variant, 'auto')}")
        xt, b = gen.rhs_for(U, m, seed=10, layout="row")
        x, it, est = _refined(P, b, m, variant, capi.FP32)
        assert est <= 1e-13 and 1 <= it <= 4, (name, it, est)
        assert O.rel_error(oracle_solve(U, b), x) <= FP64_TOL, name


@pytest.mark.parametrize("m", [1, 8, 32, 128])
def test_bf16_supernodal_refined_reaches_fp64(m):
    U = gen.random_supernodal(24, seed=6)
    with capi.Plan.from_upper(U) as P:
        P.build_supernodes(U.snodes)
        xt, b = gen.rhs_for(U, m, seed=11, layout="row")
        x, it, est = _refined(P, b, m, capi.VAR_SUPERNODAL, capi.BF16)
        assert est <= 1e-13 and it <= 12, (it, est)
        ref = oracle_solve(U, b)
        assert O.rel_error(ref, x) <= FP64_TOL
        # the FP64 kernel after the reduced-precision one on the same plan (shared epoch flags)
        for mm in (m, 2 * m + 3):
            xt2, b2 = gen.rhs_for(U, mm, seed=12, layout="row")
            x2 = P.solve_host(b2, capi.params(variant=capi.VAR_SUPERNODAL))
            assert O.rel_error(oracle_solve(U, b2), x2) <= FP64_TOL, mm


def test_refinement_reports_non_convergence():
    """diag_scale << 1 removes the diagonal dominance: FP32 solves cannot contract the error,
    and the refinement loop must say so through the returned estimate instead of claiming
    convergence."""
    U = gen.random_zipf(2000, seed=12, diag_scale=0.02)
    with capi.Plan.from_upper(U) as P:
        xt, b = gen.rhs_for(U, 1, seed=1)
        x, it, est = _refined(P, b, 1, capi.VAR_LEVEL_PULL, capi.FP32, max_iters=3)
        ref = oracle_solve(U, b)
        assert 0 <= it <= 3  # 0: the first FP32 iterate was already not finite
        if O.rel_error(ref, x) > FP64_TOL:
            assert not (est <= 1e-13), (it, est)
        d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
        info = P.solve_refined_ex(d_b, d_x, 1, capi.params(variant=capi.VAR_LEVEL_PULL), max_iters=3)
        assert info["outcome"] != "converged", info
        info = P.solve_refined_ex(d_b, d_x, 1, capi.params(variant=capi.VAR_LEVEL_PULL), max_iters=3, fallback=True)
        assert info["fell_back"], info
        # x is the FP64 AUTO solve (the chain kernel here, deterministic: one wave)
        assert np.array_equal(d_x.download(), P.solve_host(b, capi.params()), equal_nan=True)


def test_refinement_rejects_unsupported(plans):
    U, P = MATS["zipf_3000"], plans["zipf_3000"]
    xt, b = gen.rhs_for(U, 1, seed=1)
    for variant, prec in ((capi.VAR_TILE, capi.FP32), (capi.VAR_LEVEL_PULL, capi.BF16)):
        with pytest.raises(capi.SptrsvError) as e:
            _refined(P, b, 1, variant, prec)
        assert e.value.code == capi.ERR_UNSUPPORTED


def test_refinement_overflow_is_divergence_and_falls_back():
    """A factor whose solution overflows FP32 (x grows 1000x per row along a chain) but not FP64:
    the FP32 solve yields inf/NaN, which the norms must keep (a max that drops NaN would report
    convergence).  The refinement reports divergence; with the fallback x is the FP64 solve."""
    n = 40  # x reaches 1e117: beyond FP32 after 13 rows, and its squares stay finite in FP64
    rows = np.concatenate([np.arange(n), np.arange(n - 1)])
    cols = np.concatenate([np.arange(n), np.arange(1, n)])
    vals = np.concatenate([np.full(n, 1e-3), np.full(n - 1, 1.0)])
    o = np.lexsort((cols, rows))
    U = gen._from_rows(n, rows[o], cols[o], vals[o], "overflow_chain")
    b = np.ones(n)
    ref = oracle_solve(U, b)
    assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 1e100 and np.isfinite(O.rel_error(ref, ref))
    with capi.Plan.from_upper(U) as P:
        for variant in (capi.VAR_LEVEL_PULL, capi.VAR_CHAIN):
            d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
            info = P.solve_refined_ex(d_b, d_x, 1, capi.params(variant=variant))
            assert info["outcome"] == "diverged" and not info["fell_back"], info
            assert not np.isfinite(info["estimate"]) or info["estimate"] > 1e-13, info
            it, est = P.solve_refined(d_b, d_x, 1, capi.params(variant=variant))
            assert not (est <= 1e-13), (it, est)
            info = P.solve_refined_ex(d_b, d_x, 1, capi.params(variant=variant), fallback=True)
            assert info["fell_back"], info
            assert O.rel_error(ref, d_x.download()) <= FP64_TOL


@pytest.mark.parametrize("name", ["chol2d_15", "zipf_3000", "chain_4000", "csr_5"])
def test_auto_fp32_chain_like(plans, name):
    """AUTO at FP32 picks the chain kernel on chain-like factors and runs it in FP32."""
    U, P = MATS[name], plans[name]
    assert P.n_levels * 4 > U.n and P.auto_variant(1, capi.FP32) == capi.VAR_CHAIN
    xt, b = gen.rhs_for(U, 1, seed=3)
    x = P.solve_host(b, capi.params(precision=capi.FP32))
    assert O.rel_error(oracle_solve(U, b), x) <= FP32_TOL


@pytest.mark.parametrize("m", [1, 2, 5])
@pytest.mark.parametrize("name", list(MATS))
def test_exact_auto_is_bitwise(plans, name, m):
    """params.exact: AUTO restricted to kernels that reproduce the serial solver bit for bit."""
    U, P = MATS[name], plans[name]
    xt, b = gen.rhs_for(U, m, seed=31, layout="row")
    x = P.solve_host(b, capi.params(exact=True))
    assert np.array_equal(x, oracle_solve(U, b))


@pytest.mark.parametrize("name", ["csr_5", "zipf_3000", "ic0_12", "chol2d_15", "dense_70", "chain_4000"])
def test_condest_against_dense(plans, name):
    """kappa_1 estimate against the dense 1-norm condition number (numpy): a lower bound, and on
    these factors within a factor 3 (Higham's estimator is usually exact)."""
    U, P = MATS[name], plans[name]
    if U.n > 4000:
        pytest.skip("dense reference too large")
    D = U.to_scipy().toarray()
    true = np.linalg.norm(D, 1) * np.linalg.norm(np.linalg.inv(D), 1)
    kappa, ninv = P.condest()
    assert kappa <= true * (1 + 1e-8), (kappa, true)
    assert kappa >= true / 3, (kappa, true)


def relaxed_bounds_ref(U, max_width=128, max_fill=1.0):
    """The relaxed-supernode rule restated: rows join the current block while the block padded to
    a dense triangle plus the union of its rows' columns right of the block keeps at most
    max_width rows and at most (1 + max_fill) times its rows' own entries."""
    bounds, s0, cols, orig = [0], 0, set(), 0
    i = 0
    while i < U.n:
        row = U.col[U.rowptr[i] + 1:U.rowptr[i + 1]].tolist()
        w = i - s0
        uni = {c for c in cols | set(row) if c > i}
        orig2 = orig + len(row) + 1
        padded = (w + 1) * (w + 2) // 2 + (w + 1) * len(uni)
        if w > 0 and (w + 1 > max_width or padded - orig2 > max_fill * orig2):
            bounds.append(i)
            s0, cols, orig = i, set(), 0
            continue
        cols |= set(row)
        orig = orig2
        i += 1
    return bounds + [U.n]


@pytest.mark.parametrize("name", ["chol2d_15", "zipf_3000", "dense_70", "ic0_12", "snode_10", "csr_5"])
def test_relaxed_supernodes(name):
    """Relaxed supernodes on any factor: the partition follows the rule (restated above), every
    block holds its rows' entries plus explicit zeros, and the MFMA solve matches the oracle."""
    U = MATS[name]
    with capi.Plan.from_upper(U) as P:
        for width, fill in ((0, 0.0), (16, 0.25)):
            sb, zeros = P.build_supernodes_relaxed(width, fill)
            ref_b = relaxed_bounds_ref(U, width or 128, fill or 1.0)
            # blocks wider than 128 are split by the layout; the rule never makes one here
            assert sb.tolist() == ref_b, (name, width, fill)
            dense = sum((b - a) * (b - a + 1) // 2 for a, b in zip(ref_b[:-1], ref_b[1:]))
            rect = 0
            for a, b in zip(ref_b[:-1], ref_b[1:]):
                uni = set()
                for r in range(a, b):
                    uni |= {c for c in U.col[U.rowptr[r]:U.rowptr[r + 1]].tolist() if c >= b}
                rect += (b - a) * len(uni)
            assert zeros == dense + rect - U.nnz
            for m in (1, 5, 40):
                xt, bb = gen.rhs_for(U, m, seed=41, layout="row")
                x = P.solve_host(bb, capi.params(variant=capi.VAR_SUPERNODAL))
                assert O.rel_error(oracle_solve(U, bb), x) <= FP64_TOL, (name, m)
