"""ctypes front-end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (as the checker)
and bench.py's cpu_baseline leg (as the timed CPU baseline).  The shipped package never
imports this module.

All matrices are upper-triangular CSR with the diagonal first in each row (the reference's
convention, src/core.cpp:18-32); right-hand sides are column-major n x m arrays stored as
numpy arrays of shape (m, n) (x[rh, i] == x[i + rh*n]).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_c_int, _c_dbl, _c_i64 = ctypes.c_int, ctypes.c_double, ctypes.c_int64


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True, capture_output=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_base_gauss_upper": (_c_dbl, [_c_int, _dp, _u64p, _i32p, _dp, _dp, _c_int]),
            "orc_base_gauss_upper_new": (_c_dbl, [_c_int, _dp, _u64p, _i32p, _dp, _dp, _c_int]),
            "orc_base_gauss_lower": (_c_dbl, [_c_int, _dp, _u64p, _i32p, _dp, _dp]),
            "orc_set_level_up": (_c_int, [_c_int, _i32p, _u64p, _u64p]),
            "orc_level_sets": (None, [_c_int, _u64p, _c_int, _i32p, _i32p]),
            "orc_tasks_by_level": (None, [_u64p, _i32p, _i32p, _c_int, _c_int, _i32p]),
            "orc_barrier_upper": (_c_dbl, [_c_int, _dp, _dp, _dp, _i32p, _u64p, _c_int, _c_int,
                                           ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl)]),
            "orc_barrier_upper_new": (_c_dbl, [_c_int, _dp, _dp, _dp, _i32p, _u64p, _c_int, _c_int,
                                               ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl)]),
            "orc_indegree_csc": (None, [_c_int, _c_i64, _i32p, _i32p]),
            "orc_transpose": (_c_dbl, [_c_int, ctypes.c_uint64, _dp, _i32p, _u64p, _dp, _u64p, _i32p]),
            "orc_supernodal_upper": (_c_dbl, [_c_int, _i32p, _dp, _dp, _i32p, _u64p, _c_int, _c_int]),
            "orc_supernodal_upper_new": (_c_dbl, [_c_int, _i32p, _dp, _dp, _i32p, _u64p, _c_int, _c_int]),
            "orc_pad_extra": (ctypes.c_uint64, [_c_int, _i32p]),
            "orc_ccs2ccs_pad": (_c_dbl, [_dp, _i32p, _u64p, _dp, _u64p, _i32p, _i32p, _c_int,
                                         ctypes.POINTER(_c_i64)]),
            "orc_supernodal_blas_upper": (_c_dbl, [_c_int, _c_int, _i32p, _dp, _dp, _u64p, _i32p, _c_int]),
            "orc_rel_error": (_c_dbl, [_c_i64, _dp, _dp]),
            "orc_supernodal_blas_upper_cblas": (_c_dbl, [ctypes.c_char_p, _c_int, _c_int, _i32p, _dp, _dp, _u64p,
                                                         _i32p, _c_int]),
            "orc_spmv_residual": (None, [_c_int, _dp, _u64p, _i32p, _dp, _dp, _dp, _c_int]),
            "orc_wtime": (_c_dbl, []),
            "orc_max_threads": (_c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _csr(rowptr, col, val):
    return (np.ascontiguousarray(rowptr, dtype=np.uint64),
            np.ascontiguousarray(col, dtype=np.int32),
            np.ascontiguousarray(val, dtype=np.float64))


def _rhs(b, n):
    b = np.ascontiguousarray(b, dtype=np.float64)
    if b.ndim == 1:
        b = b.reshape(1, n)
    return b


def base_gauss_upper(rowptr, col, val, b, new=False):
    """Serial backward substitution (src/core.cpp:18-51). b: (n,) or (m, n). Returns (x, seconds)."""
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    bb = _rhs(b, n)
    x = np.zeros_like(bb)
    fn = lib().orc_base_gauss_upper_new if new else lib().orc_base_gauss_upper
    t = fn(n, v, rp, c, x, bb, bb.shape[0])
    return (x.reshape(np.shape(b)), t)


def base_gauss_upper_columns(rowptr, col, val, b, workers=16):
    """base_gauss_upper over the columns of b (m, n) split across host threads.  The serial
    solve's RHS loop is outermost (src/core.cpp:21-30), so every column is computed exactly as in
    one call: the result is bitwise that of base_gauss_upper.  ctypes releases the GIL."""
    from concurrent.futures import ThreadPoolExecutor
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    bb = _rhs(b, n)
    m = bb.shape[0]
    x = np.zeros_like(bb)
    w = max(1, min(workers, m))
    cuts = np.linspace(0, m, w + 1).astype(int)

    def part(k):
        k0, k1 = cuts[k], cuts[k + 1]
        if k1 > k0:
            lib().orc_base_gauss_upper(n, v, rp, c, x[k0:k1], np.ascontiguousarray(bb[k0:k1]), int(k1 - k0))

    with ThreadPoolExecutor(w) as ex:
        list(ex.map(part, range(w)))
    return x.reshape(np.shape(b))


def base_gauss_lower(rowptr, col, val, b):
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(bb)
    t = lib().orc_base_gauss_lower(n, v, rp, c, x, bb)
    return x, t


def levels(rowptr, col):
    """1-based levels and the max level (src/core.cpp:244-270)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    c = np.ascontiguousarray(col, dtype=np.int32)
    n = len(rp) - 1
    lv = np.zeros(n, dtype=np.uint64)
    L = lib().orc_set_level_up(n, c, rp, lv)
    return lv, L


def level_sets(levels_arr, max_level):
    """(level_ptr[max_level+1], level_rows[n]) (src/core.cpp:282-287)."""
    lv = np.ascontiguousarray(levels_arr, dtype=np.uint64)
    n = len(lv)
    ptr = np.zeros(max_level + 1, dtype=np.int32)
    rows = np.zeros(n, dtype=np.int32)
    lib().orc_level_sets(n, lv, max_level, ptr, rows)
    return ptr, rows


def tasks_by_level(rowptr, level_ptr, level_rows, max_level, nthreads):
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    tasks = np.zeros((max_level + 1, nthreads + 1), dtype=np.int32)
    lib().orc_tasks_by_level(rp, np.ascontiguousarray(level_ptr, dtype=np.int32),
                             np.ascontiguousarray(level_rows, dtype=np.int32),
                             max_level, nthreads, tasks)
    return tasks


def barrier_upper(rowptr, col, val, b, nthreads, new=False):
    """OpenMP level-set solve (src/core.cpp:273-483). Returns (x, t_analysis, t_solve)."""
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    bb = _rhs(b, n)
    x = np.zeros_like(bb)
    ta, ts = _c_dbl(0.0), _c_dbl(0.0)
    fn = lib().orc_barrier_upper_new if new else lib().orc_barrier_upper
    fn(n, x, bb, v, c, rp, int(nthreads), bb.shape[0], ctypes.byref(ta), ctypes.byref(ts))
    return x.reshape(np.shape(b)), ta.value, ts.value


def indegree_csc(n, rowidx):
    """In-degree per row incl. the diagonal (src/kernel_full.cl:40-52)."""
    ri = np.ascontiguousarray(rowidx, dtype=np.int32)
    out = np.zeros(n, dtype=np.int32)
    lib().orc_indegree_csc(n, len(ri), ri, out)
    return out


def transpose(rowptr, col, val):
    """CSR(U) -> CSC(U) counting sort (src/utils.cpp:128-159). Returns (colptr, rowidx, val_t)."""
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    nz = len(c)
    val_t = np.zeros(nz, dtype=np.float64)
    row_t = np.zeros(nz, dtype=np.uint64)
    colptr = np.zeros(n + 1, dtype=np.int32)
    lib().orc_transpose(n, nz, v, c, rp, val_t, row_t, colptr)
    return colptr, row_t.astype(np.int32), val_t


def supernodal_upper(rowptr, col, val, snodes, b, new=False):
    """Supernodal custom solve, in place on a copy of b (src/core.cpp:53-133)."""
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    sn_arr = np.ascontiguousarray(snodes, dtype=np.int32)
    x = _rhs(b, n).copy()
    fn = lib().orc_supernodal_upper_new if new else lib().orc_supernodal_upper
    t = fn(len(sn_arr) - 1, sn_arr, x, v, c, rp, n, x.shape[0])
    return x.reshape(np.shape(b)), t


def ccs2ccs_pad(rowptr, col, val, snodes):
    """Padded supernodal format (src/core.cpp:136-161). Returns (val_pad, col_pad, rowptr_pad)."""
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    sn_arr = np.ascontiguousarray(snodes, dtype=np.int32)
    sn = len(sn_arr) - 1
    extra = lib().orc_pad_extra(sn, sn_arr)
    total = len(c) + int(extra)
    val_pad = np.zeros(total, dtype=np.float64)
    col_pad = np.zeros(total, dtype=np.uint64)
    rowptr_pad = np.zeros(n + 1, dtype=np.int32)
    nnz_pad = _c_i64(0)
    lib().orc_ccs2ccs_pad(v, c, rp, val_pad, col_pad, rowptr_pad, sn_arr, sn, ctypes.byref(nnz_pad))
    assert nnz_pad.value == total
    return val_pad, col_pad, rowptr_pad


def supernodal_blas_upper(val_pad, col_pad, rowptr_pad, snodes, b):
    """BLAS-shaped supernodal solve on the padded format (src/core.cpp:164-201)."""
    sn_arr = np.ascontiguousarray(snodes, dtype=np.int32)
    rpp = np.ascontiguousarray(rowptr_pad, dtype=np.int32)
    n = len(rpp) - 1
    x = _rhs(b, n).copy()
    t = lib().orc_supernodal_blas_upper(n, len(sn_arr) - 1, sn_arr, x,
                                        np.ascontiguousarray(val_pad, dtype=np.float64),
                                        np.ascontiguousarray(col_pad, dtype=np.uint64), rpp, x.shape[0])
    return x.reshape(np.shape(b)), t


CBLAS_CANDIDATES = ("/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so",
                    "libopenblas.so", "libopenblas.so.0", "libcblas.so", "libcblas.so.3")


def find_cblas():
    """First CBLAS library (exporting cblas_dgemm and cblas_dtrsm) that loads here, or None."""
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")  # see orc_supernodal_blas_upper_cblas
    for path in CBLAS_CANDIDATES:
        try:
            h = ctypes.CDLL(path)
            getattr(h, "cblas_dgemm")
            getattr(h, "cblas_dtrsm")
            return path
        except (OSError, AttributeError):
            continue
    return None


def supernodal_blas_upper_cblas(blas_path, val_pad, col_pad, rowptr_pad, snodes, b):
    """supernodal_blas_upper (src/core.cpp:164-201) with the BLAS calls made through the CBLAS
    library at blas_path.  In place on a copy of b; returns (x, seconds)."""
    sn_arr = np.ascontiguousarray(snodes, dtype=np.int32)
    rpp = np.ascontiguousarray(rowptr_pad, dtype=np.int32)
    n = len(rpp) - 1
    x = _rhs(b, n).copy()
    t = lib().orc_supernodal_blas_upper_cblas(os.fsencode(blas_path), n, len(sn_arr) - 1, sn_arr, x,
                                              np.ascontiguousarray(val_pad, dtype=np.float64),
                                              np.ascontiguousarray(col_pad, dtype=np.uint64), rpp, x.shape[0])
    if t < 0:
        raise OSError(f"cannot load cblas_dgemm / cblas_dtrsm from {blas_path}")
    return x.reshape(np.shape(b)), t


def rel_error(x_ref, x):
    """Relative L2 error (src/utils.cpp:355-380), both in one layout."""
    a = np.ascontiguousarray(x_ref, dtype=np.float64).ravel()
    bb = np.ascontiguousarray(x, dtype=np.float64).ravel()
    return lib().orc_rel_error(len(a), a, bb)


def spmv_residual(rowptr, col, val, x, b):
    rp, c, v = _csr(rowptr, col, val)
    n = len(rp) - 1
    xx = _rhs(x, n)
    bb = _rhs(b, n)
    r = np.zeros_like(bb)
    lib().orc_spmv_residual(n, v, rp, c, xx, bb, r, bb.shape[0])
    return r.reshape(np.shape(b))


def wtime():
    return lib().orc_wtime()


def max_threads():
    return lib().orc_max_threads()
