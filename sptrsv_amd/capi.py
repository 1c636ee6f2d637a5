"""ctypes bindings of libsptrsv_amd.so (the C ABI in include/sptrsv_capi.h).

The shared library is the product: this module only loads it and converts numpy arrays.
Importing it without a built library raises immediately -- there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libsptrsv_amd.so")

OK = 0
ERR_BAD_INPUT = -1
ERR_HIP = -2
ERR_DEADLOCK = -3
ERR_IO = -4
ERR_NOMEM = -5
ERR_UNSUPPORTED = -6
ERR_NO_DEVICE = -7
NOALLOC_MAX_ROWS = 1 << 21
NOALLOC_MAX_ELEMS = 1 << 21

VAR_AUTO = 0
VAR_CSC_PUSH = 1
VAR_ROW_PULL = 2
VAR_LEVEL_PULL = 3
VAR_LEVELSET = 4
VAR_CHAIN = 5
VAR_SUPERNODAL = 6
VAR_TILE = 7
VAR_BLOCK = 8
VAR_PANEL = 9
VARIANT_NAMES = {VAR_AUTO: "auto", VAR_CSC_PUSH: "csc_push", VAR_ROW_PULL: "row_pull", VAR_LEVEL_PULL: "level_pull",
                 VAR_LEVELSET: "levelset", VAR_CHAIN: "chain",
                 VAR_SUPERNODAL: "supernodal", VAR_TILE: "tile", VAR_BLOCK: "block",
                 VAR_PANEL: "panel"}

FP64 = 0
FP32 = 1
BF16 = 2


class SptrsvError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: status {code}: {msg}")
        self.code = code


REFINE_CONVERGED, REFINE_FLOOR, REFINE_MAX_ITERS, REFINE_DIVERGED = range(4)
REFINE_OUTCOMES = ["converged", "floor", "max_iters", "diverged"]


class RefineOpts(ctypes.Structure):
    _fields_ = [("max_iters", ctypes.c_int), ("precision", ctypes.c_int), ("tol", ctypes.c_double),
                ("fallback", ctypes.c_int), ("reserved", ctypes.c_int)]


class RefineInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("outcome", ctypes.c_int), ("fell_back", ctypes.c_int),
                ("variant", ctypes.c_int), ("estimate", ctypes.c_double)]

    def as_dict(self):
        return {"iters": self.iters, "outcome": REFINE_OUTCOMES[self.outcome], "fell_back": bool(self.fell_back),
                "variant": self.variant, "estimate": self.estimate}


class Params(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("group", ctypes.c_int), ("unroll", ctypes.c_int),
                ("precision", ctypes.c_int), ("watchdog_us", ctypes.c_int),
                ("occupancy", ctypes.c_int), ("block", ctypes.c_int), ("exact", ctypes.c_int)]


_lib = None
_vp, _i, _i64, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_size_t
_PLAN = ctypes.c_void_p


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    """Load the shared library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m sptrsv_amd.build` "
                          "(the solver has no non-GPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "sptrsv_last_error": (ctypes.c_char_p, []),
        "sptrsv_status_string": (ctypes.c_char_p, [_i]),
        "sptrsv_device_init": (_i, [_i]),
        "sptrsv_device_info": (_i, [ctypes.c_char_p, _i, ctypes.POINTER(_i), ctypes.POINTER(_sz)]),
        "sptrsv_device_sync": (_i, []),
        "sptrsv_division_check": (_i, [_vp, _vp, _vp, _vp, ctypes.c_int64]),
        "sptrsv_malloc": (_i, [ctypes.POINTER(_vp), _sz]),
        "sptrsv_free": (_i, [_vp]),
        "sptrsv_memcpy_h2d": (_i, [_vp, _vp, _sz]),
        "sptrsv_memcpy_d2h": (_i, [_vp, _vp, _sz]),
        "sptrsv_memcpy_d2d": (_i, [_vp, _vp, _sz]),
        "sptrsv_validate_csr": (_i, [_i, _i64, _vp, _vp, _vp]),
        "sptrsv_validate_csc": (_i, [_i, _i64, _vp, _vp, _vp]),
        "sptrsv_csr_to_csc": (_i, [_i, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
        "sptrsv_plan_create_csr": (_i, [_i, _i64, _vp, _vp, _vp, ctypes.POINTER(_PLAN)]),
        "sptrsv_plan_create_csc": (_i, [_i, _i64, _vp, _vp, _vp, ctypes.POINTER(_PLAN)]),
        "sptrsv_plan_destroy": (_i, [_PLAN]),
        "sptrsv_plan_info": (_i, [_PLAN, ctypes.POINTER(_i), ctypes.POINTER(_i64), ctypes.POINTER(_i)]),
        "sptrsv_plan_analysis_time": (_i, [_PLAN, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "sptrsv_plan_get_indegree": (_i, [_PLAN, _vp]),
        "sptrsv_plan_get_levels": (_i, [_PLAN, _vp]),
        "sptrsv_plan_get_blocks": (_i, [_PLAN, _vp, _vp, _vp, _vp]),
        "sptrsv_plan_auto_variant": (_i, [_PLAN, _i, _i, _vp]),
        "sptrsv_plan_get_level_sets": (_i, [_PLAN, _vp, _vp]),
        "sptrsv_plan_get_csc": (_i, [_PLAN, _vp, _vp, _vp]),
        "sptrsv_solve": (_i, [_PLAN, ctypes.POINTER(Params), _vp, _vp, _i]),
        "sptrsv_plan_sync": (_i, [_PLAN]),
        "sptrsv_solve_host": (_i, [_PLAN, ctypes.POINTER(Params), _vp, _vp, _i, _i]),
        "sptrsv_plan_timing": (_i, [_PLAN, _i]),
        "sptrsv_plan_timing_read": (_i, [_PLAN, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i)]),
        "sptrsv_residual": (_i, [_PLAN, _vp, _vp, _vp, _i]),
        "sptrsv_solve_transposed": (_i, [_PLAN, ctypes.POINTER(Params), _vp, _vp, _i]),
        "sptrsv_read_matrix_file": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i64),
                                         ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
        "sptrsv_read_snodes_file": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_vp)]),
        "sptrsv_free_host": (None, [_vp]),
        "sptrsv_ccs2ccs_pad_device": (_i, [_PLAN, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_i64)]),
        "sptrsv_solve_cholesky": (_i, [_PLAN, ctypes.POINTER(Params), _vp, _vp, _i]),
        "sptrsv_plan_build_supernodes": (_i, [_PLAN, _i, _vp]),
        "sptrsv_batch_create": (_i, [_i, _vp, _i64, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
        "sptrsv_batch_destroy": (_i, [_vp]),
        "sptrsv_batch_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i64)]),
        "sptrsv_batch_features": (_i, [_vp, _i, _vp]),
        "sptrsv_plan_features": (_i, [_vp, _i, _vp]),
        "sptrsv_selector_set": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp]),
        "sptrsv_selector_clear": (_i, []),
        "sptrsv_batch_set_configs": (_i, [_vp, _vp, _vp]),
        "sptrsv_batch_solve": (_i, [_vp, _i, _vp, _vp, _i, _i]),
        "sptrsv_batch_sync": (_i, [_vp]),
        "sptrsv_batch_solve_refined": (_i, [_vp, ctypes.POINTER(RefineOpts), _vp, _vp, _i, ctypes.POINTER(_i),
                                            ctypes.POINTER(ctypes.c_double)]),
        "sptrsv_batch_residual": (_i, [_vp, _vp, _vp, _vp, _i]),
        "sptrsv_batch_time_config": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp]),
        "sptrsv_mlp_create": (_i, [_i, _i, _i, _i, ctypes.c_uint, ctypes.POINTER(_vp)]),
        "sptrsv_mlp_destroy": (_i, [_vp]),
        "sptrsv_mlp_num_params": (_i, [_vp, ctypes.POINTER(_i)]),
        "sptrsv_mlp_get_params": (_i, [_vp, _vp]),
        "sptrsv_mlp_set_params": (_i, [_vp, _vp]),
        "sptrsv_mlp_loss_grad": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp]),
        "sptrsv_mlp_forward": (_i, [_vp, _vp, _i, _vp]),
        "sptrsv_mlp_train": (_i, [_vp, _vp, _vp, _vp, _i, _i, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
        "sptrsv_solve_refined": (_i, [_PLAN, ctypes.POINTER(Params), ctypes.POINTER(RefineOpts), _vp, _vp, _i,
                                      ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double)]),
        "sptrsv_plan_get_supernodes": (_i, [_PLAN, ctypes.POINTER(_i), _vp]),
        "sptrsv_plan_build_supernodes_relaxed": (_i, [_PLAN, _i, ctypes.c_double]),
        "sptrsv_plan_supernodal_padding": (_i, [_PLAN, ctypes.POINTER(_i64)]),
        "sptrsv_solve_refined_ex": (_i, [_PLAN, ctypes.POINTER(Params), ctypes.POINTER(RefineOpts), _vp, _vp, _i,
                                         ctypes.POINTER(RefineInfo)]),
        "sptrsv_batch_solve_refined_ex": (_i, [_vp, ctypes.POINTER(RefineOpts), _vp, _vp, _i,
                                               ctypes.POINTER(RefineInfo)]),
        "sptrsv_condest": (_i, [_PLAN, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "sptrsv_solve_push_noalloc": (_i, [_PLAN, _vp, _vp, _i]),
        "sptrsv_gen_batch_sizes": (_i, [_i, _i, ctypes.c_uint64, _i, _i, _vp, _vp, _vp]),
        "sptrsv_gen_batch_fill": (_i, [_i, _i, ctypes.c_uint64, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, where: str):
    if rc != OK:
        raise SptrsvError(rc, where, lib().sptrsv_last_error().decode())


def params(variant=VAR_AUTO, group=0, unroll=0, precision=FP64, watchdog_us=0, occupancy=0, block=0,
           exact=False) -> Params:
    p = Params()
    p.variant, p.group, p.unroll, p.precision, p.watchdog_us = variant, group, unroll, precision, watchdog_us
    p.occupancy, p.block, p.exact = occupancy, block, int(bool(exact))
    return p


def refine_opts(precision=FP32, tol=0.0, max_iters=0, fallback=False) -> RefineOpts:
    o = RefineOpts()
    o.max_iters, o.precision, o.tol, o.fallback = max_iters, precision, tol, int(bool(fallback))
    return o


def device_info():
    name = ctypes.create_string_buffer(256)
    cus, mem = _i(0), _sz(0)
    check(lib().sptrsv_device_info(name, 256, ctypes.byref(cus), ctypes.byref(mem)), "device_info")
    return name.value.decode(), cus.value, mem.value


def device_sync():
    check(lib().sptrsv_device_sync(), "device_sync")


def selector_set(sizes, params, mu, sd, classes):
    """install a variant-selector network for AUTO (24 -> h1 -> h2 -> out, sptrsv_mlp layout)"""
    p = np.ascontiguousarray(params, dtype=np.float32)
    mu = np.ascontiguousarray(mu, dtype=np.float32)
    sd = np.ascontiguousarray(sd, dtype=np.float32)
    cl = np.ascontiguousarray(classes, dtype=np.int32)
    i, h1, h2, o = (int(v) for v in sizes)
    check(lib().sptrsv_selector_set(i, h1, h2, o, _ptr(p), _ptr(mu), _ptr(sd), _ptr(cl)), "selector_set")


def selector_clear():
    check(lib().sptrsv_selector_clear(), "selector_clear")


def selector_load(path=None):
    """install the single-system selector saved by sptrsv_amd.tune_single (default: the one
    shipped with the package); returns its training summary"""
    import json
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_single.json")
    with open(path) as f:
        doc = json.load(f)
    selector_set(doc["sizes"], doc["params"], doc["mu"], doc["sd"], doc["classes"])
    return doc.get("meta", {})


def division_check(n: np.ndarray, d: np.ndarray):
    """the kernels' division with a reused reciprocal and the compiled division, on the device:
    (fast, ieee), to be bitwise equal"""
    n = np.ascontiguousarray(n, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    bufs = [DeviceArray.from_host(n), DeviceArray.from_host(d), DeviceArray(n.shape), DeviceArray(n.shape)]
    try:
        check(lib().sptrsv_division_check(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, n.size), "division_check")
        return bufs[2].download(), bufs[3].download()
    finally:
        for b in bufs:
            b.free()


def validate_csr(n, rowptr, col, val) -> int:
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    c = np.ascontiguousarray(col, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    if len(rp) != n + 1 or len(v) != len(c):
        return ERR_BAD_INPUT
    return lib().sptrsv_validate_csr(int(n), len(c), _ptr(rp), _ptr(c), _ptr(v))


def _take_host(ptr, count, dtype):
    """copy a library-allocated host array into numpy and release it"""
    if not ptr.value:
        return np.zeros(0, dtype=dtype)
    buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr.value)
    out = np.frombuffer(bytes(buf), dtype=dtype).copy()
    lib().sptrsv_free_host(ptr)
    return out


def read_matrix(path: str):
    """(n, rowptr int64, col int32, val f64) from a .matrix / .mtx / .txt file (the reference's
    formats, read by the library's loader)."""
    n, nnz = _i(0), _i64(0)
    rp, c, v = _vp(), _vp(), _vp()
    rc = lib().sptrsv_read_matrix_file(os.fsencode(path), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(rp),
                                       ctypes.byref(c), ctypes.byref(v))
    if rc != OK:
        for q in (rp, c, v):
            if q.value:
                lib().sptrsv_free_host(q)
        raise SptrsvError(rc, "read_matrix", f"cannot read {path}")
    return (n.value, _take_host(rp, n.value + 1, np.int64), _take_host(c, nnz.value, np.int32),
            _take_host(v, nnz.value, np.float64))


def read_snodes(path: str):
    """supernode boundaries (sn + 1 ints) from a .snodes / .txt file"""
    sn, p = _i(0), _vp()
    rc = lib().sptrsv_read_snodes_file(os.fsencode(path), ctypes.byref(sn), ctypes.byref(p))
    if rc != OK:
        raise SptrsvError(rc, "read_snodes", f"cannot read {path}")
    return _take_host(p, sn.value + 1, np.int32)


def validate_csc(n, colptr, rowidx, val) -> int:
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    ri = np.ascontiguousarray(rowidx, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    if len(cp) != n + 1 or len(v) != len(ri):
        return ERR_BAD_INPUT
    return lib().sptrsv_validate_csc(int(n), len(ri), _ptr(cp), _ptr(ri), _ptr(v))


def last_error() -> str:
    return lib().sptrsv_last_error().decode()


def csr_to_csc(n, rowptr, col, val):
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    c = np.ascontiguousarray(col, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    nnz = int(rp[-1])
    cp = np.zeros(n + 1, dtype=np.int32)
    ri = np.zeros(max(nnz, 1), dtype=np.int32)
    cv = np.zeros(max(nnz, 1), dtype=np.float64)
    check(lib().sptrsv_csr_to_csc(int(n), nnz, _ptr(rp), _ptr(c), _ptr(v), _ptr(cp), _ptr(ri), _ptr(cv)),
          "csr_to_csc")
    return cp, ri[:nnz], cv[:nnz]


class DeviceArray:
    """A device allocation holding a numpy-shaped array (no torch involved)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _vp()
        check(lib().sptrsv_malloc(ctypes.byref(self.ptr), self.nbytes), "malloc")

    @classmethod
    def from_host(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().sptrsv_memcpy_h2d(self.ptr, _ptr(a), self.nbytes), "h2d")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().sptrsv_memcpy_d2h(_ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib().sptrsv_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Plan:
    """One upper-triangular factor resident on the GPU (CSR + CSC + analysis)."""

    def __init__(self, handle):
        self.handle = handle
        self._owned = True
        n, nnz, L = _i(0), _i64(0), _i(0)
        check(lib().sptrsv_plan_info(self.handle, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(L)), "info")
        self.n, self.nnz, self.n_levels = n.value, nnz.value, L.value

    def analysis_time(self):
        """(analysis ms, level computation ms): host wall time of the plan's device analysis."""
        a, l = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(lib().sptrsv_plan_analysis_time(self.handle, ctypes.byref(a), ctypes.byref(l)), "analysis_time")
        return a.value, l.value

    @classmethod
    def from_csr(cls, rowptr, col, val):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        c = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        h = _PLAN()
        if len(c) != int(rp[-1]) or len(v) != len(c):
            raise SptrsvError(ERR_BAD_INPUT, "plan_create_csr", "array lengths do not match rowptr")
        check(lib().sptrsv_plan_create_csr(len(rp) - 1, len(c), _ptr(rp), _ptr(c), _ptr(v),
                                           ctypes.byref(h)), "plan_create_csr")
        return cls(h)

    @classmethod
    def from_upper(cls, U):
        return cls.from_csr(U.rowptr, U.col, U.val)

    @classmethod
    def from_csc(cls, colptr, rowidx, val):
        cp = np.ascontiguousarray(colptr, dtype=np.int32)
        ri = np.ascontiguousarray(rowidx, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=np.float64)
        h = _PLAN()
        check(lib().sptrsv_plan_create_csc(len(cp) - 1, int(cp[-1]), _ptr(cp), _ptr(ri), _ptr(v),
                                           ctypes.byref(h)), "plan_create_csc")
        return cls(h)

    def close(self):
        if self.handle and self._owned:
            lib().sptrsv_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- analysis ----
    def indegree(self):
        out = np.zeros(self.n, dtype=np.int32)
        check(lib().sptrsv_plan_get_indegree(self.handle, _ptr(out)), "get_indegree")
        return out

    def levels(self):
        out = np.zeros(self.n, dtype=np.int32)
        check(lib().sptrsv_plan_get_levels(self.handle, _ptr(out)), "get_levels")
        return out

    def auto_variant(self, m=1, precision=FP64):
        """the variant SPTRSV_VAR_AUTO runs for m right-hand sides at `precision`"""
        out = np.zeros(1, dtype=np.int32)
        check(lib().sptrsv_plan_auto_variant(self.handle, m, precision, _ptr(out)), "auto_variant")
        return int(out[0])

    def features(self, m=1):
        """the 24 selector features of the factor at m right-hand sides"""
        out = np.zeros(24, dtype=np.float32)
        check(lib().sptrsv_plan_features(self.handle, m, _ptr(out)), "plan_features")
        return out

    def blocks(self):
        """block order of the BLOCK variant: (kind, block_ptr, perm); kind 1 = grid bricks,
        2 = level bands"""
        nb = np.zeros(1, dtype=np.int32)
        kind = np.zeros(1, dtype=np.int32)
        check(lib().sptrsv_plan_get_blocks(self.handle, _ptr(nb), _ptr(kind), None, None), "get_blocks")
        ptr = np.zeros(int(nb[0]) + 1, dtype=np.int32)
        perm = np.zeros(self.n, dtype=np.int32)
        check(lib().sptrsv_plan_get_blocks(self.handle, _ptr(nb), _ptr(kind), _ptr(ptr), _ptr(perm)), "get_blocks")
        return int(kind[0]), ptr, perm

    def level_sets(self):
        ptr = np.zeros(self.n_levels + 1, dtype=np.int32)
        rows = np.zeros(self.n, dtype=np.int32)
        check(lib().sptrsv_plan_get_level_sets(self.handle, _ptr(ptr), _ptr(rows)), "get_level_sets")
        return ptr, rows

    def build_supernodes(self, snodes=None):
        """Supernodal layout from boundaries `snodes` (sn+1 ints), or detected when None."""
        if snodes is None:
            check(lib().sptrsv_plan_build_supernodes(self.handle, 0, None), "build_supernodes")
        else:
            s = np.ascontiguousarray(snodes, dtype=np.int32)
            check(lib().sptrsv_plan_build_supernodes(self.handle, len(s) - 1, _ptr(s)), "build_supernodes")
        return self.supernodes()

    def build_supernodes_relaxed(self, max_width=0, max_fill=0.0):
        """Relaxed supernodes (explicit zeros allowed); returns (boundaries, explicit zeros)."""
        check(lib().sptrsv_plan_build_supernodes_relaxed(self.handle, int(max_width), float(max_fill)),
              "build_supernodes_relaxed")
        z = _i64(0)
        check(lib().sptrsv_plan_supernodal_padding(self.handle, ctypes.byref(z)), "supernodal_padding")
        return self.supernodes(), z.value

    def supernodes(self):
        sn = _i(0)
        check(lib().sptrsv_plan_get_supernodes(self.handle, ctypes.byref(sn), None), "get_supernodes")
        out = np.zeros(sn.value + 1, dtype=np.int32)
        check(lib().sptrsv_plan_get_supernodes(self.handle, ctypes.byref(sn), _ptr(out)), "get_supernodes")
        return out

    def csc(self):
        cp = np.zeros(self.n + 1, dtype=np.int32)
        ri = np.zeros(self.nnz, dtype=np.int32)
        cv = np.zeros(self.nnz, dtype=np.float64)
        check(lib().sptrsv_plan_get_csc(self.handle, _ptr(cp), _ptr(ri), _ptr(cv)), "get_csc")
        return cp, ri, cv

    # ---- solves ----
    def solve(self, d_b: DeviceArray, d_x: DeviceArray, m: int = 1, prm: Params | None = None):
        """Asynchronous device solve (row-major n x m)."""
        check(lib().sptrsv_solve(self.handle, ctypes.byref(prm or params()), d_b.ptr, d_x.ptr, int(m)), "solve")

    def sync(self):
        check(lib().sptrsv_plan_sync(self.handle), "sync")

    def solve_push_noalloc(self, d_b: DeviceArray, d_x: DeviceArray, m: int = 1):
        """Asynchronous CSC push solve with the library's fixed work array (no allocation)."""
        check(lib().sptrsv_solve_push_noalloc(self.handle, d_b.ptr, d_x.ptr, int(m)), "solve_push_noalloc")

    def solve_host(self, b: np.ndarray, prm: Params | None = None, col_major: bool = False):
        """Synchronous solve of host b: shape (n,), (n, m) row-major, or (m, n) with col_major."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        if b.ndim == 1:
            m = 1
        else:
            m = b.shape[0] if col_major else b.shape[1]
        x = np.empty_like(b)
        check(lib().sptrsv_solve_host(self.handle, ctypes.byref(prm or params()), _ptr(b), _ptr(x), int(m),
                                      int(bool(col_major))), "solve_host")
        return x

    def solve_refined(self, d_b, d_x, m=1, prm=None, precision=FP32, tol=0.0, max_iters=0):
        """Reduced-precision solves + FP64 refinement on the device (synchronous).
        Returns (corrections applied, estimated relative error of x; converged when <= tol)."""
        o = refine_opts(precision, tol, max_iters)
        it, ratio = _i(0), ctypes.c_double(0.0)
        check(lib().sptrsv_solve_refined(self.handle, ctypes.byref(prm or params()), ctypes.byref(o), d_b.ptr,
                                         d_x.ptr, int(m), ctypes.byref(it), ctypes.byref(ratio)), "solve_refined")
        return it.value, ratio.value

    def solve_refined_ex(self, d_b, d_x, m=1, prm=None, precision=FP32, tol=0.0, max_iters=0, fallback=False):
        """As solve_refined, returning the outcome report (dict: iters, outcome, fell_back,
        variant, estimate); with fallback a run that does not converge ends with the FP64 solve."""
        o = refine_opts(precision, tol, max_iters, fallback)
        info = RefineInfo()
        check(lib().sptrsv_solve_refined_ex(self.handle, ctypes.byref(prm or params()), ctypes.byref(o), d_b.ptr,
                                            d_x.ptr, int(m), ctypes.byref(info)), "solve_refined_ex")
        return info.as_dict()

    def condest(self):
        """(kappa_1(U) estimate, estimate of ||U^-1||_1) from FP64 solves on the device."""
        k, ni = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(lib().sptrsv_condest(self.handle, ctypes.byref(k), ctypes.byref(ni)), "condest")
        return k.value, ni.value

    def solve_transposed(self, d_b: DeviceArray, d_x: DeviceArray, m: int = 1, prm: Params | None = None):
        """Asynchronous device solve of U^T X = B (forward substitution, row-major n x m)."""
        check(lib().sptrsv_solve_transposed(self.handle, ctypes.byref(prm or params()), d_b.ptr, d_x.ptr, int(m)),
              "solve_transposed")

    def solve_cholesky(self, d_b: DeviceArray, d_x: DeviceArray, m: int = 1, prm: Params | None = None):
        """Asynchronous device solve of U^T U X = B (row-major n x m)."""
        check(lib().sptrsv_solve_cholesky(self.handle, ctypes.byref(prm or params()), d_b.ptr, d_x.ptr, int(m)),
              "solve_cholesky")

    def _solve_dev_host(self, method, b, prm):
        b = np.ascontiguousarray(b, dtype=np.float64)
        m = 1 if b.ndim == 1 else b.shape[1]
        d_b = DeviceArray.from_host(b)
        d_x = DeviceArray(b.shape)
        try:
            method(d_b, d_x, m, prm)
            self.sync()
            return d_x.download()
        finally:
            d_b.free()
            d_x.free()

    def solve_transposed_host(self, b: np.ndarray, prm: Params | None = None):
        """Synchronous U^T x = b for host b of shape (n,) or row-major (n, m)."""
        return self._solve_dev_host(self.solve_transposed, b, prm)

    def solve_cholesky_host(self, b: np.ndarray, prm: Params | None = None):
        """Synchronous U^T U x = b for host b of shape (n,) or row-major (n, m)."""
        return self._solve_dev_host(self.solve_cholesky, b, prm)

    def ccs2ccs_pad(self, snodes):
        """The padded supernodal layout built on the device: (val_pad, col_pad uint64, rowptr_pad)."""
        s = np.ascontiguousarray(snodes, dtype=np.int32)
        total = _i64(0)
        check(lib().sptrsv_ccs2ccs_pad_device(self.handle, len(s) - 1, _ptr(s), None, None, None, ctypes.byref(total)),
              "ccs2ccs_pad")
        dv, dc = DeviceArray((total.value,)), DeviceArray((total.value,), np.uint64)
        dr = DeviceArray((self.n + 1,), np.int32)
        try:
            check(lib().sptrsv_ccs2ccs_pad_device(self.handle, len(s) - 1, _ptr(s), dv.ptr, dc.ptr, dr.ptr,
                                                  ctypes.byref(total)), "ccs2ccs_pad")
            return dv.download(), dc.download(), dr.download()
        finally:
            for d in (dv, dc, dr):
                d.free()

    def residual(self, d_b: DeviceArray, d_x: DeviceArray, d_r: DeviceArray, m: int = 1):
        check(lib().sptrsv_residual(self.handle, d_b.ptr, d_x.ptr, d_r.ptr, int(m)), "residual")

    def timing(self, enable: bool):
        check(lib().sptrsv_plan_timing(self.handle, int(enable)), "timing")

    def timing_read(self):
        ms, cnt = ctypes.c_double(0.0), _i(0)
        check(lib().sptrsv_plan_timing_read(self.handle, ctypes.byref(ms), ctypes.byref(cnt)), "timing_read")
        return ms.value, cnt.value


BCFG_LS1, BCFG_LS4, BCFG_LS16, BCFG_LS64, BCFG_CHAIN = range(5)
BCFG_COUNT = 5
BCFG_NAMES = ["ls1", "ls4", "ls16", "ls64", "chain"]


class Batch:
    """Independent systems solved together (one persistent launch); see sptrsv_batch_*."""

    def __init__(self, U, sys_ptr):
        rp = np.ascontiguousarray(U.rowptr, dtype=np.int64)
        c = np.ascontiguousarray(U.col, dtype=np.int32)
        v = np.ascontiguousarray(U.val, dtype=np.float64)
        sp = np.ascontiguousarray(sys_ptr, dtype=np.int32)
        h = _vp()
        check(lib().sptrsv_batch_create(len(sp) - 1, _ptr(sp), len(c), _ptr(rp), _ptr(c), _ptr(v), ctypes.byref(h)),
              "batch_create")
        self.handle = h
        self.nsys = len(sp) - 1
        self.n = U.n
        self.nnz = len(c)

    def close(self):
        if self.handle:
            lib().sptrsv_batch_destroy(self.handle)
            self.handle = None

    def residual(self, d_b, d_x, d_r, m=1):
        """r = b - U x over every system (FP64, asynchronous; sync() waits)."""
        check(lib().sptrsv_batch_residual(self.handle, d_b.ptr, d_x.ptr, d_r.ptr, int(m)), "batch_residual")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def features(self, m=1):
        out = np.zeros((self.nsys, 24), dtype=np.float32)
        check(lib().sptrsv_batch_features(self.handle, int(m), _ptr(out)), "batch_features")
        return out

    def set_configs(self, cfg=None, order=None):
        c = None if cfg is None else np.ascontiguousarray(cfg, dtype=np.int32)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        check(lib().sptrsv_batch_set_configs(self.handle, None if c is None else _ptr(c),
                                             None if o is None else _ptr(o)), "batch_set_configs")

    def solve(self, d_b, d_x, m=1, precision=FP64, graph=False):
        check(lib().sptrsv_batch_solve(self.handle, int(precision), d_b.ptr, d_x.ptr, int(m), int(bool(graph))),
              "batch_solve")

    def sync(self):
        check(lib().sptrsv_batch_sync(self.handle), "batch_sync")

    def solve_refined(self, d_b, d_x, m=1, tol=0.0, max_iters=0):
        """FP32 batched solves + FP64 refinement (synchronous): (corrections, error estimate)."""
        o = refine_opts(FP32, tol, max_iters)
        it, est = _i(0), ctypes.c_double(0.0)
        check(lib().sptrsv_batch_solve_refined(self.handle, ctypes.byref(o), d_b.ptr, d_x.ptr, int(m), ctypes.byref(it),
                                               ctypes.byref(est)), "batch_solve_refined")
        return it.value, est.value

    def solve_refined_ex(self, d_b, d_x, m=1, tol=0.0, max_iters=0, fallback=False):
        """As solve_refined with the outcome report (see Plan.solve_refined_ex)."""
        o = refine_opts(FP32, tol, max_iters, fallback)
        info = RefineInfo()
        check(lib().sptrsv_batch_solve_refined_ex(self.handle, ctypes.byref(o), d_b.ptr, d_x.ptr, int(m),
                                                  ctypes.byref(info)), "batch_solve_refined_ex")
        return info.as_dict()

    def time_config(self, cfg, d_b, d_x, m=1, precision=FP64):
        us = np.zeros(self.nsys, dtype=np.float64)
        check(lib().sptrsv_batch_time_config(self.handle, int(cfg), int(precision), d_b.ptr, d_x.ptr, int(m), _ptr(us)),
              "batch_time_config")
        return us


class MLP:
    """The selector network on the GPU (sptrsv_mlp_*)."""

    def __init__(self, nin=24, h1=64, h2=64, nout=5, seed=0):
        h = _vp()
        check(lib().sptrsv_mlp_create(nin, h1, h2, nout, seed, ctypes.byref(h)), "mlp_create")
        self.handle = h
        self.sizes = (nin, h1, h2, nout)
        cnt = _i(0)
        check(lib().sptrsv_mlp_num_params(h, ctypes.byref(cnt)), "mlp_num_params")
        self.nparams = cnt.value

    def close(self):
        if self.handle:
            lib().sptrsv_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        check(lib().sptrsv_mlp_get_params(self.handle, _ptr(p)), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().sptrsv_mlp_set_params(self.handle, _ptr(p)), "mlp_set_params")

    def loss_grad(self, X, y, mask=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        losses = np.zeros(len(y), dtype=np.float32)
        grad = np.zeros(self.nparams, dtype=np.float32)
        check(lib().sptrsv_mlp_loss_grad(self.handle, _ptr(X), _ptr(y), None if mk is None else _ptr(mk), len(y),
                                         _ptr(losses), _ptr(grad)), "mlp_loss_grad")
        return losses, grad

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((len(X), self.sizes[3]), dtype=np.float32)
        check(lib().sptrsv_mlp_forward(self.handle, _ptr(X), len(X), _ptr(out)), "mlp_forward")
        return out

    def train(self, X, y, mask=None, epochs=500, lr=1e-2):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        fl = ctypes.c_float(0.0)
        check(lib().sptrsv_mlp_train(self.handle, _ptr(X), _ptr(y), None if mk is None else _ptr(mk), len(y),
                                     int(epochs), float(lr), ctypes.byref(fl)), "mlp_train")
        return fl.value
