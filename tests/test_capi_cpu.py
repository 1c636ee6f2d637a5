"""CPU-side checks of the shipped library: it loads, exports every entry point the headers in
include/ declare, validates input before touching a device, and the reference's loader entry
points (include/utils.h) read the reference's fixtures from C++.  No compute calls."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from sptrsv_amd import capi, gen

INCLUDE = os.path.join(ROOT, "include")


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?[\s\*]([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return out


def test_library_loads():
    assert capi.lib() is not None


@pytest.mark.parametrize("header", ["sptrsv_capi.h", "core.h", "utils.h", "sptrsv_syncfree_opencl.h"])
def test_every_declared_symbol_is_exported(header):
    names = declared(header)
    assert names, header
    table = exported()
    for name in names:
        if header == "sptrsv_capi.h":
            assert re.search(rf" T {name}$", table, flags=re.M), name
        else:  # C++ linkage, demangled "name(args)"
            assert re.search(rf" T {name}\(", table, flags=re.M), name


def test_capi_header_names_match_the_reference_drop_in():
    # the drop-in headers keep the reference's entry-point names (include/core.h:9-19,
    # include/utils.h:14-20, include/sptrsv_syncfree_opencl.h:11-37)
    assert set(declared("core.h")) == {
        "base_gauss_lower", "base_gauss_upper", "base_gauss_upper_new", "supernodal_upper",
        "supernodal_upper_new", "supernodal_blas_upper", "supernodal_blas_upper_new", "ccs2ccs_pad",
        "setLevelUpAndGetMaxLevel", "barrier_upper", "barrier_upper_new"}
    assert set(declared("sptrsv_syncfree_opencl.h")) == {
        "sptrsv_syncfree_opencl", "sptrsv_syncfree2_opencl", "sptrsv_syncfree3_opencl"}
    assert {"read_csr", "read_snodes", "fill_x_b", "transpose", "check_result"} <= set(declared("utils.h"))


def test_validation_accepts_fixtures(csr5, csr6):
    for n, rp, col, val in (csr5, csr6):
        assert capi.validate_csr(n, rp, col, val) == capi.OK
    U = gen.random_zipf(2000, seed=1)
    assert capi.validate_csr(U.n, U.rowptr, U.col, U.val) == capi.OK
    cp, ri, cv = gen.to_csc(U)
    assert capi.validate_csc(U.n, cp, ri, cv) == capi.OK


@pytest.mark.parametrize("case,fragment", [
    ("lower_entry", "first entry is not the diagonal"),
    ("unsorted", "not strictly ascending"),
    ("zero_diag", "zero or non-finite diagonal"),
    ("nan_diag", "zero or non-finite diagonal"),
    ("empty_row", "is empty"),
    ("bad_rowptr", "rowptr[n] != nnz"),
    ("col_range", "not strictly ascending"),
])
def test_validation_rejects_malformed(case, fragment):
    # a lower entry would be a dependency cycle for the sync-free kernels (the reference hangs
    # on it); every such input must be refused before any launch
    rp = np.array([0, 3, 5, 6])
    col = np.array([0, 1, 2, 1, 2, 2])
    val = np.array([2.0, 1, 1, 3, 1, 4])
    if case == "lower_entry":
        col = np.array([0, 1, 2, 0, 1, 2])
    elif case == "unsorted":
        col = np.array([0, 2, 1, 1, 2, 2])
    elif case == "zero_diag":
        val = val.copy(); val[3] = 0.0
    elif case == "nan_diag":
        val = val.copy(); val[5] = np.nan
    elif case == "empty_row":
        rp = np.array([0, 3, 3, 6]); col = np.array([0, 1, 2, 1, 2, 2])
    elif case == "bad_rowptr":
        rp = np.array([0, 3, 5, 7])
    elif case == "col_range":
        col = np.array([0, 1, 3, 1, 2, 2])
    assert capi.validate_csr(3, rp, col, val) == capi.ERR_BAD_INPUT
    assert fragment in capi.last_error()


def test_csc_validation_rejects_missing_diagonal():
    cp = np.array([0, 1, 2, 4], dtype=np.int32)
    ri = np.array([0, 0, 0, 1], dtype=np.int32)  # column 1 has no diagonal
    assert capi.validate_csc(3, cp, ri, np.ones(4)) == capi.ERR_BAD_INPUT


def _build(src, out):
    lib_dir = os.path.dirname(capi.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", INCLUDE, os.path.join(ROOT, "tests", "cpp", src), "-o", out,
                    "-L", lib_dir, "-lsptrsv_amd", f"-Wl,-rpath,{lib_dir}"], check=True)
    return out


def test_cpp_loaders_read_reference_fixtures(tmp_path):
    exe = _build("dropin_io.cpp", str(tmp_path / "dropin_io"))
    out = subprocess.run([exe, os.path.join(GOLDEN, "csr_6_upper.txt"), os.path.join(GOLDEN, "csr_6_upper_snodes.txt")],
                         capture_output=True, text=True, check=True)
    lines = out.stdout.splitlines()
    assert lines[0] == "read ok"
    assert lines[1] == "n 6 nz 13"
    assert lines[2] == "row 0 2 6 9 10 12 13"
    assert lines[3] == "col 0 3 1 2 4 5 2 4 5 3 4 5 5"
    assert lines[4].split()[1:] == ["1", "1", "1", "2", "-1", "1", "1", "2", "3", "2", "4", "1", "6"]
    assert lines[5] == "sn 4 snodes 0 1 3 4 6"
    assert lines[6] == "fill 24.0"
    assert "Relative error: 2.887e-04" in out.stderr


def test_cpp_loaders_binary_and_mtx(tmp_path):
    """.matrix / .snodes binary layouts (src/utils.cpp:17-31, :89-98) and .mtx coordinates sorted
    by column (a lower factor, src/utils.cpp:32-60), written here from the csr_5 fixture."""
    exe = _build("dropin_io.cpp", str(tmp_path / "dropin_io"))
    n, rp, col, val = 5, [0, 2, 3, 4, 5, 6], [0, 3, 1, 2, 3, 4], [1, 2, 3, 4, 5, 6.0]
    mpath, spath = tmp_path / "m.matrix", tmp_path / "m.snodes"
    with open(mpath, "wb") as f:
        f.write(struct.pack("<iQ", n, len(col)))
        f.write(np.asarray(col, dtype=np.int32).tobytes())
        f.write(np.asarray(rp, dtype=np.uint64).tobytes())
        f.write(np.asarray(val, dtype=np.float64).tobytes())
    with open(spath, "wb") as f:
        f.write(struct.pack("<i", 2))
        f.write(np.asarray([0, 3, 5], dtype=np.int32).tobytes())
    out = subprocess.run([exe, str(mpath), str(spath)], capture_output=True, text=True, check=True).stdout
    assert "n 5 nz 6" in out and "row 0 2 3 4 5 6" in out and "col 0 3 1 2 3 4" in out
    assert "sn 2 snodes 0 3 5" in out
    # Matrix Market of L = U^T: entries (row, col) 1-based sorted by column
    mtx = tmp_path / "m.mtx"
    entries = []
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            entries.append((col[p] + 1, i + 1, val[p]))  # L(col, i) = U(i, col)
    entries.sort(key=lambda e: (e[1], e[0]))
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% lower factor\n")
        f.write(f"{n} {n} {len(entries)}\n")
        for r, c, v in entries:
            f.write(f"{r} {c} {v}\n")
    out = subprocess.run([exe, str(mtx), str(spath)], capture_output=True, text=True, check=True).stdout
    assert "n 5 nz 6" in out and "row 0 2 3 4 5 6" in out and "col 0 3 1 2 3 4" in out
    bad = subprocess.run([exe, str(tmp_path / "m.xyz"), str(spath)], capture_output=True, text=True)
    assert "read failed" in bad.stdout and "Unknown extension" in bad.stderr
    missing = subprocess.run([exe, str(tmp_path / "none.matrix"), str(spath)], capture_output=True, text=True)
    assert "read failed" in missing.stdout


def _write_matrix(path, U):
    """.matrix binary layout (src/utils.cpp:17-31): int n, uint64 nz, int32 col, uint64 rowptr, f64 val"""
    with open(path, "wb") as f:
        f.write(struct.pack("<iQ", U.n, U.nnz))
        f.write(np.asarray(U.col, dtype=np.int32).tobytes())
        f.write(np.asarray(U.rowptr, dtype=np.uint64).tobytes())
        f.write(np.asarray(U.val, dtype=np.float64).tobytes())


def test_c_abi_loaders(tmp_path):
    """sptrsv_read_matrix_file / sptrsv_read_snodes_file give the reference fixtures' arrays and a
    .matrix / .snodes pair written here back bit for bit; a missing file is an error."""
    n, rp, col, val = capi.read_matrix(os.path.join(GOLDEN, "csr_6_upper.txt"))
    assert n == 6 and rp.tolist() == [0, 2, 6, 9, 10, 12, 13]
    assert col.tolist() == [0, 3, 1, 2, 4, 5, 2, 4, 5, 3, 4, 5, 5]
    assert val.tolist() == [1, 1, 1, 2, -1, 1, 1, 2, 3, 2, 4, 1, 6]
    assert capi.read_snodes(os.path.join(GOLDEN, "csr_6_upper_snodes.txt")).tolist() == [0, 1, 3, 4, 6]
    U = gen.random_supernodal(6, seed=3)
    _write_matrix(tmp_path / "s.matrix", U)
    with open(tmp_path / "s.snodes", "wb") as f:
        f.write(struct.pack("<i", len(U.snodes) - 1))
        f.write(np.asarray(U.snodes, dtype=np.int32).tobytes())
    n, rp, col, val = capi.read_matrix(str(tmp_path / "s.matrix"))
    assert n == U.n and np.array_equal(rp, U.rowptr) and np.array_equal(col, U.col) and np.array_equal(val, U.val)
    assert np.array_equal(capi.read_snodes(str(tmp_path / "s.snodes")), U.snodes)
    # file-level failures are SPTRSV_ERR_IO, malformed contents SPTRSV_ERR_BAD_INPUT
    for bad in ("missing.matrix", "missing.txt", "missing.mtx"):
        with pytest.raises(capi.SptrsvError) as e:
            capi.read_matrix(str(tmp_path / bad))
        assert e.value.code == capi.ERR_IO, bad
    (tmp_path / "m.unknown").write_text("1 1\n1.0\n0\n0 1\n")
    with pytest.raises(capi.SptrsvError) as e:
        capi.read_matrix(str(tmp_path / "m.unknown"))
    assert e.value.code == capi.ERR_IO
    with open(tmp_path / "short.matrix", "wb") as f:
        f.write(struct.pack("<iQ", 4, 10))  # header only: a short read
    with pytest.raises(capi.SptrsvError) as e:
        capi.read_matrix(str(tmp_path / "short.matrix"))
    assert e.value.code == capi.ERR_IO
    (tmp_path / "neg.txt").write_text("-3 1\n")
    with pytest.raises(capi.SptrsvError) as e:
        capi.read_matrix(str(tmp_path / "neg.txt"))
    assert e.value.code == capi.ERR_BAD_INPUT
    with pytest.raises(capi.SptrsvError) as e:
        capi.read_snodes(str(tmp_path / "missing.snodes"))
    assert e.value.code == capi.ERR_IO


def test_tables_layout(tmp_path, monkeypatch):
    """times.csv keeps the reference driver's layout (run.py:34-58): ';' after every cell, a header
    row of algorithms, one row per matrix file (supernode files skipped), min over the repeats."""
    from sptrsv_amd import tables
    for name in ("b.matrix", "a.mtx", "a.snodes", "c_snodes.txt", "c.txt", "notes.md"):
        (tmp_path / name).write_text("")
    times = iter([3.0, 1.0, 2.0, 5.0, 4.0, 6.0, 0.5, 0.7, 0.9, 8.0, 7.0, 9.0, 2.0, 2.0, 2.0, 1.0, 3.0, 1.0])
    calls = []

    def fake(algo, matrix, snodes, rhs):
        calls.append((algo, os.path.basename(matrix), os.path.basename(snodes)))
        return f"log {algo}\n", next(times)

    monkeypatch.setattr(tables, "single_launch", fake)
    monkeypatch.setattr("sys.argv", ["tables"])
    tables.main([str(tmp_path), "--algos", "base,barrier", "--repeats", "3", "--out", str(tmp_path / "exp")])
    (run,) = os.listdir(tmp_path / "exp")
    rows = open(tmp_path / "exp" / run / "times.csv").read().splitlines()
    assert rows[0] == ";base;barrier;"
    assert rows[1:] == ["a.mtx;1.0;4.0;", "b.matrix;0.5;7.0;", "c.txt;2.0;1.0;"]
    assert ("base", "a.mtx", "a.snodes") in calls and ("barrier", "c.txt", "c_snodes.txt") in calls
    assert os.path.exists(tmp_path / "exp" / run / "best_times_log.log")


PANEL_EMULATE_CASES = [
    "16 16 16 16 1 0", "12 12 12 32 2 0.1", "10 9 11 20 3 0.2", "9 6 5 3 6 0.3", "8 8 8 128 7 0",
    "2 2 2 1 8 0", "3 40 70 17 10 0.05", "17 13 1 16 4 0.1", "33 40 1 5 5 0", "40 3 1 16 9 0",
]


def test_panel_kernel_indexing_emulated(tmp_path):
    """The PANEL kernel's address arithmetic on the layout tables, emulated on the CPU under
    AddressSanitizer and UBSan (tests/cpp/panel_emulate.cpp): every dependency read comes from
    the cell or halo value written before it, no access leaves its buffer, and x is bitwise
    equal to the serial back substitution -- partial bricks, 2D and 3D grids, missing entries,
    m not a multiple of the 16-column panel."""
    exe = str(tmp_path / "panel_emulate")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(ROOT, "sptrsv_amd", "csrc"), os.path.join(ROOT, "tests", "cpp", "panel_emulate.cpp"),
                    "-o", exe], check=True)
    for case in PANEL_EMULATE_CASES:
        out = subprocess.run([exe, *case.split()], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.startswith("ok"), (case, out.stdout, out.stderr)
