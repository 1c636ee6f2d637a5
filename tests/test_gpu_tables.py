"""GPU: the time-tables driver (sptrsv_amd/tables.py, the reference's run.py layout) end to end."""
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from sptrsv_amd import gen

pytestmark = pytest.mark.gpu


def test_tables_end_to_end(tmp_path):
    """Two factors (the csr_6 fixture with its .txt supernodes, and a supernodal .matrix/.snodes
    pair): every algorithm of the reference CLI gives a positive time, one fresh process each,
    and every process's residual ||b - U x|| / ||b|| is at rounding level."""
    d = tmp_path / "mats"
    d.mkdir()
    shutil.copy(os.path.join(GOLDEN, "csr_6_upper.txt"), d / "csr_6_upper.txt")
    shutil.copy(os.path.join(GOLDEN, "csr_6_upper_snodes.txt"), d / "csr_6_upper_snodes.txt")
    U = gen.random_supernodal(12, seed=5)
    with open(d / "sn12.matrix", "wb") as f:
        f.write(struct.pack("<iQ", U.n, U.nnz))
        f.write(np.asarray(U.col, dtype=np.int32).tobytes())
        f.write(np.asarray(U.rowptr, dtype=np.uint64).tobytes())
        f.write(np.asarray(U.val, dtype=np.float64).tobytes())
    with open(d / "sn12.snodes", "wb") as f:
        f.write(struct.pack("<i", len(U.snodes) - 1))
        f.write(np.asarray(U.snodes, dtype=np.int32).tobytes())
    print("[gpu] tables driver on 2 matrices x 6 algorithms", flush=True)
    out = subprocess.run([sys.executable, "-m", "sptrsv_amd.tables", str(d), "--repeats", "1", "--out",
                          str(tmp_path / "exp")], cwd=ROOT, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    (run,) = os.listdir(tmp_path / "exp")
    rows = open(tmp_path / "exp" / run / "times.csv").read().splitlines()
    assert rows[0] == ";base;custom;blas;barrier;syncfree;write_first;"
    assert [r.split(";")[0] for r in rows[1:]] == ["csr_6_upper.txt", "sn12.matrix"]
    for r in rows[1:]:
        cells = r.split(";")[1:-1]
        assert len(cells) == 6 and all(float(c) > 0 for c in cells), r
    log = open(tmp_path / "exp" / run / "best_times_log.log").read()
    res = [float(ln.split(": ")[1]) for ln in log.splitlines() if ln.startswith("Relative residual")]
    assert len(res) == 12 and max(res) <= 1e-12, res
