import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def read_txt_csr(path):
    """The reference's .txt CSR layout: n nz / val[nz] / col[nz] / rowptr[n+1]
    (src/utils.cpp:61-76)."""
    tok = open(path).read().split()
    n, nz = int(tok[0]), int(tok[1])
    vals = np.array([float(t) for t in tok[2:2 + nz]])
    cols = np.array([int(t) for t in tok[2 + nz:2 + 2 * nz]], dtype=np.int32)
    rp = np.array([int(t) for t in tok[2 + 2 * nz:3 + 2 * nz + n]], dtype=np.int64)
    return n, rp, cols, vals


def read_txt_snodes(path):
    tok = open(path).read().split()
    sn = int(tok[0])
    return np.array([int(t) for t in tok[1:2 + sn]], dtype=np.int32)


@pytest.fixture(scope="session")
def csr5():
    return read_txt_csr(os.path.join(GOLDEN, "csr_5_upper.txt"))


@pytest.fixture(scope="session")
def csr6():
    return read_txt_csr(os.path.join(GOLDEN, "csr_6_upper.txt"))


@pytest.fixture(scope="session")
def snodes6():
    return read_txt_snodes(os.path.join(GOLDEN, "csr_6_upper_snodes.txt"))
