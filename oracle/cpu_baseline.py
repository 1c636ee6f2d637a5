"""CPU baselines of the reference's CPU methods on the benchmark factors -- test infrastructure
(the timed CPU baseline of bench.py; the product never runs it).

    python -m oracle.cpu_baseline <dir> <workload> [<workload> ...]

<dir> holds <workload>.npz files (rowptr, col, val[, snodes]) written by bench.py.  One run is
one fresh process, like the reference's run.py (run.py:24-32 starts the CLI once per repetition
and keeps the minimum); it prints one JSON object: per workload, the best times in ms of
  serial      base_gauss_upper (src/core.cpp:18-32), one host core, m = 1 (m = 8 for supernodal)
  levelset    barrier_upper (src/core.cpp:273-376), all host threads, analysis and solve
  supernodal  supernodal_upper_new (src/core.cpp:94-133), one host core (supernodal factors)
  blas        supernodal_blas_upper (src/core.cpp:164-201) with its cblas_dgemm / cblas_dtrsm
              calls made through MKL or another CBLAS found on the host (supernodal factors)
each timed around the solve only (omp_get_wtime, as the reference does), repeated within a
small time budget.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from oracle import oracle as O


def _best(fn, budget_s, max_reps):
    times = []
    t0 = time.perf_counter()
    while (time.perf_counter() - t0 < budget_s or not times) and len(times) < max_reps:
        times.append(fn())
    return min(times) * 1e3, len(times)


def run(path, name):
    d = np.load(path, allow_pickle=False)
    rp, col, val = d["rowptr"], d["col"], d["val"]
    n = len(rp) - 1
    sn = d["snodes"] if "snodes" in d.files else None
    m = 8 if sn is not None else 1
    rng = np.random.default_rng(1)
    b = rng.uniform(-1.0, 1.0, (m, n)) if m > 1 else rng.uniform(-1.0, 1.0, n)
    out = {"n": int(n), "nnz": int(rp[-1]), "m": m}
    out["serial_ms"], out["serial_reps"] = _best(lambda: O.base_gauss_upper(rp, col, val, b)[1], 2.0, 50)
    nth = O.max_threads()

    def ls():
        _, ta, ts = O.barrier_upper(rp, col, val, b, nth)
        return ta + ts

    out["levelset_ms"], out["levelset_reps"] = _best(ls, 1.0, 5)
    out["levelset_threads"] = nth
    if sn is not None:
        out["supernodal_ms"], out["supernodal_reps"] = _best(
            lambda: O.supernodal_upper(rp, col, val, sn, b, new=True)[1], 1.0, 5)
        lib = O.find_cblas()
        vp, cp, rpp = O.ccs2ccs_pad(rp, col, val, sn)
        if lib:
            out["blas_ms"], out["blas_reps"] = _best(
                lambda: O.supernodal_blas_upper_cblas(lib, vp, cp, rpp, sn, b)[1], 1.0, 5)
            out["blas_library"] = lib
            out["blas_threading"] = os.environ.get("MKL_THREADING_LAYER", "")
        else:  # no CBLAS on the host: the plain-loop restatement of the same method
            out["blas_ms"], out["blas_reps"] = _best(lambda: O.supernodal_blas_upper(vp, cp, rpp, sn, b)[1], 1.0, 5)
            out["blas_library"] = "none (plain loops)"
    return out


def main(argv):
    folder, names = argv[0], argv[1:]
    res = {name: run(os.path.join(folder, name + ".npz"), name) for name in names}
    print(json.dumps(res))


if __name__ == "__main__":
    main(sys.argv[1:])
