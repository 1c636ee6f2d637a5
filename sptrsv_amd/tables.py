"""Time tables per algorithm x matrix, in the layout of the reference's benchmark driver
(run.py:34-58): for each matrix file of a directory (supernode files beside it, `<name>.snodes`)
and each algorithm, the minimum over `repeats` fresh processes of the time the solve reports,
written to experiments/<timestamp>/times.csv (';'-separated, a header row of algorithms, one row
per matrix) with every process's log in best_times_log.log.

The algorithm names are the reference CLI's (app/main.cpp:6-7); here they select the kernel the
drop-in entry point of that name runs:

    base         base_gauss_upper          ROW_PULL (exact serial summation order)
    custom       supernodal_upper          SUPERNODAL (supernodes from the .snodes file)
    blas         supernodal_blas_upper     SUPERNODAL (the padded layout solves on the same kernel)
    barrier      barrier_upper             LEVELSET (one launch per level)
    syncfree     sptrsv_syncfree_opencl    AUTO (BLOCK on natural-order grid factors)
    write_first  sptrsv_syncfree3_opencl   ROW_PULL
    block / level_pull / tile / chain / csc_push    the library's other kernels

A process prints "Time: <seconds>" last on stderr, like the reference CLI (src/utils.cpp:353),
where the time is the device solve time of one solve with data resident on the GPU (the
reference's GPU functions report their kernel time the same way).  Unlike the reference's
barrier_upper, the level analysis is not inside it: a plan analyses once, at creation.  With
`check`, the process also prints the relative residual ||b - U x|| / ||b|| computed on the GPU.

    python -m sptrsv_amd.tables <matrix-dir> [--algos base,barrier,...] [--repeats 3] [--rhs 1]
"""
import argparse
import datetime
import glob
import os
import subprocess
import sys

import numpy as np

ALGOS = ["base", "custom", "blas", "barrier", "syncfree", "write_first"]


def _variant(capi, algo):
    table = {"base": capi.VAR_ROW_PULL, "custom": capi.VAR_SUPERNODAL, "blas": capi.VAR_SUPERNODAL,
             "barrier": capi.VAR_LEVELSET, "syncfree": capi.VAR_AUTO, "write_first": capi.VAR_ROW_PULL,
             "level_pull": capi.VAR_LEVEL_PULL, "tile": capi.VAR_TILE, "chain": capi.VAR_CHAIN, "block": capi.VAR_BLOCK,
             "csc_push": capi.VAR_CSC_PUSH}
    if algo not in table:
        raise SystemExit(f"unknown algorithm {algo}; one of {', '.join(table)}")
    return table[algo]


def run_one(algo, matrix, snodes, rhs, check):
    """one measurement in this process; prints the log and returns seconds"""
    from sptrsv_amd import capi
    variant = _variant(capi, algo)
    n, rp, col, val = capi.read_matrix(matrix)
    print(f"matrix {os.path.basename(matrix)} n {n} nz {len(col)} algo {algo} rhs {rhs}", file=sys.stderr)
    with capi.Plan.from_csr(rp, col, val) as P:
        if variant == capi.VAR_SUPERNODAL:
            P.build_supernodes(capi.read_snodes(snodes) if snodes and os.path.exists(snodes) else None)
        b = np.ones((n, rhs)) if rhs > 1 else np.ones(n)  # the reference fills b with ones (src/utils.cpp:113-126)
        d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
        prm = capi.params(variant=variant)
        P.solve(d_b, d_x, rhs, prm)  # first solve builds lazily created layouts
        P.sync()
        P.timing(True)
        P.solve(d_b, d_x, rhs, prm)
        P.sync()
        ms, _ = P.timing_read()
        P.timing(False)
        if check:
            d_r = capi.DeviceArray(b.shape)
            P.residual(d_b, d_x, d_r, rhs)
            P.sync()
            r = d_r.download()
            print(f"Relative residual: {np.linalg.norm(r) / np.linalg.norm(b):.3e}", file=sys.stderr)
            d_r.free()
        d_b.free()
        d_x.free()
    seconds = ms / 1e3
    print(f"Time: {seconds}", file=sys.stderr)
    return seconds


def single_launch(algo, matrix, snodes, rhs):
    cmd = [sys.executable, "-m", "sptrsv_amd.tables", "--one", algo, matrix, snodes, str(rhs), "check"]
    proc = subprocess.run(cmd, capture_output=True, text=True, cwd=os.path.dirname(os.path.dirname(__file__)))
    log = proc.stderr
    lines = [ln for ln in log.splitlines() if ln.startswith("Time: ")]
    if proc.returncode != 0 or not lines:
        return log, None
    return log, float(lines[-1].split("Time: ")[-1])


def get_times(matrices, algos, repeats, rhs, out_root):
    run_name = datetime.datetime.now().strftime("%m_%d_%H_%M_%S")
    out_dir = os.path.join(out_root, run_name)
    os.makedirs(out_dir, exist_ok=False)
    logs = []
    with open(os.path.join(out_dir, "times.csv"), "w") as f:
        f.write(";" + "".join(a + ";" for a in algos) + "\n")
        for m in matrices:
            stem = os.path.splitext(m)[0]
            snodes = next((s for s in (stem + ".snodes", stem + "_snodes.txt") if os.path.exists(s)), stem + ".snodes")
            f.write(os.path.basename(m) + ";")
            for a in algos:
                best, best_log = None, ""
                for _ in range(repeats):
                    log, t = single_launch(a, m, snodes, rhs)
                    if t is not None and (best is None or t < best):
                        best, best_log = t, log
                    elif best is None:
                        best_log = log
                logs.append(best_log + "-" * 100 + "\n")
                f.write(("error" if best is None else repr(best)) + ";")
                print(f"{os.path.basename(m)} {a}: {best}", flush=True)
            f.write("\n")
    with open(os.path.join(out_dir, "best_times_log.log"), "w") as f:
        f.write("".join(logs))
    return out_dir


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--one", nargs="+", help=argparse.SUPPRESS)
    ap.add_argument("directory", nargs="?")
    ap.add_argument("--algos", default=",".join(ALGOS))
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--rhs", type=int, default=1)
    ap.add_argument("--out", default="experiments")
    args = ap.parse_args(argv)
    if args.one:
        algo, matrix, snodes, rhs = args.one[:4]
        run_one(algo, matrix, snodes, int(rhs), len(args.one) > 4 and args.one[4] == "check")
        return
    if not args.directory:
        ap.error("a directory of matrix files is required")
    matrices = sorted(p for p in glob.glob(os.path.join(args.directory, "*"))
                      if os.path.splitext(p)[1] in (".matrix", ".mtx", ".txt") and "snodes" not in os.path.basename(p))
    out = get_times(matrices, [a for a in args.algos.split(",") if a], args.repeats, args.rhs, args.out)
    print(f"tables in {out}")


if __name__ == "__main__":
    main()
