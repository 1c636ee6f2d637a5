"""Benchmark of the sparse upper-triangular solve on one MI355X.

Default workload: U = IC(0) factor of the 7-point Laplacian on a 128^3 grid (n = 2,097,152,
nnz = 8,339,456), FP64, one right-hand side, b = U x_true with x_true ~ U(-1,1) (seed 1) --
the single-GPU sync-free configuration the headline metric is quoted on.  One step = one full
solve (sentinel fill + ticket reset + solve kernel) with U, b and x resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant level_pull] [--m 1]

Prints ONE JSON line.  `value` is achieved GB/s over whole steps, with the algorithmic byte
count  nnz*(8+4) + 4*(n+1) + 8*n*m (b) + 8*n*m (x)  (index and value of every stored entry, the
row pointers, b read once, x written once).  `roofline.achieved` divides the same bytes by the
solve kernel's own average duration, timed with HIP events on the plan's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6


def progress(msg):
    """One line per finished section on stderr (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_probe(nbytes, reps):
    """Plain device-to-device copies of a known size: the calibration point for the
    FETCH_SIZE / WRITE_SIZE counters (read nbytes and write nbytes per copy)."""
    from sptrsv_amd import capi
    src = capi.DeviceArray.from_host(np.ones(nbytes // 8))
    dst = capi.DeviceArray((nbytes // 8,))
    for _ in range(reps):
        capi.check(capi.lib().sptrsv_memcpy_d2d(dst.ptr, src.ptr, nbytes), "memcpy_d2d")
    capi.device_sync()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def alg_bytes(n, nnz, m, vbytes=8):
    return nnz * (vbytes + 4) + 4 * (n + 1) + 2 * vbytes * n * m


def supernodal_flops(U, m):
    """Algorithmic flops of a supernodal solve: per supernode of width w with r rectangular
    columns, 2*w*r*m (rectangular product) + w*w*m (triangle)."""
    s = np.asarray(U.snodes, dtype=np.int64)
    w = np.diff(s)
    r = (U.rowptr[s[:-1] + 1] - U.rowptr[s[:-1]]) - w
    return float(np.sum((2 * w * r + w * w) * m))


_WORKLOADS = {}


def make_workload(name):
    """(factor, description), generated once per process"""
    if name not in _WORKLOADS:
        _WORKLOADS[name] = _make_workload(name)
    return _WORKLOADS[name]


def _make_workload(name):
    from sptrsv_amd import gen
    if name == "ic0_128":
        return gen.laplace3d_ic0(128), "IC(0) of 7-point Laplacian, 128^3 grid"
    if name == "zipf_1e6":
        return gen.random_zipf(1_000_000, seed=1), "random upper, n=1e6, Poisson(8) rows, Zipf(1.5) offsets"
    if name == "chol2d_100":
        return gen.laplace2d_cholesky(100), "exact Cholesky of 5-point Laplacian, 100^2 grid"
    if name == "supernodal_4096":
        return gen.random_supernodal(4096, seed=2), "random supernodal, 4096 supernodes, widths U[16,128]"
    raise SystemExit(f"unknown workload {name}")


def time_solves(capi, P, d_b, d_x, m, prm, steps, warmup, first=None):
    for i in range(warmup):
        t0 = time.perf_counter()
        P.solve(d_b, d_x, m, prm)
        if first is not None:
            P.sync()
            first.append(round((time.perf_counter() - t0) * 1e3, 3))
    P.sync()
    capi.device_sync()
    P.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        P.solve(d_b, d_x, m, prm)
    P.sync()
    capi.device_sync()
    t1 = time.perf_counter()
    kms, launches = P.timing_read()
    P.timing(False)
    return (t1 - t0) * 1000.0 / steps, kms / max(launches, 1)


def read_traffic(workload, variant, m):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        table = json.load(open(path))
        return table.get(f"{workload}/{variant}/m{m}")
    except Exception:
        return None


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS", "")}


def cpu_baseline(workloads, headline, m):
    """The reference's CPU methods restated in the oracle, timed on the host in 3 fresh processes
    (python -m oracle.cpu_baseline, like run.py:24-32), minimum over the processes.  The headline
    figure is the serial solve (one core) of the bench workload; the OpenMP level-set solve (all
    threads) and, on the supernodal factor, the supernodal and CBLAS supernodal solves ride along."""
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="sptrsv_cpu_")
    names = []
    for name, U in workloads.items():
        arrays = {"rowptr": U.rowptr, "col": U.col, "val": U.val}
        if U.snodes is not None:
            arrays["snodes"] = U.snodes
        np.savez(os.path.join(tmp, name + ".npz"), **arrays)
        names.append(name)
    runs = []
    for _ in range(3):
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", tmp, *names], capture_output=True, text=True,
                           cwd=ROOT, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"cpu baseline process failed: {r.stderr[-2000:]}")
        runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    for name in names:
        os.remove(os.path.join(tmp, name + ".npz"))
    os.rmdir(tmp)
    best = {}
    for name in names:
        rec = dict(runs[0][name])
        for key in rec:
            if key.endswith("_ms"):
                rec[key] = round(min(r[name][key] for r in runs), 4)
        best[name] = rec
    h = best[headline]
    U = workloads[headline]
    return {
        "value": alg_bytes(U.n, U.nnz, m) / (h["serial_ms"] / 1e3) / 1e9,
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"serial solve of the same factor (m={m}) repeated within 2 s per process, "
                   f"minimum over 3 fresh processes; {h['serial_ms']:.3f} ms"),
        "serial_ms": h["serial_ms"],
        "openmp_levelset_ms": h["levelset_ms"],
        "openmp_threads": h["levelset_threads"],
        "host": host_info(),
        "workloads": best,
    }


def plan_timing(O, U, P, plan_s):
    """Plan creation (upload + device analysis) and its parts, next to the oracle's level
    analysis on the host (setLevelUpAndGetMaxLevel, src/core.cpp:244-270; one core)."""
    a_ms, l_ms = P.analysis_time()
    t0 = time.perf_counter()
    O.levels(U.rowptr, U.col)
    host_ms = (time.perf_counter() - t0) * 1e3
    return {"plan_ms": round(plan_s * 1e3, 2), "analysis_ms": round(a_ms, 2), "levels_ms": round(l_ms, 2),
            "levels": P.n_levels, "oracle_host_levels_ms": round(host_ms, 2)}


def bench_mixed(capi, gen, names, out, only=""):
    """Wall time of the refined solve (reduced-precision solves + FP64 residual and update on the
    GPU, stopped at an estimated relative error <= 1e-13) against the plain FP64 solve of the
    same kernel family and (m <= 32) the plain FP64 sync-free CSC push solve, data resident on
    the GPU; rel_err_vs_fp64 is measured, not estimated."""
    from sptrsv_amd import gen as G
    cases = [("ic0_128", "level_pull", 32, capi.FP32), ("ic0_128", "level_pull", 128, capi.FP32),
             ("zipf_1e6", "chain", 1, capi.FP32), ("supernodal_4096", "supernodal", 128, capi.BF16),
             ("supernodal_4096", "supernodal", 32, capi.BF16)]
    cache = {}
    for wl, vname, m, prec in cases:
        key = f"{wl}/{vname}/m{m}"
        if only and key != only:
            continue
        try:
            if wl not in cache:
                U, _ = make_workload(wl)
                P = capi.Plan.from_upper(U)
                if wl.startswith("supernodal"):
                    P.build_supernodes(U.snodes)
                cache = {wl: (U, P)}  # keep one workload resident at a time
            U, P = cache[wl]
            _, b = G.rhs_for(U, m, seed=1, layout="row")
            d_b = capi.DeviceArray.from_host(b)
            d_x = capi.DeviceArray(b.shape)
            d_r = capi.DeviceArray(b.shape)
            prm = capi.params(variant=names[vname])
            P.solve(d_b, d_x, m, prm)
            P.sync()
            t64 = []
            for _ in range(3):
                t0 = time.perf_counter()
                P.solve(d_b, d_x, m, prm)
                P.sync()
                t64.append(time.perf_counter() - t0)
            x64 = d_x.download()
            # the fastest FP64 solve the library has for this system (AUTO), for reference
            prm_auto = capi.params()
            P.solve(d_b, d_r, m, prm_auto)
            P.sync()
            tauto = []
            for _ in range(3):
                t0 = time.perf_counter()
                P.solve(d_b, d_r, m, prm_auto)
                P.sync()
                tauto.append(time.perf_counter() - t0)
            tpush = None
            if m <= 32:
                # the plain FP64 sync-free kernel (CSC push, the reference's algorithm) on the same
                # system: the refined solve's other point of comparison
                prm_push = capi.params(variant=capi.VAR_CSC_PUSH)
                P.solve(d_b, d_r, m, prm_push)
                P.sync()
                t0 = time.perf_counter()
                P.solve(d_b, d_r, m, prm_push)
                P.sync()
                tpush = time.perf_counter() - t0
            P.solve_refined(d_b, d_r, m, prm, precision=prec)  # warm-up (builds reduced copies)
            tref = []
            for _ in range(3):
                t0 = time.perf_counter()
                it, ratio = P.solve_refined(d_b, d_r, m, prm, precision=prec)
                tref.append(time.perf_counter() - t0)
            xr = d_r.download()
            err = float(np.linalg.norm(xr - x64) / np.linalg.norm(x64))
            out[key] = {"fp64_ms": round(min(t64) * 1e3, 3), "refined_ms": round(min(tref) * 1e3, 3),
                        "speedup": round(min(t64) / min(tref), 3), "iters": it, "error_estimate": ratio,
                        "rel_err_vs_fp64": err, "low_precision": "bf16+fp32" if prec == capi.BF16 else "fp32",
                        "fp64_auto_ms": round(min(tauto) * 1e3, 3),
                        "fp64_auto_variant": capi.VARIANT_NAMES.get(P.auto_variant(m), "?")}
            if tpush is not None:
                out[key]["fp64_syncfree_ms"] = round(tpush * 1e3, 3)
                out[key]["speedup_vs_syncfree"] = round(tpush / min(tref), 3)
            for d in (d_b, d_x, d_r):
                d.free()
        except Exception as exc:
            out[key] = {"error": str(exc)}
        progress(f"mixed {key}: {out[key]}")


KAPPA_TARGETS = (1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8)


def bench_kappa(capi, gen, out, plan_ms, targets=None):
    """Mixed precision with FP64 refinement across the condition-number sweep: the Zipf factor
    (n = 1e6) and the supernodal factor (4,096 supernodes) with the diagonal scaled so that
    kappa_1 spans 1e2 ... 1e8 (gen.KAPPA_SCALES).  Per factor: kappa re-estimated on the GPU
    (sptrsv_condest); per m: the plain FP64 solve, the refined solve (FP32 inner solves for Zipf,
    BF16 rectangular products for the supernodal factor) with the FP64 fallback on, wall times
    with data resident, and the refined x against the serial oracle.  plan_ms: family -> m list;
    targets: the kappa values (all of KAPPA_TARGETS by default)."""
    from oracle import oracle as O
    for fam, ms in plan_ms.items():
        base = gen.random_zipf(1_000_000, seed=1) if fam == "zipf_1e6" else gen.random_supernodal(4096, seed=2)
        sn = fam.startswith("supernodal")
        for tgt in (targets or KAPPA_TARGETS):
            key = f"{fam}/kappa{tgt:.0e}"
            res = {"diag_scale": gen.KAPPA_SCALES[fam][tgt]}
            try:
                U = gen.scale_diagonal(base, gen.KAPPA_SCALES[fam][tgt])
                t0 = time.perf_counter()
                P = capi.Plan.from_upper(U)
                if sn:
                    P.build_supernodes(U.snodes)
                res["plan_s"] = round(time.perf_counter() - t0, 3)
                t0 = time.perf_counter()
                res["kappa1_gpu"], _ = P.condest()
                res["condest_s"] = round(time.perf_counter() - t0, 3)
                for m in ms:
                    xt, b = gen.rhs_for(U, m, seed=1, layout="row")
                    ref = (O.base_gauss_upper(U.rowptr, U.col, U.val, b)[0] if m == 1 else
                           np.ascontiguousarray(O.base_gauss_upper_columns(U.rowptr, U.col, U.val,
                                                                           np.ascontiguousarray(b.T)).T))
                    d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
                    prm64 = capi.params(variant=capi.VAR_SUPERNODAL if sn else capi.VAR_AUTO)
                    P.solve(d_b, d_x, m, prm64)
                    P.sync()
                    t64 = []
                    for _ in range(2):
                        t0 = time.perf_counter()
                        P.solve(d_b, d_x, m, prm64)
                        P.sync()
                        t64.append(time.perf_counter() - t0)
                    e64 = O.rel_error(ref, d_x.download())
                    prec = capi.BF16 if sn else capi.FP32
                    prm_lo = capi.params(variant=capi.VAR_SUPERNODAL if sn else capi.VAR_AUTO)
                    P.solve_refined_ex(d_b, d_x, m, prm_lo, precision=prec, fallback=True)  # warm-up
                    tr, info = [], None
                    for _ in range(2):
                        t0 = time.perf_counter()
                        info = P.solve_refined_ex(d_b, d_x, m, prm_lo, precision=prec, fallback=True)
                        tr.append(time.perf_counter() - t0)
                    err = O.rel_error(ref, d_x.download())
                    res[f"m{m}"] = {"fp64_ms": round(min(t64) * 1e3, 3), "fp64_rel_err_vs_oracle": e64,
                                    "refined_ms": round(min(tr) * 1e3, 3), "speedup": round(min(t64) / min(tr), 3),
                                    "inner": "bf16+fp32" if sn else "fp32", "inner_variant":
                                        capi.VARIANT_NAMES.get(info["variant"], "?"), "iters": info["iters"],
                                    "outcome": info["outcome"], "fell_back": info["fell_back"],
                                    "estimate": info["estimate"], "rel_err_vs_oracle": err}
                    d_b.free()
                    d_x.free()
                P.close()
            except Exception as exc:  # report, never hide
                res["error"] = str(exc)
            out[key] = res
            progress(f"kappa {key}: {res}")


SELECTOR_PATH = os.path.join(ROOT, "sptrsv_amd", "selector.json")


def bench_batch(capi, gen, out, count=4096, nmax=65536, ms=(1,)):
    """Batched solve of `count` systems from the native generator (seed 3; n log-uniform
    [256, nmax]; banded / Zipf / supernodal shaped like the 4,096-supernode factor) in one
    persistent graph-replayed launch, for each m: the plain fallback configuration (LS1),
    exhaustive autotuning, the MLP selector (trained by `python -m sptrsv_amd.tune` on 20,000
    systems of seed 4, loaded from sptrsv_amd/selector.json), the selector with its longest
    systems re-decided, and the FP32 + FP64-refinement solve with the selector's choices.
    b = U x_true is formed on the device (the batch residual with b = 0 gives -U x_true)."""
    from sptrsv_amd import tune
    sel = tune.Selector.load(SELECTOR_PATH)
    t0 = time.perf_counter()
    U, sp, kinds = gen.batch_systems_native(count, seed=3, nmin=256, nmax=nmax)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    B = capi.Batch(U, sp)
    t_create = time.perf_counter() - t0
    res = {"systems": count, "n_total": U.n, "nnz_total": U.nnz, "nmax": nmax,
           "kinds": {k: kinds.count(k) for k in gen.BATCH_KINDS}, "generate_s": round(t_gen, 1),
           "batch_create_s": round(t_create, 1), "selector": sel.meta}
    progress(f"batch: {count} systems, n={U.n}, nnz={U.nnz}, generated {t_gen:.1f} s, created {t_create:.1f} s")
    rng = np.random.default_rng(1)
    for m in ms:
        shape = (U.n, m) if m > 1 else (U.n,)
        xt = rng.uniform(-1.0, 1.0, shape)
        d_xt = capi.DeviceArray.from_host(xt)
        d_0 = capi.DeviceArray.from_host(np.zeros(shape))
        d_b, d_x = capi.DeviceArray(shape), capi.DeviceArray(shape)
        B.residual(d_0, d_xt, d_b, m)  # d_b = -U x_true: the solution is -x_true
        B.sync()
        d_0.free()
        T = tune.exhaustive(B, d_b, d_x, m=m, reps=2)
        feats = B.features(m)
        mlp = sel.predict(feats, m=m)
        choices = {"fallback_ls1": np.full(B.nsys, capi.BCFG_LS1), "exhaustive": np.argmin(T, axis=1),
                   "mlp": mlp, "mlp_guarded": tune.guard_longest(U, sp, feats, mlp, m)}
        rm = {}
        walls = {name: [] for name in choices}
        for _ in range(2):  # the choices alternate, so a transient slows all of them alike
            for name, ch in choices.items():
                B.set_configs(ch, tune.lpt_order(T[np.arange(B.nsys), ch]))
                B.solve(d_b, d_x, m, capi.FP64, True)
                B.sync()
                for _ in range(3 if name != "fallback_ls1" else 1):
                    t0 = time.perf_counter()
                    B.solve(d_b, d_x, m, capi.FP64, True)
                    B.sync()
                    walls[name].append(time.perf_counter() - t0)
        for name, ch in choices.items():
            cost = T[np.arange(B.nsys), ch]
            best = min(walls[name])
            # the launch cannot end before its longest system: one slow pick shows here, not in
            # the summed regret
            rm[name] = {"ms": round(best * 1e3, 3), "gnnz_per_s": round(U.nnz * m / best / 1e9, 3),
                        "sum_device_us": round(float(cost.sum()), 1), "longest_system_us": round(float(cost.max()), 1),
                        "regret": round(tune.regret(T, ch), 4)}
        rm["mlp_vs_exhaustive_throughput"] = round(rm["exhaustive"]["ms"] / rm["mlp"]["ms"], 4)
        rm["mlp_guarded_vs_exhaustive_throughput"] = round(rm["exhaustive"]["ms"] / rm["mlp_guarded"]["ms"], 4)
        rm["mlp_vs_fallback_throughput"] = round(rm["fallback_ls1"]["ms"] / rm["mlp"]["ms"], 4)
        rm["mlp_accuracy"] = float((choices["mlp"] == choices["exhaustive"]).mean())
        ch = choices["mlp"]
        B.set_configs(ch, tune.lpt_order(T[np.arange(B.nsys), ch]))
        B.solve(d_b, d_x, m, capi.FP64, True)
        B.sync()
        x = d_x.download()
        rm["fp64_forward_err_vs_xtrue"] = float(np.linalg.norm(x + xt) / np.linalg.norm(xt))
        # FP32 solves + FP64 refinement with the selector's configurations, FP64 fallback on
        B.solve_refined_ex(d_b, d_x, m, fallback=True)
        t0 = time.perf_counter()
        info = B.solve_refined_ex(d_b, d_x, m, fallback=True)
        t_ref = time.perf_counter() - t0
        xr = d_x.download()
        rm["fp32_refined"] = {"ms": round(t_ref * 1e3, 3), **info,
                              "rel_err_vs_fp64": float(np.linalg.norm(xr - x) / np.linalg.norm(x))}
        res[f"m{m}"] = rm
        progress(f"batch m={m}: {rm}")
        for d in (d_b, d_x, d_xt):
            d.free()
        del x, xr, xt
    out["batch"] = res
    sel.close()
    B.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ic0_128")
    ap.add_argument("--variant", default="block")
    ap.add_argument("--m", type=int, default=1)
    ap.add_argument("--occupancy", type=int, default=0, help="params.occupancy of the measured solve")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run N-byte device copies (counter calibration)")
    ap.add_argument("--suite", default="", help="kappa: only the full condition-number sweep (m in 1, 32, 128); "
                    "batch: only the batched solve at m in 1, 4, 16")
    ap.add_argument("--no-kappa", action="store_true")
    ap.add_argument("--mixed-case", default="", help="run only this refined-solve case, e.g. "
                    "supernodal_4096/supernodal/m128 (profiling)")
    args = ap.parse_args()
    t_start = time.perf_counter()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the solve path here is single-GPU "
              "(multi-GPU sharding is not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe, 5)
        return
    from oracle import oracle as O
    from sptrsv_amd import capi, gen

    if args.suite == "kappa":
        sweep = {}
        bench_kappa(capi, gen, sweep, {"zipf_1e6": (1, 32, 128), "supernodal_4096": (1, 32, 128)})
        print(json.dumps({"suite": "kappa", "kappa_sweep": sweep, "bench_wall_s": round(time.perf_counter() - t_start, 1)}))
        return
    if args.suite == "batch":
        batch = {}
        bench_batch(capi, gen, batch, ms=(1, 4, 16))
        print(json.dumps({"suite": "batch", "batched": batch, "bench_wall_s": round(time.perf_counter() - t_start, 1)}))
        return

    names = {v: k for k, v in capi.VARIANT_NAMES.items()}
    variant = names[args.variant]
    U, desc = make_workload(args.workload)
    m = args.m
    t0 = time.perf_counter()
    P = capi.Plan.from_upper(U)
    analysis = {args.workload: plan_timing(O, U, P, time.perf_counter() - t0)}
    progress(f"analysis {args.workload}: {analysis[args.workload]}")
    if variant == capi.VAR_SUPERNODAL:
        P.build_supernodes(U.snodes)
    xt, b = gen.rhs_for(U, m, seed=1, layout="row")
    d_b = capi.DeviceArray.from_host(b)
    d_x = capi.DeviceArray(b.shape)
    prm = capi.params(variant=variant, occupancy=args.occupancy)
    first = []
    ms, kms = time_solves(capi, P, d_b, d_x, m, prm, args.steps, args.warmup, first)
    x = d_x.download()
    bb = b if m == 1 else np.ascontiguousarray(b.T)
    ref, _ = O.base_gauss_upper(U.rowptr, U.col, U.val, bb)
    err = O.rel_error(ref if m == 1 else ref.T, x)
    B = alg_bytes(U.n, U.nnz, m)
    gbs = B / (ms / 1e3) / 1e9
    achieved = B / (kms / 1e3) / 1e9
    name, cus, mem = capi.device_info()
    progress(f"{args.workload}/{args.variant}/m{m}: {ms:.4f} ms/step, kernel {kms:.4f} ms, rel err {err:.2e}")

    extras = {}
    if not args.no_extras:
        runs = [("ic0_128", "tile", 1), ("ic0_128", "level_pull", 1), ("ic0_128", "row_pull", 1),
                ("ic0_128", "levelset", 1), ("ic0_128", "levelset", 32),
                ("ic0_128", "level_pull", 32), ("ic0_128", "panel", 32), ("ic0_128", "panel", 64),
                ("ic0_128", "level_pull", 128), ("ic0_128", "panel", 128), ("zipf_1e6", "tile", 1),
                ("zipf_1e6", "chain", 1), ("zipf_1e6", "auto", 1), ("chol2d_100", "chain", 1),
                ("chol2d_100", "auto", 1), ("chol2d_100", "level_pull", 1),
                ("supernodal_4096", "supernodal", 8), ("supernodal_4096", "supernodal", 32),
                ("supernodal_4096", "supernodal", 128)]
        cache = {args.workload: (U, P)}
        col_refs = {}  # (workload, m) -> oracle x for several right-hand sides (row-major)
        for wl, vname, mm in runs:
            if (wl, vname, mm) == (args.workload, args.variant, m):
                continue
            try:
                if wl not in cache:
                    UU, _ = make_workload(wl)
                    t0 = time.perf_counter()
                    cache[wl] = (UU, capi.Plan.from_upper(UU))
                    analysis[wl] = plan_timing(O, UU, cache[wl][1], time.perf_counter() - t0)
                    progress(f"analysis {wl}: {analysis[wl]}")
                    if wl.startswith("supernodal"):
                        cache[wl][1].build_supernodes(UU.snodes)
                UU, PP = cache[wl]
                _, bx = gen.rhs_for(UU, mm, seed=1, layout="row")
                db = capi.DeviceArray.from_host(bx)
                dx = capi.DeviceArray(bx.shape)
                e_first = []
                e_ms, e_kms = time_solves(capi, PP, db, dx, mm, capi.params(variant=names[vname]),
                                          3 if mm > 1 or wl != "ic0_128" else 10, 2, e_first)
                BB = alg_bytes(UU.n, UU.nnz, mm)
                xg = dx.download()
                if mm == 1 and wl != "ic0_128":
                    ref_w, t_cpu = O.base_gauss_upper(UU.rowptr, UU.col, UU.val, bx)
                    cpu_ms = {f"{wl}/cpu_serial_ms": round(t_cpu * 1e3, 3)}
                    extras.update(cpu_ms)
                    e_err = O.rel_error(ref_w, xg)
                elif wl.startswith("supernodal") and mm == 8:
                    # CPU baselines on the same factor (one host core each, one solve): the serial
                    # solver and the reference's supernodal method, restated in the oracle
                    bcol = np.ascontiguousarray(bx.T)
                    ref_w, t_ser = O.base_gauss_upper(UU.rowptr, UU.col, UU.val, bcol)
                    _, t_sn = O.supernodal_upper(UU.rowptr, UU.col, UU.val, UU.snodes, bcol, new=True)
                    extras[f"{wl}/m{mm}/cpu_serial_ms"] = round(t_ser * 1e3, 3)
                    extras[f"{wl}/m{mm}/cpu_supernodal_ms"] = round(t_sn * 1e3, 3)
                    e_err = O.rel_error(np.ascontiguousarray(ref_w.T), xg)
                else:
                    # the serial oracle (several right-hand sides: column by column on host
                    # threads, bitwise the serial solve), once per workload and m
                    if (wl, mm) not in col_refs:
                        col_refs[(wl, mm)] = O.base_gauss_upper(UU.rowptr, UU.col, UU.val, bx)[0] if mm == 1 else \
                            np.ascontiguousarray(O.base_gauss_upper_columns(UU.rowptr, UU.col, UU.val,
                                                                            np.ascontiguousarray(bx.T)).T)
                    e_err = O.rel_error(col_refs[(wl, mm)], xg)
                if vname == "supernodal":
                    extras[f"{wl}/{vname}/m{mm}/TFLOP_s"] = round(supernodal_flops(UU, mm) / (e_kms / 1e3) / 1e12, 3)
                extras[f"{wl}/{vname}/m{mm}"] = {"ms_per_step": round(e_ms, 4), "kernel_ms": round(e_kms, 4),
                                                 "rel_err_vs_oracle": e_err,
                                                 "GB_s": round(BB / (e_ms / 1e3) / 1e9, 2),
                                                 "nnz_per_s": UU.nnz * mm / (e_ms / 1e3), "first_ms": e_first}
                db.free()
                dx.free()
            except Exception as exc:  # report, never hide
                extras[f"{wl}/{vname}/m{mm}"] = {"error": str(exc)}
            progress(f"extra {wl}/{vname}/m{mm}: {extras[f'{wl}/{vname}/m{mm}']}")

    if not args.no_extras and m == 1:
        # forward substitution with U^T (companion plan) and the full Cholesky solve U^T U x = b
        # on the same plan and kernel family, against the oracle's triangular solves
        cp, ri, cv = O.transpose(U.rowptr, U.col, U.val)
        y_ref = O.base_gauss_lower(cp, ri, cv, b.reshape(-1))[0]
        refs = {"transposed": y_ref, "cholesky": O.base_gauss_upper(U.rowptr, U.col, U.val, y_ref)[0]}
        for kind, fn in (("transposed", P.solve_transposed), ("cholesky", P.solve_cholesky)):
            key = f"{args.workload}/{args.variant}/m1/{kind}"
            try:
                d_b, d_x = capi.DeviceArray.from_host(b), capi.DeviceArray(b.shape)
                for _ in range(2):
                    fn(d_b, d_x, 1, prm)
                P.sync()
                reps = 10
                P.timing(True)
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn(d_b, d_x, 1, prm)
                P.sync()
                wall = (time.perf_counter() - t0) / reps
                tot_ms, launches = P.timing_read()
                P.timing(False)
                xk = d_x.download().reshape(-1)
                extras[key] = {"ms_per_step": round(wall * 1e3, 4), "kernel_ms": round(tot_ms / reps, 4),
                               "solve_launches": launches // reps,
                               "rel_err_vs_oracle": O.rel_error(refs[kind], xk)}
                d_b.free()
                d_x.free()
            except Exception as exc:  # report, never hide
                extras[key] = {"error": str(exc)}
            progress(f"extra {key}: {extras[key]}")

    if not args.no_extras:
        # the host-buffer boundary (b uploaded, x downloaded inside the call): PCIe-inclusive rate
        P.solve_host(b, prm)
        th = []
        for _ in range(5):
            t0 = time.perf_counter()
            P.solve_host(b, prm)
            th.append(time.perf_counter() - t0)
        extras[f"{args.workload}/{args.variant}/m{m}/host_buffers"] = {
            "ms": round(min(th) * 1e3, 4), "GB_s": round(B / min(th) / 1e9, 2),
            "note": "sptrsv_solve_host: b host->device, solve, x device->host (plan resident)"}
        progress(f"host-buffer solve: {extras[f'{args.workload}/{args.variant}/m{m}/host_buffers']}")
    cpu = None
    if not args.no_cpu:
        try:
            wl = {args.workload: U}
            if m == 1 and not args.no_extras:
                for extra in ("chol2d_100", "zipf_1e6", "supernodal_4096"):
                    wl[extra] = make_workload(extra)[0]
            cpu = cpu_baseline(wl, args.workload, m)
            progress(f"cpu baseline: {cpu['workloads']}")
        except Exception as exc:  # report, never hide
            cpu = {"error": str(exc)}
    del P
    mixed, batch, sweep = {}, {}, {}
    if not args.no_mixed:
        bench_mixed(capi, gen, names, mixed, only=args.mixed_case)
    if not args.no_kappa:
        bench_kappa(capi, gen, sweep, {"zipf_1e6": (1,), "supernodal_4096": (32,)}, targets=(1e2, 1e8))
    if not args.no_batch:
        try:
            bench_batch(capi, gen, batch)
        except Exception as exc:
            batch["error"] = str(exc)
    if variant == capi.VAR_SUPERNODAL:
        flops = supernodal_flops(U, m)
        tfs = flops / (kms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": round(tfs, 3), "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(tfs / FP64_MFMA_PEAK_TFS, 5), "traffic": read_traffic(args.workload, args.variant, m),
                "kernel_ms": round(kms, 5), "alg_flops": flops, "alg_bytes": B,
                "hbm_GB_s": round(achieved, 3)}
    else:
        roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": read_traffic(args.workload, args.variant, m),
                "kernel_ms": round(kms, 5), "alg_bytes": B}
    out = {
        "metric": "SpTRSV GB/s & % HBM peak; time-to-1e-12 error, mixed+refine vs FP64 sync-free",
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prm.precision == capi.FP64 else "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "n": U.n, "nnz": U.nnz, "m": m,
                   "variant": args.variant, "device": name, "compute_units": cus},
        "pct_hbm_peak": round(100.0 * gbs / HBM_PEAK_GBS, 3),
        "nnz_per_s": U.nnz * m / (ms / 1e3),
        "rel_err_vs_oracle": err,
        "warmup_solves_ms": first,
        "roofline": roof,
        "cpu_baseline": cpu,
        "extras": extras,
        "analysis": analysis,
        "mixed_refine": mixed,
        "kappa_sweep": sweep,
        "batched": batch,
        "bench_wall_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
