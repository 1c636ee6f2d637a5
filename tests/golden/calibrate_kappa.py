"""Calibrate the diagonal scales of the condition-number sweep (sptrsv_amd.gen.KAPPA_SCALES).

    python tests/golden/calibrate_kappa.py      # prints the table to paste into gen.py

For each family, kappa_1 of the diagonally scaled factor is estimated on the CPU (oracle.condest,
Hager/Higham) on a grid of scales s = 2^k, then each decade 1e2 ... 1e8 is located by secant steps
in log(s) / log(kappa) (kappa decreases monotonically with s) until it is within 10% of the target.
"""
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import condest  # noqa: E402
from sptrsv_amd import gen  # noqa: E402

TARGETS = [1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8]


def calibrate(base, name):
    cache = {}

    def kappa(s):
        if s not in cache:
            t = time.time()
            cache[s] = condest.kappa1(gen.scale_diagonal(base, s))
            print(f"  {name} s={s:.6g} kappa={cache[s]:.4g} ({time.time() - t:.1f} s)", flush=True)
        return cache[s]

    ks = list(range(12, -9, -1))  # s = 4096 ... 1/256
    grid = []
    for k in ks:
        s = 2.0 ** k
        grid.append((s, kappa(s)))
        if grid[-1][1] > 10 * TARGETS[-1]:
            break
    out = {}
    for tgt in TARGETS:
        lo = hi = None  # lo: kappa below target (larger s), hi: kappa above (smaller s)
        for s, kv in grid:
            if kv <= tgt:
                lo = (s, kv)
            elif hi is None:
                hi = (s, kv)
        if lo is None or hi is None:
            print(f"  {name}: target {tgt:g} outside the grid", flush=True)
            continue
        best = min((lo, hi), key=lambda p: abs(math.log(p[1] / tgt)))
        for _ in range(8):
            if abs(math.log(best[1] / tgt)) < math.log(1.1):
                break
            # secant in log-log between the bracket ends
            a, b = (math.log(lo[0]), math.log(lo[1])), (math.log(hi[0]), math.log(hi[1]))
            f = (math.log(tgt) - a[1]) / (b[1] - a[1]) if b[1] != a[1] else 0.5
            f = min(max(f, 0.05), 0.95)
            s = math.exp(a[0] + f * (b[0] - a[0]))
            s = float(f"{s:.6g}")
            kv = kappa(s)
            if kv <= tgt:
                lo = (s, kv)
            else:
                hi = (s, kv)
            best = min((best, (s, kv)), key=lambda p: abs(math.log(p[1] / tgt)))
        out[tgt] = best
    return out


def main():
    fams = {"zipf_1e6": lambda: gen.random_zipf(1_000_000, seed=1),
            "supernodal_4096": lambda: gen.random_supernodal(4096, seed=2)}
    table = {}
    for name, make in fams.items():
        t = time.time()
        base = make()
        print(f"{name}: n={base.n} nnz={base.nnz} ({time.time() - t:.1f} s to generate)", flush=True)
        table[name] = calibrate(base, name)
    print("KAPPA_SCALES = {")
    for name, res in table.items():
        body = ", ".join(f"{t:.0e}: {s:.6g}" for t, (s, kv) in sorted(res.items()))
        print(f'    "{name}": {{{body}}},')
        print("    # estimated kappa: " + ", ".join(f"{kv:.3g}" for t, (s, kv) in sorted(res.items())))
    print("}")


if __name__ == "__main__":
    main()
