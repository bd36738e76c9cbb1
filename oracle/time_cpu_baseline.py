#!/usr/bin/env python3
"""CPU timings of the reference's phase-space path (the reference harness
built from the reference headers with the reference's own flags,
oracle/_ref/libhydra_ref_native.so, OMP backend) and of the oracle's C
restatement on the same calls, for the CPU table of the baseline document.

Run where the reference harness is built (make -C oracle ref):
    python oracle/time_cpu_baseline.py [--threads 8]
Prints a markdown table: median of 5 timed runs after one warm-up; events/s
= n / seconds of the timed call(s).  Test and documentation infrastructure
only; the product never imports it."""
import argparse
import os
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from oracle import REF_NATIVE_LIB, Oracle, Reference  # noqa: E402

M3, MASSES3 = 5.279, [3.0969, 0.493677, 0.13957]
SEED, UW_SEED = 159753, 0x180EC6D33CFD0ABA
BW3 = [(0.892, 0.05, 0.4967141530, -0.1382643012, 1, 2),
       (1.43, 0.11, 0.6476885381, 1.5230298564, 1, 2),
       (0.775, 0.149, -0.2341533747, -0.2341368981, 0, 2)]


def bw_par():
    flat = [float(len(BW3))]
    for t in BW3:
        flat.extend(float(x) for x in t)
    return np.array(flat)


def median_of(fn, reps=5):
    fn()
    return statistics.median(fn() for _ in range(reps))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    ref = Reference(REF_NATIVE_LIB)
    ora = Oracle()
    rows = []

    def ref_gen(M, m, n, threads):
        ref.set_threads(threads)
        return median_of(lambda: ref.generate(M, m, SEED, n)[2])

    def ora_gen(M, m, n, threads):
        def one():
            t0 = time.perf_counter()
            ora.generate(M, m, SEED, n, nthreads=threads)
            return time.perf_counter() - t0
        return median_of(one)

    n1 = 1000000
    for threads in sorted({args.threads, 1}, reverse=True):
        t = ref_gen(M3, MASSES3, n1, threads)
        rows.append(("C1: 3-body, Generate", n1, threads, "reference (OMP)", n1 / t))
        to = ora_gen(M3, MASSES3, n1, threads)
        rows.append(("C1 (same)", n1, threads, "oracle restatement (OpenMP)", n1 / to))
        rows.append(("C1 restatement / reference speed", n1, threads, "ratio", t / to))
    ref.set_threads(args.threads)
    n2 = 1 << 20
    t = median_of(lambda: ref.unweight(M3, MASSES3, SEED, n2, UW_SEED, want_output=True)[3])
    rows.append(("C2-CPU: Generate + Unweight()", n2, args.threads, "reference (OMP)", n2 / t))
    for tag, m in (("C3-CPU: 10-body Generate", [0.13957] * 10),
                   ("C3-CPU near threshold", [0.999] * 10)):
        t = ref_gen(10.0, m, n2, args.threads)
        rows.append((tag, n2, args.threads, "reference (OMP)", n2 / t))
    n4 = 1 << 24
    t = median_of(lambda: ref.average_on(M3, MASSES3, SEED, n4, fid=1, fpar=bw_par())[2], reps=3)
    rows.append(("C4-CPU: AverageOn, 3-BW Dalitz functor", n4, args.threads, "reference (OMP)",
                 n4 / t))
    print("| config | events | threads | code | events/s |")
    print("|---|---|---|---|---|")
    for name, n, th, code, v in rows:
        val = f"{v:.3g}" if code != "ratio" else f"{v:.2f}"
        print(f"| {name} | {n} | {th} | {code} | {val} |")


if __name__ == "__main__":
    main()
