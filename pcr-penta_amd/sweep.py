#!/usr/bin/env python3
"""Timing sweep over the reference's grid of batch and system sizes.

Counterpart of the reference's driver (main.cu:236-305): 17 batch sizes x 10
system sizes (main.cu:243-244), each point the mean of `runs` event-timed
solves (main.cu:254, 276-286), one row per point in the format of
time_perform_analysis_oddeven.dat:

    m \\t n \\t t1 \\t t2 \\t t3          (milliseconds, "%d \\t %d \\t %f \\t %f \\t %f")

The reference's three columns are its three stage-1 variants, which give the
same result (kernel.cu:158-320, 325-526, 530-659). Here one solver replaces
them, so the columns are:

    t1  plain FP64, system-contiguous layout (the reference's, kernel.cu:43),
        on the reference's own input, initialize()'s T^2 system with y = 1
        (kernel.cu:5-26);
    t2  plain FP64, interleaved layout, same input;
    t3  FP32 inner solve + FP64 refinement (max_sweeps 8, tol 1e-12),
        interleaved, on the seeded strictly dominant generator (T^2 is too
        ill-conditioned, cond ~ (2(n+1)/pi)^4, for an FP32 inner solve).

The reference writes zeros for m > 1456 with n = 131072 (main.cu:263-271: out
of memory on its GPU); on one MI355X every point fits and is measured.
Timing covers the solve launches only (inputs already in device memory;
main.cu:42-52 times the kernel pair the same way).
"""
from __future__ import annotations

import argparse
import os
import sys

BATCH = [26, 52, 104, 156, 208, 260, 312, 364, 416, 624, 832, 1040, 1248, 1456, 1664, 1872,
         2080]
SIZES = [256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072]
AVG_RUN = 10
OUT_NAME = "time_perform_analysis_oddeven.dat"


def grid(batches=None, sizes=None):
    """(m, n) in the reference's loop order: batch size outer, system size inner."""
    for m in (batches or BATCH):
        for n in (sizes or SIZES):
            yield m, n


def format_row(m: int, n: int, t1: float, t2: float, t3: float) -> str:
    return "%d \t %d \t %f \t %f \t %f\n" % (m, n, t1, t2, t3)


def _mean_ms(fn, reset, runs, stream):
    """Mean event-timed duration (ms) of fn() on `stream`; reset() restores the
    right-hand side before each run, outside the timed interval."""
    import torch
    reset()
    fn()                                  # first launch outside the timing
    total = 0.0
    for _ in range(runs):
        reset()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        total += e0.elapsed_time(e1)
    return total / runs


def time_point(m: int, n: int, runs: int = AVG_RUN):
    """Mean solve times (ms) of the three columns at one (m, n)."""
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import pcrpenta
    stream = torch.cuda.current_stream()
    out = []
    for layout in (pcrpenta.CONTIGUOUS, pcrpenta.INTERLEAVED):
        a, b, c, d, e, y0 = pcrpenta.init_t2(n, m, layout)
        y = y0.clone()
        ws, _ = pcrpenta._workspace(n, m, pcrpenta.PREC_F64, y.device)
        out.append(_mean_ms(
            lambda: pcrpenta.solve(a, b, c, d, e, y, n, m, layout, stream=stream, ws=ws),
            lambda: y.copy_(y0), runs, stream))
        del a, b, c, d, e, y0, y, ws
    a, b, c, d, e, y0 = pcrpenta.generate(0, 0, n, m, pcrpenta.INTERLEAVED)
    y = y0.clone()
    ws, _ = pcrpenta._workspace(n, m, pcrpenta.PREC_F32_REFINE, y.device)
    out.append(_mean_ms(
        lambda: pcrpenta.solve_refined(a, b, c, d, e, y, n, m, pcrpenta.INTERLEAVED,
                                       max_sweeps=8, tol=1e-12, stream=stream, ws=ws),
        lambda: y.copy_(y0), runs, stream))
    return tuple(out)


def read_dat(path):
    """{(m, n): (t1, t2, t3)} from a file in the row format above; header
    lines and the reference's zero rows (points it did not run) are skipped."""
    rows = {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) != 5:
                continue
            try:
                m, n = int(parts[0]), int(parts[1])
                t = tuple(float(v) for v in parts[2:])
            except ValueError:
                continue
            if all(v > 0 for v in t):
                rows[(m, n)] = t
    return rows


def compare(ours_path, ref_path, cols=(26, 260, 1040, 1248, 2080)):
    """Markdown table: the reference's fastest column (Splitting 1, its t1)
    against this build's t1 (same input, same system-contiguous layout) and
    t2 (interleaved), as speed-ups, at the batch sizes `cols`; then the
    median / min / max speed-up over every point both files hold."""
    import statistics
    ours, ref = read_dat(ours_path), read_dat(ref_path)
    common = sorted(set(ours) & set(ref))
    out = ["| n \\ m | " + " | ".join(f"m={m}" for m in cols) + " |",
           "|---|" + "---|" * len(cols)]
    for n in SIZES:
        cells = []
        for m in cols:
            if (m, n) in ours and (m, n) in ref:
                r, o = ref[(m, n)][0], ours[(m, n)]
                cells.append(f"{r:.3f} / {o[0]:.3f} ms = {r / o[0]:.0f}x")
            elif (m, n) in ours:
                cells.append(f"- / {ours[(m, n)][0]:.3f} ms")
            else:
                cells.append("-")
        out.append(f"| {n} | " + " | ".join(cells) + " |")
    for k, label in ((0, "t1 (contiguous)"), (1, "t2 (interleaved)"), (2, "t3 (FP32+refine)")):
        sp = [ref[p][0] / ours[p][k] for p in common]
        out.append(f"\nSpeed-up of {label} over the reference's Splitting 1 at the {len(sp)} "
                   f"common points: median {statistics.median(sp):.1f}x, min {min(sp):.1f}x, "
                   f"max {max(sp):.1f}x.")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--compare", nargs=2, metavar=("OURS", "REFERENCE"), default=None,
                    help="print a speed-up table of two .dat files and exit")
    ap.add_argument("--out", default=OUT_NAME)
    ap.add_argument("--runs", type=int, default=AVG_RUN)
    ap.add_argument("--batches", type=int, nargs="*", default=None)
    ap.add_argument("--sizes", type=int, nargs="*", default=None)
    args = ap.parse_args(argv)
    if args.compare:
        print(compare(*args.compare))
        return 0
    with open(args.out, "w"):
        pass
    for m, n in grid(args.batches, args.sizes):
        t1, t2, t3 = time_point(m, n, args.runs)
        print(" %d \t %d" % (m, n), flush=True)
        with open(args.out, "a") as f:
            f.write(format_row(m, n, t1, t2, t3))
    return 0


if __name__ == "__main__":
    sys.exit(main())
