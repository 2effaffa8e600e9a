#!/usr/bin/env python3
"""Benchmark of the batched PCR-Penta solve on one MI355X.

Workload (default): m = 16384 systems of n = 1024, interleaved layout, plain
FP64, seeded strictly dominant generator (off-diagonals U[-1,1], main
U[4.5,5.5], RHS U[-1,1], seed 0). One "step" = one solve of the whole batch
(y <- A^{-1} y in place). Every step gets its own pristine right-hand side,
generated into HBM before the timed region, so nothing but solves is timed.

Prints ONE JSON line: unknowns/s (value), ms per step, the dominant kernel's
roofline (56 algorithmic bytes per unknown: a,b,c,d,e,y read + x written, all
FP64) and the CPU baseline (the oracle's FP64 PCR-Penta, OpenMP over systems)
timed on a sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pcr-penta_amd"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_UNKNOWN = 56     # 7 FP64 words

WORKLOADS = {
    # name: (n, m, precision, generator kind); kind 0 = strictly dominant
    # (off-diagonals U[-1,1], diagonal U[4.5,5.5]), kind 1 = weakly dominant
    # (off-diagonals U[-0.5,0.5], diagonal U[2.2,2.6])
    "lds_f64_m16384_n1024": (1024, 16384, "f64", 0),
    "lds_f32r_m16384_n2048": (2048, 16384, "f32r", 0),
    "lds_f32r_m16384_n1024": (1024, 16384, "f32r", 0),
    "hbm_f64_m16384_n2048": (2048, 16384, "f64", 0),
    "lds_f64_m256_n512": (512, 256, "f64", 0),
    "hbm_f64_m32_n1048576": (1 << 20, 32, "f64", 1),
    "hbm_f32r_m32_n1048576": (1 << 20, 32, "f32r", 1),
    # uneven: n = 0 marks the size draw of 10000 systems, n log-uniform in [64, 131072]
    "uneven_f64_m10000": (0, 10000, "f64", 0),
    "uneven_f32r_m10000": (0, 10000, "f32r", 0),
    # Heston ADI (Ns = 1024, Nv = 512): n = -1 marks the ADI driver; one step = one whole
    # 200-step run
    "heston_adi_f64": (-1, 0, "f64", 0),
    "heston_adi_f32r": (-1, 0, "f32r", 0),
}
MAX_SWEEPS, TOL = 8, 1e-12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lds_f64_m16384_n1024", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-systems", type=int, default=2048,
                    help="systems in the CPU-baseline sample")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner; use the built-in default variant")
    ap.add_argument("--tune-save", default=None, help="write the tuning table to this file")
    ap.add_argument("--tune-load", default=None,
                    help="load a tuning table (implies --no-tune)")
    ap.add_argument("--split", default=None, metavar="K_FUSE,SHAPE",
                    help="fix the split path's steps per pass and pass tile shape "
                         "(implies --no-tune for the split path)")
    args = ap.parse_args()
    if args.split:
        kf, sh = (int(v) for v in args.split.split(","))
        args.split = (kf, sh)
        args.no_tune = True
    return args


def apply_split(args):
    if args.split:
        import pcrpenta
        lds_max = pcrpenta.split_tune_get()[0]
        pcrpenta.split_tune_set(lds_max, *args.split)


CPU_SECONDS = 10.0   # CPU-baseline sample: repeat the host solve for about this long


def timed_repeats(fn, seconds=CPU_SECONDS, min_runs=3):
    """Run fn() after one warm-up until `seconds` have passed (at least
    min_runs times); returns (runs, elapsed seconds, last result)."""
    out = fn()
    runs, t0 = 0, time.perf_counter()
    while runs < min_runs or time.perf_counter() - t0 < seconds:
        out = fn()
        runs += 1
    return runs, time.perf_counter() - t0, out


def cpu_baseline(n, m, layout, kind, seed, gpu_x_sample, nsys, prec):
    """The oracle's FP64 PCR-Penta on the host (all threads, OpenMP over
    systems): the first `nsys` systems of the same workload, solved over and
    over for about CPU_SECONDS; also checks the GPU's result on those systems
    (bit-identical for FP64, error against the FP64 banded-LU oracle for
    FP32+refine)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    oracle.load()
    arrs = oracle.gen_batch(kind, seed, n, nsys, oracle.INTERLEAVED)
    threads = oracle.max_threads()
    runs, t, x = timed_repeats(
        lambda: oracle.pcr_batch(*arrs, n, nsys, oracle.INTERLEAVED, threads))
    agree = err = None
    if gpu_x_sample is not None:
        agree = bool(np.array_equal(x.reshape(n, nsys), gpu_x_sample))
        xg = oracle.gbsv_batch(*arrs, n, nsys, oracle.INTERLEAVED)
        err = oracle.normwise_relerr(gpu_x_sample.reshape(-1), xg, n, nsys, oracle.INTERLEAVED)
    return {
        "value": n * nsys * runs / t,
        "unit": "unknowns/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {nsys} of the {m} systems (n={n}, same generator) solved {runs} times "
                  f"in {t:.1f} s, FP64 PCR-Penta restated from kernel.cu:32-154, gcc -O2 + "
                  f"OpenMP over systems",
        "seconds": t,
        "bit_identical_to_gpu": agree,
        "gpu_relerr_vs_dgbsv": err,
    }


def measured_traffic(workload):
    """HBM bytes of one solve (every solver kernel of one step) measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE and calibrated against a plain copy;
    None for workloads not profiled."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        return json.load(f).get(workload, {}).get("bytes_per_step")


def bench_uneven(args, m, prec, kind):
    """One uneven batch (contiguous systems back to back, n log-uniform in [64, 131072]),
    seed 1."""
    import numpy as np
    import torch
    import pcrpenta
    seed = 1
    K, W = args.steps, args.warmup
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    sizes = pcrpenta.uneven_log_uniform_sizes(m, seed)
    off, nn, total = pcrpenta.uneven_layout(sizes)
    a, b, c, d, e, y0 = pcrpenta.generate_uneven(kind, seed, off, nn, device=dev)
    nbuf = max(1, min(K + W, int(8e9 // (8 * total))))
    ys = [y0.clone() for _ in range(nbuf)]
    y_check = y0.clone()
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    tuned = None
    refined = prec != "f64"
    sweeps = torch.zeros(m, dtype=torch.int32, device=dev) if refined else None
    if not args.no_tune and not refined:
        # autotune steps fused per pass and the later rounds' tile shape on this batch
        lds_max, _, _, n_shapes = pcrpenta.split_tune_get()
        best = None
        ys_t = y0.clone()
        for kf in (2, 3, 4, 5):
            for sh in range(n_shapes):
                try:
                    pcrpenta.split_tune_set(lds_max, kf, sh)
                except pcrpenta.PCRPError:
                    continue
                nb = pcrpenta.uneven_workspace_size(off, nn)
                ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
                ys_t.copy_(y0)
                pcrpenta.solve_uneven(a, b, c, d, e, ys_t, off, nn, stream=stream, ws=ws)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                ys_t.copy_(y0)
                e0.record(stream)
                pcrpenta.solve_uneven(a, b, c, d, e, ys_t, off, nn, stream=stream, ws=ws)
                e1.record(stream)
                torch.cuda.synchronize()
                t = e0.elapsed_time(e1)
                if best is None or t < best[0]:
                    best = (t, kf, sh)
                del ws
        pcrpenta.split_tune_set(lds_max, best[1], best[2])
        tuned = {"k_fuse": best[1], "shape": best[2], "ms": best[0]}
        print(json.dumps({"split_autotune": tuned}), file=sys.stderr, flush=True)
        del ys_t
    nb = pcrpenta.uneven_workspace_size(
        off, nn, pcrpenta.PREC_F32_REFINE if refined else pcrpenta.PREC_F64)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def step(k, y=None):
        y = ys[k % nbuf] if y is None else y
        if refined:
            return pcrpenta.solve_uneven_refined(a, b, c, d, e, y, off, nn, max_sweeps=MAX_SWEEPS,
                                                 tol=TOL, sweeps=sweeps, stream=stream, ws=ws)
        return pcrpenta.solve_uneven(a, b, c, d, e, y, off, nn, stream=stream, ws=ws)

    for w in range(W):
        step(w)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(K)]
    t0 = time.perf_counter()
    for k in range(K):
        evs[k][0].record(stream)
        step(W + k)
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    solve_s = sum(s_.elapsed_time(f) for s_, f in evs) / K / 1e3
    step(0, y_check)
    torch.cuda.synchronize()
    achieved = BYTES_PER_UNKNOWN * total / solve_s
    out = {
        "metric": "unknowns solved/s + % HBM roofline; FP32+refine vs FP64 at rel err <=1e-12",
        "value": total / (wall / K), "unit": "unknowns/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": prec,
        "data": "synthetic (seeded counter-based generator, strictly dominant distribution, 10000 "
                "sizes log-uniform in [64, 131072])",
        "config": {"workload": args.workload, "m": m, "unknowns": total, "n_min": int(nn.min()),
                   "n_max": int(nn.max()), "large_systems": int((nn > 1024).sum()),
                   "layout": "contiguous per system", "precision": prec, "seed": seed,
                   "split": None if refined else list(pcrpenta.split_tune_get()[:3]),
                   "refine": ({"max_sweeps": MAX_SWEEPS, "tol": TOL,
                               "sweeps_hist": torch.bincount(sweeps.long()).tolist()}
                              if refined else None)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(args.workload),
                     "solve_ms": solve_s * 1e3,
                     "note": ("whole solve per step: three window-path refinement sweeps "
                              "(one kernel each) and the final copy" if refined else
                              "whole solve (pass rounds + persistent LDS launch) per step")},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.load()
        # sample: the first systems of the draw up to ~4e6 unknowns
        cnt = int(np.searchsorted(np.cumsum(nn.astype(np.int64)), 4_000_000)) + 1
        cnt = min(cnt, m)
        sub_total = int(off[cnt - 1] + nn[cnt - 1])
        ins = [t[:sub_total].cpu().numpy() for t in (a, b, c, d, e, y0)]
        runs, tc, x = timed_repeats(lambda: oracle.pcr_uneven(*ins, off[:cnt], nn[:cnt]))
        xg = y_check[:sub_total].cpu().numpy()
        check = {"bit_identical_to_gpu": bool(np.array_equal(x, xg))}
        if refined:
            errs = []
            for o, nj in zip(off[:cnt], nn[:cnt]):
                sl = slice(int(o), int(o) + int(nj))
                ref = oracle.gbsv(*[t[sl] for t in ins])
                errs.append(float(np.abs(xg[sl] - ref).max() / np.abs(ref).max()))
            check = {"gpu_max_relerr_vs_dgbsv": max(errs)}
        out["cpu_baseline"] = {
            "value": sub_total * runs / tc, "unit": "unknowns/s", "cores": oracle.max_threads(),
            "kind": "port",
            "sample": f"first {cnt} systems of the draw ({sub_total} unknowns) solved {runs} "
                      f"times in {tc:.1f} s, FP64 PCR-Penta restated from kernel.cu:32-154, "
                      f"OpenMP dynamic schedule over systems",
            "seconds": tc, **check,
        }
    print(json.dumps(out), flush=True)


def bench_heston(args, prec):
    """Heston ADI, Douglas theta = 1/2, Ns = 1024, Nv = 512, 200 steps,
    each step replayed from one captured hipGraph. Unknowns per step: Nv S-lines
    of 1024 plus Ns v-lines of 512 = 2 * 524288."""
    import torch
    import pcrpenta
    pid = pcrpenta.PREC_F64 if prec == "f64" else pcrpenta.PREC_F32_REFINE
    p = pcrpenta.heston_params()
    K, W = args.steps, args.warmup
    s = torch.cuda.Stream()
    u = torch.empty(p.ns * p.nv, dtype=torch.float64, device="cuda")
    ws = torch.empty(int(pcrpenta.load().pcrp_heston_workspace_size(
        pcrpenta.ctypes.byref(p), pid)), dtype=torch.uint8, device="cuda")
    with torch.cuda.stream(s):
        for _ in range(W):
            pcrpenta.heston(p, pid, True, stream=s, ws=ws, u=u)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            pcrpenta.heston(p, pid, True, stream=s, ws=ws, u=u)
        s.synchronize()
        wall = time.perf_counter() - t0
    unknowns = 2 * p.ns * p.nv * p.steps
    hS, hv = p.smax / (p.ns + 1), p.vmax / p.nv
    i0, j0 = int(round(100.0 / hS)), int(round(0.04 / hv))
    out = {
        "metric": "unknowns solved/s + % HBM roofline; FP32+refine vs FP64 at rel err <=1e-12",
        "value": unknowns / (wall / K), "unit": "unknowns/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": prec,
        "data": "synthetic (Heston European call, Ns = 1024, Nv = 512, 200 steps of 1/200)",
        "config": {"workload": args.workload, "ns": p.ns, "nv": p.nv, "time_steps": p.steps,
                   "dt": p.dt, "scheme": "Douglas theta=0.5", "graph": "one hipGraph per time "
                   "step, replayed", "unknowns_per_run": unknowns,
                   "price_near_S100_v0.04": float(u.view(p.nv, p.ns)[j0, i0 - 1].item())},
        "roofline": {"bound": "hbm", "achieved": 56.0 * unknowns / (wall / K) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": 56.0 * unknowns / (wall / K) / HBM_PEAK, "traffic": None,
                     "note": "one solve is 29 MB (fits the 256 MB Infinity Cache); whole ADI "
                             "run incl. explicit stencil kernels"},
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(systems are independent: shard the batch across GPUs, "
                                   "one process each)"}))
        sys.exit(2)
    import torch
    import pcrpenta

    n, m, prec, kind = WORKLOADS[args.workload]
    apply_split(args)
    if n == 0:
        return bench_uneven(args, m, prec, kind)
    if n == -1:
        return bench_heston(args, prec)
    layout = pcrpenta.INTERLEAVED
    seed = 0
    K, W = args.steps, args.warmup
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    a, b, c, d, e, y0 = pcrpenta.generate(kind, seed, n, m, layout, device=dev)
    nbuf = max(1, min(K + W, int(8e9 // (8 * n * m))))
    ys = [y0.clone() for _ in range(nbuf)]
    y_check = y0.clone()
    pid = pcrpenta.PREC_F64 if prec == "f64" else pcrpenta.PREC_F32_REFINE
    ws, nb = pcrpenta._workspace(n, m, pid, dev)
    sweeps = torch.zeros(m, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    tuned = None
    if args.tune_load:
        pcrpenta.tune_load(args.tune_load)
    elif not args.no_tune and prec == "f64" and n > 1024:
        # split path: the autotuner picks steps fused per pass and the pass tile shape
        kf, sh, ms = pcrpenta.split_autotune(a, b, c, d, e, y0, n, m, layout, reps=2, ws=ws)
        tuned = {"split": [kf, sh], "ms": ms}
        print(json.dumps({"split_autotune": tuned}), file=sys.stderr, flush=True)
        if args.tune_save:
            pcrpenta.tune_save(args.tune_save)
    elif not args.no_tune and prec == "f64" and n <= 1024:
        # the autotuner times every LDS variant on this very batch (outside the timed region)
        best, best_ms, times = pcrpenta.autotune(a, b, c, d, e, y0, n, m, layout, reps=3)
        vs = pcrpenta.variants()
        tuned = {"variant": vs[best], "ms": best_ms,
                 "all_ms": {f"{v['kind']}-{v['nt']}x{v['rpt']}": round(t, 4)
                            for v, t in zip(vs, times) if t >= 0}}
        print(json.dumps({"autotune": tuned}), file=sys.stderr, flush=True)
        if args.tune_save:
            pcrpenta.tune_save(args.tune_save)
    vid = pcrpenta.tune_get(min(n, 1024)) if prec == "f64" else -1
    variant = pcrpenta.variants()[vid] if vid >= 0 else None

    def step(k, y=None):
        y = ys[k % nbuf] if y is None else y
        if prec == "f64":
            return pcrpenta.solve(a, b, c, d, e, y, n, m, layout, stream=stream, ws=ws)
        return pcrpenta.solve_refined(a, b, c, d, e, y, n, m, layout, max_sweeps=MAX_SWEEPS,
                                      tol=TOL, sweeps=sweeps, stream=stream, ws=ws)

    for w in range(W):
        step(w)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(K)]
    t0 = time.perf_counter()
    for k in range(K):
        evs[k][0].record(stream)
        step(W + k)
        evs[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(f) for s, f in evs]
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    ms_per_step = wall * 1e3 / K
    unknowns = n * m
    value = unknowns / (wall / K)

    # one more solve on a pristine RHS for the CPU-baseline agreement check
    step(0, y_check)
    torch.cuda.synchronize()
    sweep_hist = None
    if prec != "f64":
        sweep_hist = torch.bincount(sweeps.long()).tolist()

    achieved = BYTES_PER_UNKNOWN * unknowns / avg_kernel_s
    traffic = measured_traffic(args.workload)
    out = {
        "metric": "unknowns solved/s + % HBM roofline; FP32+refine vs FP64 at rel err <=1e-12",
        "value": value,
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic (seeded counter-based generator, "
                + ("strictly dominant distribution)" if kind == 0 else "weakly dominant distribution)"),
        "config": {"workload": args.workload, "n": n, "m": m, "layout": "interleaved",
                   "precision": prec, "seed": seed,
                   "generator": "dominant" if kind == 0 else "weak",
                   "kernel_variant": (f"{variant['kind']}-{variant['nt']}x{variant['rpt']}"
                                      if variant else None),
                   "split": list(pcrpenta.split_tune_get()[:3]) if n > 1024 else None,
                   "refine": ({"max_sweeps": MAX_SWEEPS, "tol": TOL, "sweeps_hist": sweep_hist}
                              if prec != "f64" else None)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel_ms": avg_kernel_s * 1e3},
    }
    if not args.no_cpu:
        nsys = min(args.cpu_systems, m)
        xs = y_check.view(n, m)[:, :nsys].cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(n, m, layout, kind, seed, xs, nsys, prec)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
