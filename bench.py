#!/usr/bin/env python3
"""Benchmark of the batched PCR-Penta solve on one MI355X.

Default workload: the uneven batch, 10000 systems with n drawn log-uniform in
[64, 131072] (seed 1, about 1.7e8 unknowns), strictly dominant generator
(off-diagonals U[-1,1], main U[4.5,5.5], RHS U[-1,1]), systems stored back to
back. One "step" = one solve of the whole batch (y <- A^{-1} y in place).
`value` is plain FP64 PCR-Penta (the reference's precision); the same line
carries the FP32 PCR + FP64 refinement form on the same batch (its unknowns/s,
its ratio to FP64, its sweep counts), each form's error against the FP64
banded-LU (dgbsv) oracle over every system, each form's HBM roofline (56
algorithmic bytes per unknown: a, b, c, d, e, y read + x written, all FP64),
and the CPU baseline (the oracle's FP64 PCR-Penta, OpenMP over systems) on the
whole batch on this host. Every timed step gets its own pristine right-hand
side, generated into HBM before the timed region, so nothing but solves is
timed. `also` adds the other workloads (Heston ADI, the n = 1024 / 2048
batches, n = 2^20) timed the same way in both forms.

--workload NAME runs one form of one workload (the profiles use these).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pcr-penta_amd"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_UNKNOWN = 56     # 7 FP64 words
VALU_PEAK = 256 * 2.4e9    # wave64 VALU instructions/s: 1 per CU per cycle
METRIC = "unknowns solved/s + % HBM roofline; FP32+refine vs FP64 at rel err <=1e-12"
HEADLINE = "uneven_m10000"

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
    # n = -2: the ADI driver with the line matrices factored once per run
    "heston_adi_f64_reuse": (-2, 0, "f64", 0),
    "heston_adi_f32r_reuse": (-2, 0, "f32r", 0),
}
MAX_SWEEPS, TOL = 8, 1e-12
RHS_BYTES_CAP = 64e9       # pristine right-hand sides kept in HBM for the timed steps


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=HEADLINE, choices=sorted(WORKLOADS) + [HEADLINE])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and checks")
    ap.add_argument("--no-also", action="store_true", help="headline: skip the other workloads")
    ap.add_argument("--cpu-systems", type=int, default=2048,
                    help="uniform workloads: systems in the CPU-baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=5, help="CPU baseline: timed runs (median)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner; use the built-in default variant")
    ap.add_argument("--tune-save", default=None, help="write the tuning table to this file")
    ap.add_argument("--tune-load", default=None,
                    help="load a tuning table (implies --no-tune)")
    ap.add_argument("--split", default=None, metavar="[LDS_MAX,]K_FUSE,SHAPE",
                    help="fix the split path's steps per pass and pass tile shape, and "
                         "optionally its LDS/HBM threshold (implies --no-tune for the split path)")
    ap.add_argument("--win-levels", type=int, default=None, metavar="L",
                    help="fix the FP32 refinement's window level count (2..7)")
    args = ap.parse_args()
    if args.split:
        vals = tuple(int(v) for v in args.split.split(","))
        if len(vals) not in (2, 3):
            ap.error("--split takes K_FUSE,SHAPE or LDS_MAX,K_FUSE,SHAPE")
        args.split = vals
        args.no_tune = True
    if args.tune_load:
        args.no_tune = True
    return args


def log(obj):
    print(json.dumps(obj), file=sys.stderr, flush=True)


# ---------------------------------------------------------------- timing

def time_steps(step, K, W):
    """W untimed steps, then K timed steps between two device synchronisations,
    each bracketed by HIP events on the solve stream. Returns (wall seconds of the
    K steps, mean event-timed seconds per step)."""
    import torch
    stream = torch.cuda.current_stream()
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
    ev = sum(s.elapsed_time(f) for s, f in evs) / K / 1e3
    return wall, ev


def measured_traffic(workload):
    """HBM bytes of one solve (every solver kernel of one step) measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE and calibrated against a plain copy;
    None for workloads not profiled."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        return json.load(f).get(workload, {}).get("bytes_per_step")


def _traffic_entry(workload):
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(tpath):
        return {}
    with open(tpath) as f:
        return json.load(f).get(workload, {})


def measured_valu(workload):
    """VALU wave-instructions of one solve (SQ_INSTS_VALU summed over every
    solver kernel of one step, rocprofv3); None for workloads not profiled."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        return json.load(f).get(workload, {}).get("valu_insts_per_step")


def rocprof_summary(workload, solve_s):
    """This workload's rocprofv3 figures (profiles/hbm_traffic.json): HBM bytes
    of one solve over this run's solve time, and LDS bank-conflict cycles
    (SQ_LDS_BANK_CONFLICT summed over the solver kernels of one solve)."""
    e = _traffic_entry(workload)
    if not e.get("bytes_per_step"):
        return None
    return {"hbm_gb_per_solve": e["bytes_per_step"] / 1e9,
            "hbm_gb_per_s": e["bytes_per_step"] / solve_s / 1e9,
            "lds_bank_conflict_cycles_per_solve": e.get("lds_bank_conflicts_per_step"),
            "profile": e.get("profile")}


def roofline(unknowns, solve_s, workload, note):
    achieved = BYTES_PER_UNKNOWN * unknowns / solve_s
    out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": achieved / HBM_PEAK, "traffic": measured_traffic(workload),
           "solve_ms": solve_s * 1e3, "note": note}
    dom = _traffic_entry(workload).get("dominant_kernel")
    if dom:
        # the kernel moving the most HBM bytes of one solve, from its rocprofv3
        # summary (profiles/): its average duration and share of the solve
        out["dominant_kernel"] = {"name": dom["name"], "profiled_avg_ms": dom["avg_us"] / 1e3,
                                  "hbm_bytes_per_dispatch": dom["hbm_bytes_per_dispatch"],
                                  "dispatches_per_solve": dom["dispatches_per_step"],
                                  "share_of_solve": dom["avg_us"] * dom["dispatches_per_step"]
                                  / 1e3 / (solve_s * 1e3)}
    rp = rocprof_summary(workload, solve_s)
    if rp:
        out["rocprof"] = rp
    insts = measured_valu(workload)
    if insts:
        # issue bound: one wave64 VALU instruction per CU per cycle (4 SIMDs,
        # 16 lanes each), 256 CUs at 2.4 GHz; FP64 FMA issues at that rate
        out["valu"] = {"insts_per_solve": insts, "achieved": insts / solve_s / 1e9,
                       "peak": VALU_PEAK / 1e9, "unit": "G wave-instructions/s",
                       "frac": insts / solve_s / VALU_PEAK}
    return out


def apply_split(args):
    import pcrpenta
    if args.split:
        if len(args.split) == 3:
            pcrpenta.split_tune_set(*args.split)
        else:
            pcrpenta.split_tune_set(pcrpenta.split_tune_get()[0], *args.split)
    if args.win_levels is not None:
        pcrpenta.f32r_tune_set(args.win_levels, pcrpenta.f32r_tune_get()[1])


# ---------------------------------------------------------------- CPU side

def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


_NATIVE = {}


def oracle_native():
    """The oracle built for this host (gcc -O3 -march=native + OpenMP, threads
    pinned one per core), loaded once; it is the CPU baseline and the checker."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if "path" not in _NATIVE:
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")
        _NATIVE["dir"] = tempfile.mkdtemp(prefix="pcrp_oracle_")
        _NATIVE["path"] = oracle.load_native(_NATIVE["dir"])
    return oracle


def median_runs(fn, runs, min_seconds=10.0):
    """One warm-up call, then at least `runs` timed calls and at least
    min_seconds in total; returns (median seconds, number of runs, last result)."""
    out = fn()
    times = []
    t_start = time.perf_counter()
    while len(times) < runs or time.perf_counter() - t_start < min_seconds:
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), len(times), out


def cpu_fields(oracle, runs, tmed):
    return {"cores": oracle.max_threads(), "kind": "port", "cpu_model": cpu_info(),
            "pinning": f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} "
                       f"OMP_PLACES={os.environ.get('OMP_PLACES')}",
            "build": "gcc -O3 -march=native -ffp-contract=off -fopenmp, built on this host",
            "runs": runs, "median_s": tmed}


# ---------------------------------------------------------------- uneven batch

def uneven_batch(kind=0, seed=1, m=10000):
    import torch
    import pcrpenta
    sizes = pcrpenta.uneven_log_uniform_sizes(m, seed)
    off, nn, total = pcrpenta.uneven_layout(sizes)
    dev = torch.device("cuda:0")
    arrs = pcrpenta.generate_uneven(kind, seed, off, nn, device=dev)
    return {"sizes": sizes, "off": off, "nn": nn, "total": total, "arrs": arrs, "m": m,
            "seed": seed, "kind": kind}


def tune_uneven_split(B):
    """Autotune the split path's steps fused per pass and tile shape on this batch."""
    import torch
    import pcrpenta
    a, b, c, d, e, y0 = B["arrs"]
    off, nn = B["off"], B["nn"]
    stream = torch.cuda.current_stream()
    _, _, _, n_shapes = pcrpenta.split_tune_get()
    best = None
    ys_t = y0.clone()
    for lds_max, kf, sh in ((lm, kf, sh) for lm in (512, 1024) for kf in (2, 3, 4, 5)
                            for sh in range(n_shapes)):
        try:
            pcrpenta.split_tune_set(lds_max, kf, sh)
        except pcrpenta.PCRPError:
            continue
        nb = pcrpenta.uneven_workspace_size(off, nn)
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=y0.device)
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
            best = (t, kf, sh, lds_max)
        del ws
    pcrpenta.split_tune_set(best[3], best[1], best[2])
    return {"lds_max": best[3], "k_fuse": best[1], "shape": best[2], "ms": best[0]}


def tune_uneven_f32r(B):
    """Window level count for the FP32 refinement of this batch (halo 2^L - 1
    block rows): the fastest of L = 4, 5, 6 on one solve each."""
    import torch
    import pcrpenta
    a, b, c, d, e, y0 = B["arrs"]
    off, nn = B["off"], B["nn"]
    keep = pcrpenta.f32r_tune_get()
    nb = pcrpenta.uneven_workspace_size(off, nn, pcrpenta.PREC_F32_REFINE)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=y0.device)
    y = y0.clone()
    stream = torch.cuda.current_stream()
    best = None
    for wl in (4, 5, 6):
        pcrpenta.f32r_tune_set(wl, keep[1])
        y.copy_(y0)
        pcrpenta.solve_uneven_refined(a, b, c, d, e, y, off, nn, stream=stream, ws=ws)
        y.copy_(y0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        pcrpenta.solve_uneven_refined(a, b, c, d, e, y, off, nn, stream=stream, ws=ws)
        e1.record(stream)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1)
        if best is None or t < best[0]:
            best = (t, wl)
    pcrpenta.f32r_tune_set(best[1], keep[1])
    return {"win_levels": best[1], "ms": best[0]}


def plan_execute_ms(B, prec, K, win_levels=None, keep_x=False):
    """Mean event-timed execute of a plan (pcrp_plan_create once: routing
    tables, window levels; executes launch kernels only) over K pristine
    right-hand sides. win_levels: FP32 window level count for this plan
    (2 makes the config generators' batches non-local, i.e. routes them to the
    pass/stage path that systems without decaying couplings take)."""
    import torch
    import pcrpenta
    a, b, c, d, e, y0 = B["arrs"]
    pid = pcrpenta.PREC_F64 if prec == "f64" else pcrpenta.PREC_F32_REFINE
    keep = pcrpenta.f32r_tune_get()
    if win_levels is not None:
        pcrpenta.f32r_tune_set(win_levels, keep[1])
    if "off" in B:
        plan = pcrpenta.Plan(a, b, c, d, e, precision=pid, offsets=B["off"], sizes=B["nn"])
    else:
        plan = pcrpenta.Plan(a, b, c, d, e, B["n"], B["m"], B["layout"], precision=pid)
    pcrpenta.f32r_tune_set(*keep)   # the plan keeps what it was created with
    ys = [y0.clone() for _ in range(2)]
    m = len(B["nn"]) if "off" in B else B["m"]
    sweeps = torch.zeros(m, dtype=torch.int32, device=y0.device)
    stream = torch.cuda.current_stream()
    plan.execute(ys[0], max_sweeps=MAX_SWEEPS, tol=TOL, sweeps=sweeps)
    x = ys[0].clone() if keep_x else None
    ys[0].copy_(y0)
    torch.cuda.synchronize()
    tot = 0.0
    for k in range(K):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        plan.execute(ys[k % 2], max_sweeps=MAX_SWEEPS, tol=TOL)
        e1.record(stream)
        ys[k % 2].copy_(y0)
        torch.cuda.synchronize()
        tot += e0.elapsed_time(e1)
    route = plan.route
    plan.destroy()
    out = {"ms": tot / K, "route": ["lds", "split", "window", "uneven-queue", "uneven-stage"][route]}
    if prec != "f64":
        out["sweeps_hist"] = torch.bincount(sweeps.long()).tolist()
    if keep_x:
        out["x"] = x
    return out


def run_uneven(B, prec, K, W):
    """K timed solves of the uneven batch in one precision form; returns the
    timings, the sweep histogram and x of one more solve (device tensor)."""
    import torch
    import pcrpenta
    a, b, c, d, e, y0 = B["arrs"]
    off, nn, total, m = B["off"], B["nn"], B["total"], B["m"]
    refined = prec != "f64"
    nbuf = max(1, min(K + W, int(RHS_BYTES_CAP // (8 * total))))
    ys = [y0.clone() for _ in range(nbuf)]
    sweeps = torch.zeros(m, dtype=torch.int32, device=y0.device) if refined else None
    nb = pcrpenta.uneven_workspace_size(
        off, nn, pcrpenta.PREC_F32_REFINE if refined else pcrpenta.PREC_F64)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=y0.device)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()

    def step(k, y=None):
        y = ys[k % nbuf] if y is None else y
        if refined:
            return pcrpenta.solve_uneven_refined(a, b, c, d, e, y, off, nn, max_sweeps=MAX_SWEEPS,
                                                 tol=TOL, sweeps=sweeps, stream=stream, ws=ws)
        return pcrpenta.solve_uneven(a, b, c, d, e, y, off, nn, stream=stream, ws=ws)

    wall, solve_s = time_steps(step, K, W)
    y_check = y0.clone()
    step(0, y_check)
    torch.cuda.synchronize()
    del ys
    out = {"value": total / (wall / K), "ms_per_step": wall * 1e3 / K, "solve_s": solve_s,
           "x": y_check}
    if refined:
        out["sweeps_hist"] = torch.bincount(sweeps.long()).tolist()
    return out


def bench_headline(args):
    """Uneven batch, both forms in one line (value: plain FP64)."""
    import numpy as np
    import torch
    import pcrpenta
    K, W = args.steps, args.warmup
    B = uneven_batch()
    off, nn, total, m = B["off"], B["nn"], B["total"], B["m"]
    tuned = tuned32 = None
    if not args.no_tune:
        tuned = tune_uneven_split(B)
        log({"split_autotune": tuned})
        tuned32 = tune_uneven_f32r(B)
        log({"f32r_autotune": tuned32})
    f64 = run_uneven(B, "f64", K, W)
    f32 = run_uneven(B, "f32r", K, W)
    plan32 = plan_execute_ms(B, "f32r", 5)
    stage32 = plan_execute_ms(B, "f32r", 3, win_levels=2, keep_x=not args.no_cpu)
    x_stage = stage32.pop("x", None)
    ratio = f32["value"] / f64["value"]
    out = {
        "metric": METRIC, "value": f64["value"], "unit": "unknowns/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": f64["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded counter-based generator, strictly dominant distribution, "
                "10000 sizes log-uniform in [64, 131072], seed 1)",
        "config": {"workload": HEADLINE, "m": m, "unknowns": total, "n_min": int(nn.min()),
                   "n_max": int(nn.max()), "large_systems": int((nn > 1024).sum()),
                   "layout": "contiguous per system (back to back)", "seed": 1,
                   "value_precision": "f64 (plain PCR-Penta)",
                   "split": list(pcrpenta.split_tune_get()[:3]), "split_autotune": tuned},
        "roofline": roofline(total, f64["solve_s"], "uneven_f64_m10000",
                             "plain FP64, whole solve per step (HBM pass rounds + one persistent "
                             "work-queue LDS launch), HIP events on the solve stream"),
        "f32r": {
            "value": f32["value"], "unit": "unknowns/s", "ms_per_step": f32["ms_per_step"],
            "ratio_vs_f64": ratio, "dtype": "f32 inner solve, f64 residual and update",
            "refine": {"max_sweeps": MAX_SWEEPS, "tol": TOL, "sweeps_hist": f32["sweeps_hist"],
                       "win_levels": pcrpenta.f32r_tune_get()[0], "autotune": tuned32},
            "roofline": roofline(total, f32["solve_s"], "uneven_f32r_m10000",
                                 "FP32+refine, whole solve per step (window-path sweeps, one "
                                 "kernel each, and the final copy)"),
            "plan_execute": {**plan32,
                             "unknowns_per_s": total / (plan32["ms"] / 1e3),
                             "note": "repeat solves of the same matrices: pcrp_plan_create "
                                     "once, then capturable executes"},
            "stage_path": {**stage32,
                           "ratio_vs_f64": f64["ms_per_step"] / stage32["ms"],
                           "note": "the pass/stage path that batches whose couplings do not "
                                   "decay take (forced here with a 2-level window plan)"},
        },
    }
    if not args.no_cpu:
        oracle = oracle_native()
        ins = [t.cpu().numpy() for t in B["arrs"]]
        x64 = f64["x"].cpu().numpy()
        x32 = f32["x"].cpu().numpy()
        ref = oracle.gbsv_uneven(*ins, off, nn)
        e64 = oracle.uneven_relerr(x64, ref, off, nn)
        e32 = oracle.uneven_relerr(x32, ref, off, nn)
        out["accuracy"] = {
            "systems": f"all {m} systems, normwise max-norm relative error vs dgbsv "
                       "(kl = ku = 2, partial pivoting)",
            "f64_max_relerr": float(e64.max()), "f64_systems_above_1e-12": int((e64 > 1e-12).sum()),
            "f32r_max_relerr": float(e32.max()),
            "f32r_systems_above_1e-12": int((e32 > 1e-12).sum()),
        }
        out["f32r"]["max_relerr_vs_dgbsv"] = float(e32.max())
        if x_stage is not None:
            es = oracle.uneven_relerr(x_stage.cpu().numpy(), ref, off, nn)
            out["f32r"]["stage_path"]["max_relerr_vs_dgbsv"] = float(es.max())
            out["accuracy"]["f32r_stage_path_max_relerr"] = float(es.max())
            del x_stage
        tmed, runs, xc = median_runs(lambda: oracle.pcr_uneven(*ins, off, nn), args.cpu_runs)
        out["cpu_baseline"] = {
            "value": total / tmed, "unit": "unknowns/s",
            "sample": f"the whole batch ({m} systems, {total} unknowns), FP64 PCR-Penta "
                      f"restated from kernel.cu:32-154, OpenMP dynamic schedule over systems, "
                      f"median of {runs} runs",
            **cpu_fields(oracle, runs, tmed),
            "bit_identical_to_gpu_f64": bool(np.array_equal(xc, x64)),
        }
        del ins
    del f64["x"], f32["x"]
    if not args.no_also:
        out["also"] = also_workloads(args)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------- uniform batches

def run_uniform(n, m, prec, kind, K, W, tune=True, layout=None):
    """K timed solves of a uniform batch; returns timings (and x of one more
    solve on a pristine right-hand side, as a device tensor)."""
    import torch
    import pcrpenta
    layout = pcrpenta.INTERLEAVED if layout is None else layout
    dev = torch.device("cuda:0")
    a, b, c, d, e, y0 = pcrpenta.generate(kind, 0, n, m, layout, device=dev)
    pid = pcrpenta.PREC_F64 if prec == "f64" else pcrpenta.PREC_F32_REFINE
    ws, nb = pcrpenta._workspace(n, m, pid, dev)
    tuned = None
    if tune and prec == "f64" and n > 1024:
        kf, sh, ms = pcrpenta.split_autotune(a, b, c, d, e, y0, n, m, layout, reps=2, ws=ws)
        tuned = {"split": [kf, sh], "ms": ms}
    elif tune and prec == "f64":
        best, best_ms, times = pcrpenta.autotune(a, b, c, d, e, y0, n, m, layout, reps=3)
        vs = pcrpenta.variants()
        tuned = {"variant": vs[best], "ms": best_ms,
                 "all_ms": {f"{v['kind']}-{v['nt']}x{v['rpt']}": round(t, 4)
                            for v, t in zip(vs, times) if t >= 0}}
    nbuf = max(1, min(K + W, int(RHS_BYTES_CAP // (8 * n * m))))
    ys = [y0.clone() for _ in range(nbuf)]
    sweeps = torch.zeros(m, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()

    def step(k, y=None):
        y = ys[k % nbuf] if y is None else y
        if prec == "f64":
            return pcrpenta.solve(a, b, c, d, e, y, n, m, layout, stream=stream, ws=ws)
        return pcrpenta.solve_refined(a, b, c, d, e, y, n, m, layout, max_sweeps=MAX_SWEEPS,
                                      tol=TOL, sweeps=sweeps, stream=stream, ws=ws)

    wall, solve_s = time_steps(step, K, W)
    del ys
    y_check = y0.clone()
    step(0, y_check)
    torch.cuda.synchronize()
    out = {"value": n * m / (wall / K), "ms_per_step": wall * 1e3 / K, "solve_s": solve_s,
           "tuned": tuned, "x": y_check, "layout": layout,
           "batch": {"arrs": (a, b, c, d, e, y0), "n": n, "m": m, "layout": layout}}
    if prec != "f64":
        out["sweeps_hist"] = torch.bincount(sweeps.long()).tolist()
    return out


def run_heston(prec, K, W, reuse=False, separate=False):
    """K timed Heston ADI runs (200 Douglas steps each, captured 20 steps per
    hipGraph); reuse: the line matrices factored once per run."""
    import torch
    import pcrpenta
    pid = pcrpenta.PREC_F64 if prec == "f64" else pcrpenta.PREC_F32_REFINE
    p = pcrpenta.heston_params(reuse_factor=1 if reuse else 0,
                               separate_explicit=1 if separate else 0)
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
    return {"value": unknowns / (wall / K), "ms_per_run": wall * 1e3 / K, "u": u, "p": p,
            "unknowns": unknowns}


def also_workloads(args):
    """The other workloads, both forms each, fewer steps (timing only: the GPU
    tests hold their parity checks)."""
    import numpy as np
    K, W = 10, 3
    out = {}
    h64 = run_heston("f64", 5, 2)
    h32 = run_heston("f32r", 5, 2)
    h64r = run_heston("f64", 5, 2, reuse=True)
    h32r = run_heston("f32r", 5, 2, reuse=True)
    s64 = run_heston("f64", 5, 2, separate=True)
    s32 = run_heston("f32r", 5, 2, separate=True)
    p = h64["p"]
    hS, hv = p.smax / (p.ns + 1), p.vmax / p.nv
    U64 = h64["u"].view(p.nv, p.ns).cpu().numpy()
    U32 = h32["u"].view(p.nv, p.ns).cpu().numpy()
    out["heston_adi"] = {
        "grid": [p.ns, p.nv], "time_steps": p.steps,
        "graph": "20 Douglas steps per hipGraph, replayed",
        "stage": "explicit stage computed by the S-line solve kernels (two launches per step)",
        "f64_ms_per_run": h64["ms_per_run"], "f32r_ms_per_run": h32["ms_per_run"],
        "f64_unknowns_per_s": h64["value"], "f32r_unknowns_per_s": h32["value"],
        "ratio_f32r_vs_f64": h64["ms_per_run"] / h32["ms_per_run"],
        "factor_reuse": {
            "f64_ms_per_run": h64r["ms_per_run"], "f32r_ms_per_run": h32r["ms_per_run"],
            "ratio_f32r_reuse_vs_plain_f64": h64["ms_per_run"] / h32r["ms_per_run"],
            "ratio_f32r_reuse_vs_f64_reuse": h64r["ms_per_run"] / h32r["ms_per_run"],
            "note": "the shared v-line matrix factored once per run, in both precisions "
                    "(bit-identical results to the full solves)"},
        "separate_explicit": {
            "f64_ms_per_run": s64["ms_per_run"], "f32r_ms_per_run": s32["ms_per_run"],
            "ratio_f32r_vs_f64": s64["ms_per_run"] / s32["ms_per_run"],
            "note": "a separate explicit-stage kernel before each S-line solve (three launches "
                    "per step); the same bits as the folded stage"},
        "rocprof": {f: rocprof_summary("heston_adi_" + f, r["ms_per_run"] / 1e3)
                    for f, r in (("f64", h64), ("f32r", h32), ("f64_reuse", h64r),
                                 ("f32r_reuse", h32r))},
        "f32r_vs_f64_grid_relerr": float(np.abs(U32 - U64).max() / np.abs(U64).max()),
        "price_at_S100_v0.04_f64": float(_interp_price(U64, p, hS, hv)),
        "refine": {"tol": 1e-12, "max_sweeps": 8},
    }
    for label, n, m, kind, wl in (
            ("lds_n1024_m16384", 1024, 16384, 0, ("lds_f64_m16384_n1024", "lds_f32r_m16384_n1024")),
            ("n2048_m16384", 2048, 16384, 0, ("hbm_f64_m16384_n2048", "lds_f32r_m16384_n2048")),
            ("hbm_n1048576_m32", 1 << 20, 32, 1, ("hbm_f64_m32_n1048576", "hbm_f32r_m32_n1048576"))):
        r64 = run_uniform(n, m, "f64", kind, K, W)
        r32 = run_uniform(n, m, "f32r", kind, K, W)
        p32 = plan_execute_ms(r32["batch"], "f32r", 5)
        s32 = plan_execute_ms(r32["batch"], "f32r", 3, win_levels=2) if n > 1024 else None
        out[label] = {
            "n": n, "m": m, "generator": "dominant" if kind == 0 else "weak",
            "layout": "interleaved",
            "f64_unknowns_per_s": r64["value"], "f64_ms": r64["ms_per_step"],
            "f64_roofline_frac": BYTES_PER_UNKNOWN * n * m / r64["solve_s"] / HBM_PEAK,
            "f32r_unknowns_per_s": r32["value"], "f32r_ms": r32["ms_per_step"],
            "f32r_roofline_frac": BYTES_PER_UNKNOWN * n * m / r32["solve_s"] / HBM_PEAK,
            "ratio_f32r_vs_f64": r32["value"] / r64["value"],
            "f32r_sweeps_hist": r32["sweeps_hist"],
            "f32r_plan_execute_ms": p32["ms"], "f32r_plan_route": p32["route"],
            "ratio_f32r_plan_vs_f64": r64["ms_per_step"] / p32["ms"],
            "f32r_stage_path": s32 and {**s32, "ratio_vs_f64": r64["ms_per_step"] / s32["ms"]},
            "f64_tuned": (r64["tuned"] or {}).get("variant", (r64["tuned"] or {}).get("split")),
            "rocprof": {"f64": rocprof_summary(wl[0], r64["solve_s"]),
                        "f32r": rocprof_summary(wl[1], r32["solve_s"])},
        }
        del r64, r32
    return out


def _interp_price(U, p, hS, hv):
    """Bilinear price at S = 100, v = 0.04 from the grid nodes."""
    fi = 100.0 / hS - 1.0     # fractional column index (node i at column i - 1)
    fj = 0.04 / hv
    i0, j0 = int(fi), int(fj)
    ti, tj = fi - i0, fj - j0
    return ((1 - tj) * ((1 - ti) * U[j0, i0] + ti * U[j0, i0 + 1])
            + tj * ((1 - ti) * U[j0 + 1, i0] + ti * U[j0 + 1, i0 + 1]))


# ---------------------------------------------------------------- single-form workloads

def bench_single(args):
    """One form of one workload (what the rocprof runs use)."""
    import numpy as np
    import torch
    import pcrpenta
    n, m, prec, kind = WORKLOADS[args.workload]
    K, W = args.steps, args.warmup
    torch.cuda.set_device(0)
    if args.tune_load:
        pcrpenta.tune_load(args.tune_load)
    if n < 0:
        r = run_heston(prec, K, W, reuse=n == -2)
        p = r["p"]
        out = {"metric": METRIC, "value": r["value"], "unit": "unknowns/s", "n_gpus": 1,
               "steps": K, "warmup": W, "ms_per_step": r["ms_per_run"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": prec,
               "data": "synthetic (Heston European call, Ns = 1024, Nv = 512, 200 steps of 1/200)",
               "config": {"workload": args.workload, "ns": p.ns, "nv": p.nv,
                          "time_steps": p.steps, "graph": "20 Douglas steps per hipGraph",
                          "factor_reuse": n == -2,
                          "unknowns_per_run": r["unknowns"]}}
        print(json.dumps(out), flush=True)
        return
    if n == 0:
        B = uneven_batch(kind)
        if not args.no_tune and prec == "f64":
            log({"split_autotune": tune_uneven_split(B)})
        elif not args.no_tune:
            log({"f32r_autotune": tune_uneven_f32r(B)})
        r = run_uneven(B, prec, K, W)
        unknowns = B["total"]
        cfg = {"workload": args.workload, "m": m, "unknowns": unknowns,
               "layout": "contiguous per system", "precision": prec,
               "split": list(pcrpenta.split_tune_get()[:3])}
    else:
        r = run_uniform(n, m, prec, kind, K, W, tune=not args.no_tune)
        if r["tuned"]:
            log({"autotune": r["tuned"]})
        unknowns = n * m
        cfg = {"workload": args.workload, "n": n, "m": m, "layout": "interleaved",
               "precision": prec, "generator": "dominant" if kind == 0 else "weak",
               "split": list(pcrpenta.split_tune_get()[:3]) if n > 1024 else None}
    if args.tune_save:
        pcrpenta.tune_save(args.tune_save)
    if prec != "f64":
        cfg["refine"] = {"max_sweeps": MAX_SWEEPS, "tol": TOL, "sweeps_hist": r["sweeps_hist"]}
    out = {"metric": METRIC, "value": r["value"], "unit": "unknowns/s", "n_gpus": 1, "steps": K,
           "warmup": W, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": prec,
           "data": "synthetic (seeded counter-based generator)", "config": cfg,
           "roofline": roofline(unknowns, r["solve_s"], args.workload, "whole solve per step")}
    if not args.no_cpu and n > 0:
        oracle = oracle_native()
        nsys = min(args.cpu_systems, m)
        arrs = oracle.gen_batch(kind, 0, n, nsys, oracle.INTERLEAVED)
        tmed, runs, x = median_runs(
            lambda: oracle.pcr_batch(*arrs, n, nsys, oracle.INTERLEAVED), args.cpu_runs)
        xs = r["x"].view(n, m)[:, :nsys].cpu().numpy()
        chk = {"bit_identical_to_gpu": bool(np.array_equal(x.reshape(n, nsys), xs))}
        if prec != "f64":
            xg = oracle.gbsv_batch(*arrs, n, nsys, oracle.INTERLEAVED)
            chk = {"gpu_relerr_vs_dgbsv": oracle.normwise_relerr(xs.reshape(-1), xg, n, nsys,
                                                                 oracle.INTERLEAVED)}
        out["cpu_baseline"] = {"value": n * nsys / tmed, "unit": "unknowns/s",
                               "sample": f"first {nsys} of the {m} systems (n={n}, same "
                                         f"generator), median of {runs} runs",
                               **cpu_fields(oracle, runs, tmed), **chk}
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
    pcrpenta.load()
    torch.cuda.set_device(0)
    apply_split(args)
    if args.workload == HEADLINE:
        return bench_headline(args)
    return bench_single(args)


if __name__ == "__main__":
    main()
