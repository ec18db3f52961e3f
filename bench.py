"""GP iterations/s of the device-resident adaptive global placement on one MI355X.

A step is one outer iteration of the adaptive loop (AdaptiveGlobalPlacer.cpp:735-928): a
gradient evaluation, an optional Nesterov look-ahead gradient evaluation, and the backtracking
value evaluations, with the controller's vector passes on the device.

Workload (default): the single large design at real shape -- 210,000 movable cells, 12 fixed
macros (15 % of the core), 220,000 nets (+0.2 % at fanout 50-500, ~0.9 M pins), 1,200 rows x
1,200 sites, 512 x 512 bins, target density 0.8, seeded synthetic netlist, quadratic initial
placement (InitialPlacer quadraticPlacement, solved on the GPU before the session), 280 adaptive
iterations (10 warm-up + 270 timed) with early stop disabled.  Inputs are resident in HBM before
the timed region starts.

--workload precision: 1.3 M cells, 20 macros (12 %), 1.4 M nets (~5.7 M pins), 3,000 rows x 3,000
sites, 1024 x 1024 bins, target density 0.7, quadratic initial placement, DCT products in split
BF16 (bf16x3).  --workload plain: 2,000 cells, 64 x 64 bins, random initial placement.
--workload batch: 256 independent blocks (see batch_sample) through the placement service.
`roofline_dct` reports the DCT products against the MFMA peak of their arithmetic.

Prints one JSON line.  `roofline` reports k_wire, the kernel with the largest share of GPU time,
against MI355X's HBM peak: its algorithmic bytes (per evaluation 12 per pin, 8 per net, 16 per
particle with the gradient and 8 without -- one pass over each array at 4-byte indices and 8-byte
positions) over its launch durations, timed with HIP events on the session's stream inside the
timed region; `traffic` is its measured HBM bytes per launch from the counter summary under
profiles/.  `cpu_baseline` is the FP64 CPU restatement of the
reference path (oracle/) timed on one host core on a bounded number of iterations of the same
workload.  `kernel_timing_us` (average per launch of every kernel class) comes from a separate
stretch of iterations after the timed region, so that its events do not slow the timed steps.

`serving` reports the second half of the metric on a bounded sample of the served mixed-size
stream (p50 / p99 submit-to-result latency, jobs per hour; see serving_sample).

--copy-calibration MB runs one plain device-to-device copy of MB MiB (torch) and exits: a kernel
of known traffic whose FETCH_SIZE / WRITE_SIZE counters scale those of k_wire in a profiling run.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
METRICS = {
    "large": "GP iters/s @210k cells 512\u00b2 bins; p50/p99 placement latency on served mix",
    "precision": "GP iters/s @1.3M cells 1024\u00b2 bins, bf16x3 DCT products",
    "plain": "GP iters/s @2k cells 64\u00b2 bins",
    "batch": "GP iters/s aggregate over 256 independent blocks (500-50k cells) on one GPU",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str):
    from myplacement_amd import abi, design
    from myplacement_amd import backend as hb
    if name == "large":
        d = design.large_design(seed=2026)
        o = abi.default_options(target_density=0.8, bins_x=512, bins_y=512)
        desc = {"workload": "synthetic 210k cells, 12 macros (15%), 220k nets (~0.9M pins), 1200x1200 sites, "
                            "512x512 bins, t=0.8, quadratic init", "cells": 210000, "bins": 512, "target_density": 0.8}
        d.center, _ = hb.quadratic_placement(d)
    elif name == "precision":
        d = design.large_design(seed=2026, cells=1_300_000, nets=1_400_000, row_count=3000, sites=3000, macros=20,
                                macro_fraction=0.12)
        o = abi.default_options(target_density=0.7, bins_x=1024, bins_y=1024, dct_precision=abi.MP_DCT_BF16X3)
        desc = {"workload": "synthetic 1.3M cells, 20 macros (12%), 1.4M nets (~5.7M pins), 3000x3000 sites, "
                            "1024x1024 bins, t=0.7, quadratic init, split-BF16 (bf16x3) DCT products",
                "cells": 1_300_000, "bins": 1024, "target_density": 0.7}
        d.center, _ = hb.quadratic_placement(d)
    elif name == "plain":
        d = design.plain_design(seed=2026)
        o = abi.default_options(target_density=0.9, bins_x=64, bins_y=64)
        desc = {"workload": "synthetic 2k cells, 2.3k nets, 40x400 sites, 64x64 bins, t=0.9, random init",
                "cells": 2000, "bins": 64, "target_density": 0.9}
        d.center = hb.random_placement(d, 2026)
    else:
        raise SystemExit(f"unknown workload {name}")
    o.iterations = 1_000_000
    o.feasible_refinement_iterations = 2 ** 31 - 1
    return d, o, desc


def cpu_baseline(name: str, iterations: int) -> dict:
    """The oracle's single-threaded FP64 GP loop on the same design (checker, not product)."""
    from oracle import bindings as ob
    d, o, _ = make_workload(name)
    o.iterations = iterations
    _, r, _ = ob.global_place(d, o)
    return {"value": iterations / r["loop_seconds"], "unit": "GP iterations/s", "cores": 1, "kind": "port",
            "sample": f"{iterations} adaptive iterations of the same workload, FP64 oracle (own FFT DCT), 1 thread",
            "loop_seconds": r["loop_seconds"]}


def serving_sample(jobs: int, rate: float, slots: int, seed: int = 2026, mode: str = "streams") -> dict:
    """A bounded sample of the served mixed-size stream: `jobs` requests with Poisson arrivals at
    `rate` per second, sizes log-uniform in 1k-200k cells (designs.served_design), bin grids by
    designs.served_bins, target density uniform in 0.7-0.9, random initial placement, each run until
    design overflow <= 0.10 (checked from iteration 21 on) or 280 iterations.  Jobs are generated
    before the stream starts; latency is submit -> result.  mode: the service's mode ("streams": a
    stream per job; "batched": lockstep rounds of batched launches, jobs joining between rounds)."""
    import threading
    from myplacement_amd import abi, design
    from myplacement_amd import backend as hb
    specs, gaps = design.served_stream_specs(jobs, rate, seed)
    work = []
    for job_seed, cells, side, t in specs:
        d = design.served_design(job_seed, cells)
        d.center = hb.random_placement(d, job_seed)
        o = abi.default_options(target_density=t, bins_x=side, bins_y=side, iterations=280,
                                feasible_refinement_iterations=1, seed=job_seed)
        work.append((d, o, cells, side))
    server = hb.PlacementServer(max_active_jobs=slots, memory_budget_bytes=200 << 30, mode=mode)
    reports = [None] * jobs
    threads = []

    def collect(i, job, m):
        reports[i] = server.wait(job, m)[1]

    t0 = time.perf_counter()
    due = 0.0
    for i, (d, o, cells, side) in enumerate(work):
        due += gaps[i]
        delay = due - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        job = server.submit(d, o)
        th = threading.Thread(target=collect, args=(i, job, d.module_count))
        th.start()
        threads.append(th)
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    server.close()
    lat = np.array([r["latency_seconds"] for r in reports])
    conv = [r["result"]["overflow_after"] <= 0.10 + 1e-12 for r in reports]
    first = min(r["submit_seconds"] for r in reports)
    last = max(r["finish_seconds"] for r in reports)
    # capacity: the same requests all submitted at once, to the plain backend one job at a time and
    # to the batched service
    capacity = {}
    for name, m, k in (("plain_one_at_a_time", "streams", 1), ("batched", "batched", jobs)):
        span, reps = run_all_at_once([(d, o) for d, o, _, _ in work], m, k)
        capacity[name] = {"mode": m, "slots": k, "makespan_s": span, "jobs_per_hour": 3600.0 * jobs / span,
                          "gp_iterations_per_s": sum(r["result"]["completed_iterations"] for r in reps) / span}
        log(f"serving capacity: {name}, makespan {span:.2f}s")
    capacity["ratio"] = capacity["batched"]["jobs_per_hour"] / capacity["plain_one_at_a_time"]["jobs_per_hour"]
    return {
        "capacity_all_at_once": capacity,
        "jobs": jobs, "arrival_rate_per_s": rate, "slots": slots, "mode": mode,
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()), "jobs_per_hour": 3600.0 * jobs / (last - first),
        "converged_fraction": float(np.mean(conv)),
        "mean_iterations": float(np.mean([r["result"]["completed_iterations"] for r in reports])),
        "cells_range": [int(min(w[2] for w in work)), int(max(w[2] for w in work))], "wall_s": wall,
        "sample": f"{jobs} requests, Poisson {rate}/s, 1k-200k cells log-uniform, t~U(0.7,0.9), "
                  "bins 32-256 by >=4 cells/bin, stop at overflow<=0.10 or 280 iterations",
    }


def run_all_at_once(work, mode: str, slots: int):
    """Submits every (design, options) of `work` at once to a placement service; returns the
    makespan and the job reports."""
    from myplacement_amd import backend as hb
    server = hb.PlacementServer(max_active_jobs=slots, memory_budget_bytes=200 << 30, mode=mode)
    t0 = time.perf_counter()
    ids = [server.submit(d, o) for d, o in work]
    reports = [server.wait(j, w[0].module_count)[1] for j, w in zip(ids, work)]
    wall = time.perf_counter() - t0
    server.close()
    return wall, reports


def parse_batch_runs(text: str):
    """"mode:slots[:precision],..." -> [(mode, slots, precision)]."""
    runs = []
    for item in text.split(","):
        f = item.split(":")
        runs.append((f[0], int(f[1]), f[2] if len(f) > 2 else "fp32"))
    return runs


def batch_sample(jobs: int, runs_spec, iterations: int, seed: int = 2026) -> dict:
    """The batching configuration's workload: `jobs` independent blocks, each from its own seed,
    cells log-uniform in 500-50k (designs.served_design), bin grids by designs.served_bins (32 or 64
    at these sizes), target density 0.8, random initial placement, a fixed `iterations` adaptive
    iterations each (early stop off).  All jobs are submitted at once to the placement service, for
    each (mode, slots, precision) of `runs_spec`: mode "streams" runs each admitted job on its own
    stream, "batched" runs the admitted jobs in lockstep rounds with one batched launch per stage;
    slots bounds the jobs running at once; precision "bf16x3" gives the jobs whose grid allows it
    (both sides multiples of 64) split-BF16 DCT products, "fp32" none.  The aggregate rate is the
    total of the jobs' GP iterations over the makespan."""
    from myplacement_amd import abi, design
    from myplacement_amd import backend as hb
    work = []
    for job_seed, cells, side in design.batch_job_specs(jobs, seed):
        d = design.served_design(job_seed, cells)
        d.center = hb.random_placement(d, job_seed)
        o = abi.default_options(target_density=0.8, bins_x=side, bins_y=side, iterations=iterations,
                                feasible_refinement_iterations=2 ** 31 - 1, seed=job_seed)
        work.append((d, o, cells, side))
    log(f"batch: {jobs} blocks generated, {sum(w[2] for w in work)} cells in total")
    runs = {}
    for mode, slots, precision in runs_spec:
        jobs_here = []
        for d, o, _, side in work:
            if precision == "bf16x3" and side % 64 == 0:
                o = abi.MpOptions.from_buffer_copy(o)
                o.dct_precision = abi.MP_DCT_BF16X3
            jobs_here.append((d, o))
        wall, reports = run_all_at_once(jobs_here, mode, slots)
        its = sum(r["result"]["completed_iterations"] for r in reports)
        key = f"{mode}/{slots}" + ("" if precision == "fp32" else f"/{precision}")
        runs[key] = {"mode": mode, "slots": slots, "precision": precision, "makespan_s": wall, "gp_iterations": its,
                     "gp_iterations_per_s": its / wall, "jobs_per_hour": 3600.0 * jobs / wall}
        log(f"batch: {mode}, {slots} slots, {precision}, makespan {wall:.2f}s, {its / wall:.0f} GP it/s")
    sides = [w[3] for w in work]
    return {"jobs": jobs, "iterations_per_job": iterations, "runs": runs,
            "cells_range": [int(min(w[2] for w in work)), int(max(w[2] for w in work))],
            "grids": {str(g): sides.count(g) for g in sorted(set(sides))},
            "sample": f"{jobs} blocks, cells log-uniform 500-50k, bins 32/64 by >=4 cells/bin, t=0.8, random init, "
                      f"{iterations} iterations each, all submitted at once"}


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def dct_roofline(t, precision: int):
    """MFMA roofline of the DCT products over the all-class timing stretch: the multiply-adds the
    half-length products execute (2 flops each, FP32-equivalent) over their event-timed launch
    durations, against FP32 MFMA or, for split-BF16, dense BF16 MFMA (which runs three bf16
    products per FP32-equivalent product)."""
    if not t or not t["launches"] or t["total_us"] <= 0:
        return None
    achieved = t["work"] / (t["total_us"] * 1e-6) / 1e12
    bf16 = precision == 1
    peak = BF16_MFMA_PEAK_TFLOPS / 3.0 if bf16 else FP32_MFMA_PEAK_TFLOPS
    return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "arithmetic": "bf16x3: 3 bf16 products per FP32-equivalent product, peak = 2500/3" if bf16 else "fp32",
            "kernel": "k_dct", "avg_launch_us": t["total_us"] / t["launches"], "launches": t["launches"]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=270)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="large")
    ap.add_argument("--cpu-iterations", type=int, default=60)
    ap.add_argument("--copy-calibration", type=int, default=0)
    ap.add_argument("--serve-jobs", type=int, default=40, help="served-stream sample size (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-mode", default="batched", choices=("streams", "batched"))
    ap.add_argument("--serve-only", action="store_true", help="only the served-stream sample")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="no HIP-event timing (for hardware-counter runs, where the profiler rejects it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--batch-jobs", type=int, default=256, help="batching-form sample size (0: skip)")
    ap.add_argument("--batch-runs", default="streams:1,streams:16,batched:64,batched:256,batched:256:bf16x3",
                    help="service mode:slots[:bf16x3] runs of the batch workload")
    ap.add_argument("--batch-iterations", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU split is a design note, "
              "not a measured path", file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        import torch
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} MiB read, {args.copy_calibration} MiB written")
        return 0
    from myplacement_amd import backend as hb
    if args.serve_only:
        sv = serving_sample(args.serve_jobs, args.serve_rate, args.serve_slots, mode=args.serve_mode)
        print(json.dumps({"serving": sv}), flush=True)
        return 0
    if args.workload == "batch":
        b = batch_sample(args.batch_jobs, parse_batch_runs(args.batch_runs), args.batch_iterations)
        best = max(b["runs"].values(), key=lambda r: r["gp_iterations_per_s"])
        line = {"metric": METRICS["batch"], "value": best["gp_iterations_per_s"], "unit": "GP iterations/s",
                "n_gpus": 1, "steps": b["jobs"] * b["iterations_per_job"], "warmup": 0,
                "ms_per_step": 1000.0 / best["gp_iterations_per_s"], "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64 positions/reductions, f32 pin terms, f32 DCT products",
                "data": "synthetic", "config": {"workload": b["sample"], "mode": best["mode"], "slots": best["slots"],
                           "dct": best["precision"]},
                "batch": b}
        print(json.dumps(line), flush=True)
        return 0
    d, o, desc = make_workload(args.workload)
    t0 = time.perf_counter()
    prep = hb.PreparedDesign(d, o)
    sess = hb.PlacementSession(prep, o)
    setup_s = time.perf_counter() - t0
    log(f"setup {setup_s:.2f}s P={prep.info.particle_count} pins={prep.info.pin_count}")
    sess.step(args.warmup)
    log(f"warmup {args.warmup} done")
    h0 = len(sess.history())
    timed = not args.no_kernel_timing
    if timed:
        sess.enable_kernel_timing(True, classes=["k_wire"])
    t1 = time.perf_counter()
    done = sess.step(args.steps)
    sess.positions()  # device work of the last step is complete
    elapsed = time.perf_counter() - t1
    log(f"timed {done} steps in {elapsed:.3f}s")
    hist = sess.history()[h0:]
    roof, timing = None, {}
    if timed:
        live = sess.kernel_timing()["k_wire"]
        sess.enable_kernel_timing(False)
        # every kernel class, timed over a separate stretch of the same run
        sess.enable_kernel_timing(True)
        sess.step(args.timing_steps)
        timing = sess.kernel_timing()
        sess.enable_kernel_timing(False)
        timing["k_wire"] = {kk: timing["k_wire"][kk] - live[kk] for kk in ("total_us", "launches", "work")}
        avg_us = live["total_us"] / live["launches"]
        achieved = live["work"] / (live["total_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_wire") if args.workload == "large" else None, "kernel": "k_wire",
                "avg_launch_us": avg_us, "launches": live["launches"],
                "algorithmic_bytes_per_launch": live["work"] / live["launches"]}
    _, res = sess.finish()

    line = {
        "metric": METRICS[args.workload],
        "value": done / elapsed,
        "unit": "GP iterations/s",
        "n_gpus": 1,
        "steps": done,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / max(done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 positions/reductions, f32 pin terms, "
                 + ("bf16x3 DCT products" if o.dct_precision == 1 else "f32 DCT products"),
        "data": "synthetic",
        "config": dict(desc, evaluations_per_step=(
            sum(1 for _ in hist) and (res["gradient_evaluations"] + res["value_evaluations"]) / max(1, res["completed_iterations"]))),
        "roofline": roof,
        "kernel_timing_us": {k: round(v["total_us"] / max(1, v["launches"]), 3) for k, v in timing.items()
                             if v["launches"]},
        "roofline_dct": dct_roofline(timing.get("k_gemm"), o.dct_precision),
        "setup_seconds": setup_s,
    }
    sess.close()
    if args.serve_jobs > 0:
        log(f"served stream sample: {args.serve_jobs} jobs ...")
        line["serving"] = serving_sample(args.serve_jobs, args.serve_rate, args.serve_slots, mode=args.serve_mode)
    if args.batch_jobs > 0:
        log(f"batching sample: {args.batch_jobs} blocks ...")
        b = batch_sample(args.batch_jobs, parse_batch_runs(args.batch_runs), args.batch_iterations)
        best = max(b["runs"].values(), key=lambda r: r["gp_iterations_per_s"])
        b["metric"] = METRICS["batch"]
        b["best"] = f"{best['mode']}/{best['slots']}/{best['precision']}"
        line["batching"] = b
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cb = cpu_baseline(args.workload, args.cpu_iterations)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
