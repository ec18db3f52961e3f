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

Prints one JSON line.  `roofline` reports k_wire_dbins, the kernel with the largest share of GPU
time (the wirelength of the candidate and look-ahead evaluations with the density's bin pass in the
same launch), against MI355X's HBM peak: its algorithmic bytes (per evaluation of the wirelength 12
per pin, 8 per net, 16 per particle with the gradient and 8 without, and of the bin pass 24 per bin
-- one pass over each array at 4-byte indices and 8-byte values) over its launch durations, timed
with HIP events on the session's stream inside the timed region (every fifth launch, every third
in runs under 100 steps, so that the events slow the timed steps less); `traffic` is its measured HBM
bytes per launch from the counter summary under profiles/.  `cpu_baseline` is the FP64 CPU restatement of the
reference path (oracle/) timed on one host core on a bounded number of iterations of the same
workload.  `kernel_timing_us` (average per launch of every kernel class) comes from a separate
stretch of iterations after the timed region, so that its events do not slow the timed steps.

`serving` reports the second half of the metric on a bounded sample of the served mixed-size
stream (p50 / p99 submit-to-result latency, jobs per hour; see serving_sample).

--copy-calibration MB runs one plain device-to-device copy of MB MiB (torch) and exits: a kernel
of known traffic whose FETCH_SIZE / WRITE_SIZE counters scale those of k_wire_dbins in a profiling run.
"""

from __future__ import annotations

import argparse
import hashlib
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


def make_workload(name: str, center=None):
    """The design, options and description of a workload; `center` replaces the computed start (the
    CPU baseline's child runs from the start the GPU run used)."""
    from myplacement_amd import abi, design
    from myplacement_amd import backend as hb
    if name == "large":
        d = design.large_design(seed=2026)
        o = abi.default_options(target_density=0.8, bins_x=512, bins_y=512)
        desc = {"workload": "synthetic 210k cells, 12 macros (15%), 220k nets (~0.9M pins), 1200x1200 sites, "
                            "512x512 bins, t=0.8, quadratic init", "cells": 210000, "bins": 512, "target_density": 0.8}
        d.center = center if center is not None else hb.quadratic_placement(d)[0]
        desc["start_sha256"] = hashlib.sha256(np.ascontiguousarray(d.center, dtype="<f8").tobytes()).hexdigest()
    elif name == "precision":
        d = design.large_design(seed=2026, cells=1_300_000, nets=1_400_000, row_count=3000, sites=3000, macros=20,
                                macro_fraction=0.12)
        o = abi.default_options(target_density=0.7, bins_x=1024, bins_y=1024, dct_precision=abi.MP_DCT_BF16X3)
        desc = {"workload": "synthetic 1.3M cells, 20 macros (12%), 1.4M nets (~5.7M pins), 3000x3000 sites, "
                            "1024x1024 bins, t=0.7, quadratic init, split-BF16 (bf16x3) DCT products",
                "cells": 1_300_000, "bins": 1024, "target_density": 0.7}
        d.center = center if center is not None else hb.quadratic_placement(d)[0]
    elif name == "plain":
        d = design.plain_design(seed=2026)
        o = abi.default_options(target_density=0.9, bins_x=64, bins_y=64)
        desc = {"workload": "synthetic 2k cells, 2.3k nets, 40x400 sites, 64x64 bins, t=0.9, random init",
                "cells": 2000, "bins": 64, "target_density": 0.9}
        d.center = center if center is not None else hb.random_placement(d, 2026)
    else:
        raise SystemExit(f"unknown workload {name}")
    o.iterations = 1_000_000
    o.feasible_refinement_iterations = 2 ** 31 - 1
    return d, o, desc


def dropin_sample(prep, calls: int) -> dict:
    """The drop-in boundary as the reference defines it (CudaPlacementBackend::evaluate): positions
    copied in from host memory and, with gradients, two gradient vectors copied out, on every call.
    PCIe-inclusive time per call on the workload's design at its starting positions; never `value`."""
    from myplacement_amd import backend as hb
    layout = prep.layout()
    be = hb.HipPlacementBackend.create(layout)
    xy = prep.initial_positions()
    r = layout["region"]
    gamma = 0.5 * max((r[2] - r[0]) / layout["bins_x"], (r[3] - r[1]) / layout["bins_y"])
    out = {}
    for grad in (True, False):
        for _ in range(10):
            be.evaluate(xy, gamma, grad)
        t = time.perf_counter()
        for _ in range(calls):
            be.evaluate(xy, gamma, grad)
        out["gradient_ms" if grad else "value_ms"] = 1000.0 * (time.perf_counter() - t) / calls
    be.close()
    out["calls"] = calls
    out["note"] = ("mp_gp_evaluate with pageable host buffers: positions up through the runtime's pageable path, "
                   "gradients down through pinned staging")
    log(f"drop-in evaluate: {out['gradient_ms']:.3f} ms with gradients, {out['value_ms']:.3f} ms value-only")
    return out


def host_description() -> dict:
    """nproc, the CPU model and the cores this process may use (the CPU baseline's hardware)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_baseline_run(name: str, iterations: int, start_file: str) -> dict:
    """One oracle run (checker, not product) of the workload in this process: FP64, single thread."""
    from oracle import bindings as ob
    d, o, _ = make_workload(name, center=np.load(start_file))
    o.iterations = iterations
    _, r, _ = ob.global_place(d, o)
    return {"loop_seconds": r["loop_seconds"], "elapsed_seconds": r["elapsed_seconds"],
            "iterations": r["completed_iterations"]}


def cpu_baseline(name: str, iterations: int, repeats: int, start_file: str) -> dict:
    """The oracle's FP64 GP loop on the same design and start, `repeats` runs, each in a child
    process pinned to one core (the reference's CPU path is single-threaded), median reported:
    loop-only iterations/s (the metric's unit) and set-up-inclusive iterations/s (the reference's
    elapsed_seconds counts set-up too)."""
    import statistics
    import subprocess
    core = min(os.sched_getaffinity(0))
    runs = []
    for k in range(repeats):
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", name,
                              "--cpu-iterations", str(iterations), "--start-file", start_file],
                             capture_output=True, text=True, preexec_fn=lambda: os.sched_setaffinity(0, {core}))
        if out.returncode != 0:
            raise RuntimeError("cpu baseline run failed: " + out.stderr[-2000:])
        runs.append(json.loads(out.stdout.strip().splitlines()[-1]))
        log(f"cpu baseline repeat {k + 1}/{repeats}: {runs[-1]['iterations'] / runs[-1]['loop_seconds']:.2f} it/s")
    loop = statistics.median(r["iterations"] / r["loop_seconds"] for r in runs)
    total = statistics.median(r["iterations"] / r["elapsed_seconds"] for r in runs)
    return dict({"value": loop, "unit": "GP iterations/s", "cores": 1, "kind": "port",
                 "sample": f"{iterations} adaptive iterations of the same workload from the same start, FP64 oracle "
                           f"(own FFT DCT, FFTW absent), 1 thread pinned to one core, median of {repeats} runs",
                 "setup_inclusive_value": total, "repeats": [r["iterations"] / r["loop_seconds"] for r in runs],
                 "pinned_core": core}, **host_description())


def hpwl_deviation_sample(work, reports, pick, max_cells: int) -> dict:
    """Per-job final HPWL and overflow of served jobs against the FP64 oracle's run of the same job
    (same design, start and options), outside the timed stream, for the requests `pick` (chosen
    evenly spaced in stream order among those of at most `max_cells` cells: the oracle runs ~10^4 x
    slower than the device).  Each pair is one run per side, so at small sizes the chaotic spread
    between runs (~0.5 % at 2k cells) is part of the deviation."""
    from oracle import bindings as ob
    if not pick:
        return {"jobs": 0}
    rows = []
    t = time.perf_counter()
    for i in pick:
        d, o, cells, side = work[i]
        _, r, _ = ob.global_place(d, o)
        mine = reports[i]["result"]
        rows.append({"request": i, "cells": cells, "bins": side,
                     "hpwl_deviation": mine["hpwl_after"] / r["hpwl_after"] - 1.0,
                     "overflow_deviation": mine["overflow_after"] - r["overflow_after"],
                     "iterations": [mine["completed_iterations"], r["completed_iterations"]]})
    dev = np.array([r["hpwl_deviation"] for r in rows])
    ovf = np.array([r["overflow_deviation"] for r in rows])
    log(f"served-job HPWL deviation vs the FP64 oracle: {len(rows)} jobs in {time.perf_counter() - t:.1f}s, "
        f"mean {dev.mean():+.3%}, max |{np.abs(dev).max():.3%}|")
    return {"jobs": len(rows), "max_cells": max_cells, "mean_hpwl_deviation": float(dev.mean()),
            "mean_abs_hpwl_deviation": float(np.abs(dev).mean()), "max_abs_hpwl_deviation": float(np.abs(dev).max()),
            "mean_overflow_deviation": float(ovf.mean()), "max_abs_overflow_deviation": float(np.abs(ovf).max()),
            "per_job": rows}


def serving_sample(jobs: int, rate: float, slots: int, seed: int = 2026, mode: str = "streams",
                   capacity_jobs: int = 200, deviation_jobs: int = 16, deviation_max_cells: int = 10000) -> dict:
    """The served mixed-size stream: `jobs` requests with Poisson arrivals at `rate` per second,
    sizes log-uniform in 1k-200k cells (design.served_stream_specs, design.served_design), bin grids
    by design.served_bins, target density uniform in 0.7-0.9, random initial placement, each run
    until design overflow <= 0.10 (checked from iteration 21 on) or 280 iterations.  A producer
    thread builds each request's design ahead of its arrival time (the designs of ~75M cells in all
    are never held at once); latency is submit -> result.  mode: the service's mode ("streams": a
    stream per job; "batched": lockstep rounds of batched launches, jobs joining between rounds)."""
    import queue
    import threading
    from myplacement_amd import abi, design
    from myplacement_amd import backend as hb
    specs, gaps = design.served_stream_specs(jobs, rate, seed)
    eligible = [i for i, sp in enumerate(specs) if sp[1] <= deviation_max_cells]
    picks = {eligible[int(k * len(eligible) / deviation_jobs)] for k in range(min(deviation_jobs, len(eligible)))}
    keep = set(range(min(capacity_jobs, jobs))) | picks

    def build(i):
        job_seed, cells, side, t = specs[i]
        d = design.served_design(job_seed, cells)
        d.center = hb.random_placement(d, job_seed)
        o = abi.default_options(target_density=t, bins_x=side, bins_y=side, iterations=280,
                                feasible_refinement_iterations=1, seed=job_seed)
        return d, o, cells, side

    ready = queue.Queue(maxsize=64)

    def produce():
        for i in range(jobs):
            ready.put(build(i))

    producer = threading.Thread(target=produce, daemon=True)
    producer.start()
    server = hb.PlacementServer(max_active_jobs=slots, memory_budget_bytes=200 << 30, mode=mode)
    reports = [None] * jobs
    kept = {}
    threads = []

    def collect(i, job, m):
        reports[i] = server.wait(job, m)[1]

    log(f"served stream: {jobs} requests, {sum(sp[1] for sp in specs)} cells in total")
    late = 0.0
    t0 = time.perf_counter()
    due = 0.0
    for i in range(jobs):
        w = ready.get()
        due += gaps[i]
        delay = due - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        else:
            late = max(late, -delay)
        job = server.submit(w[0], w[1])
        if i in keep:
            kept[i] = w
        th = threading.Thread(target=collect, args=(i, job, w[0].module_count))
        th.start()
        threads.append(th)
        if (i + 1) % 100 == 0:
            log(f"served stream: {i + 1} requests submitted at {time.perf_counter() - t0:.0f}s, "
                f"{sum(r is not None for r in reports)} done")
    for th in threads:
        th.join()
    producer.join()
    work = [kept.get(i) for i in range(jobs)]
    wall = time.perf_counter() - t0
    server.close()
    lat = np.array([r["latency_seconds"] for r in reports])
    conv = [r["result"]["overflow_after"] <= 0.10 + 1e-12 for r in reports]
    first = min(r["submit_seconds"] for r in reports)
    last = max(r["finish_seconds"] for r in reports)
    # capacity: the stream's first `capacity_jobs` requests all submitted at once, to the plain
    # backend one job at a time and to the batched service
    capacity = {"jobs": min(capacity_jobs, jobs)}
    sub = [(w[0], w[1]) for w in work[:capacity["jobs"]]]
    for name, m, k in (("plain_one_at_a_time", "streams", 1), ("batched", "batched", len(sub))):
        span, reps = run_all_at_once(sub, m, k)
        capacity[name] = {"mode": m, "slots": k, "makespan_s": span, "jobs_per_hour": 3600.0 * len(sub) / span,
                          "gp_iterations_per_s": sum(r["result"]["completed_iterations"] for r in reps) / span}
        log(f"serving capacity: {name}, makespan {span:.2f}s")
    capacity["ratio"] = capacity["batched"]["jobs_per_hour"] / capacity["plain_one_at_a_time"]["jobs_per_hour"]
    deviation = hpwl_deviation_sample(work, reports, sorted(picks), deviation_max_cells)
    return {
        "capacity_all_at_once": capacity,
        "jobs": jobs, "arrival_rate_per_s": rate, "slots": slots, "mode": mode,
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()), "jobs_per_hour": 3600.0 * jobs / (last - first),
        "converged_fraction": float(np.mean(conv)),
        "mean_iterations": float(np.mean([r["result"]["completed_iterations"] for r in reports])),
        "cells_range": [int(min(sp[1] for sp in specs)), int(max(sp[1] for sp in specs))], "wall_s": wall,
        "latest_submission_s": late,
        "sample": f"{jobs} requests, Poisson {rate}/s, 1k-200k cells log-uniform, t~U(0.7,0.9), "
                  "bins 32-256 by >=4 cells/bin, stop at overflow<=0.10 or 280 iterations",
        "hpwl_vs_oracle": deviation,
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


# the roofline's kernel: its timing class and its name in the profiler's traces; every roof_every(steps)-th
# launch of it in the timed region is bracketed by events (each event pair costs the stream ~1.5 us).
# The class's launches alternate look-ahead (gradient) and candidate (value-only) evaluations in
# momentum iterations, so the period is odd: the sample then holds both kinds.  Short timed regions
# (under 100 steps) time every third launch, so that they still hold several.
ROOF_CLASS, ROOF_KERNEL = "k_wire_bins", "k_wire_dbins"


def roof_every(steps: int) -> int:
    return 5 if steps >= 100 else 3


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
    ap.add_argument("--cpu-iterations", type=int, default=40, help="oracle iterations per CPU-baseline repeat")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--cpu-baseline-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--start-file", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--dropin-evaluations", type=int, default=30,
                    help="timed calls of the reference-shaped mp_gp_evaluate with host buffers (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0)
    ap.add_argument("--serve-jobs", type=int, default=2000, help="served-stream requests (0: skip)")
    ap.add_argument("--serve-capacity-jobs", type=int, default=200,
                    help="requests of the stream submitted at once for the capacity comparison")
    ap.add_argument("--serve-deviation-jobs", type=int, default=16,
                    help="served jobs re-run in the FP64 oracle for the per-job HPWL deviation")
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
    if args.cpu_baseline_child:
        print(json.dumps(cpu_baseline_run(args.cpu_baseline_child, args.cpu_iterations, args.start_file)), flush=True)
        return 0
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
        sv = serving_sample(args.serve_jobs, args.serve_rate, args.serve_slots, mode=args.serve_mode,
                            capacity_jobs=args.serve_capacity_jobs, deviation_jobs=args.serve_deviation_jobs)
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
    start_center = d.center.copy()
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
        sess.enable_kernel_timing(True, classes=[ROOF_CLASS], every=roof_every(args.steps))
    t1 = time.perf_counter()
    done = sess.step(args.steps)
    sess.positions()  # device work of the last step is complete
    elapsed = time.perf_counter() - t1
    log(f"timed {done} steps in {elapsed:.3f}s")
    hist = sess.history()[h0:]
    roof, timing = None, {}
    if timed:
        live = sess.kernel_timing()[ROOF_CLASS]
        sess.enable_kernel_timing(False)
        # every kernel class, timed over a separate stretch of the same run
        sess.enable_kernel_timing(True)
        sess.step(args.timing_steps)
        timing = sess.kernel_timing()
        sess.enable_kernel_timing(False)
        timing[ROOF_CLASS] = {kk: timing[ROOF_CLASS][kk] - live[kk] for kk in ("total_us", "launches", "work")}
        if live["launches"] == 0:  # a timed region too short to reach a sampled launch
            log(f"no {ROOF_KERNEL} launch was timed in {done} steps: no roofline")
        else:
            avg_us = live["total_us"] / live["launches"]
            achieved = live["work"] / (live["total_us"] * 1e-6) / 1e9
            roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic(ROOF_KERNEL) if args.workload == "large" else None,
                    "kernel": ROOF_KERNEL, "avg_launch_us": avg_us, "launches": live["launches"],
                    "algorithmic_bytes_per_launch": live["work"] / live["launches"]}
    _, res = sess.finish()
    dropin = dropin_sample(prep, args.dropin_evaluations) if args.dropin_evaluations > 0 else None

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
        "dropin_boundary": dropin,
    }
    sess.close()
    if args.serve_jobs > 0:
        log(f"served stream sample: {args.serve_jobs} jobs ...")
        line["serving"] = serving_sample(args.serve_jobs, args.serve_rate, args.serve_slots, mode=args.serve_mode,
                                         capacity_jobs=args.serve_capacity_jobs,
                                         deviation_jobs=args.serve_deviation_jobs)
    if args.batch_jobs > 0:
        log(f"batching sample: {args.batch_jobs} blocks ...")
        b = batch_sample(args.batch_jobs, parse_batch_runs(args.batch_runs), args.batch_iterations)
        best = max(b["runs"].values(), key=lambda r: r["gp_iterations_per_s"])
        b["metric"] = METRICS["batch"]
        b["best"] = f"{best['mode']}/{best['slots']}/{best['precision']}"
        line["batching"] = b
    if not args.no_cpu_baseline:
        import tempfile
        log("cpu baseline ...")
        with tempfile.TemporaryDirectory() as tmp:
            start_file = os.path.join(tmp, "start.npy")
            np.save(start_file, start_center)
            line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_iterations, args.cpu_repeats, start_file)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
