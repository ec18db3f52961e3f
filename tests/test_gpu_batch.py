"""Placements batched over jobs (mp_gp_global_place_batch and the service's batched mode): every
stage of every job's evaluation chain runs as one batched launch over all running jobs, and each
job's result must be bit-equal to its own mp_gp_global_place run (each batched workgroup runs the
job's own workgroup body; every evaluation is deterministic).

The mix covers the batched kernels' variants side by side in one launch: dense DCT products (32 x
32 and 96 x 80 grids) next to half-length ones (64 x 64, 128 x 128, 64 x 128), FP32 next to bf16x3
products, and a job that rebuilds its row order for every evaluation (its look-ahead evaluation then
runs after the current one: a second pass of the round)."""

import sys
import threading

import numpy as np
import pytest

from myplacement_amd import abi, design
from myplacement_amd import backend as hb

pytestmark = pytest.mark.gpu

RESULT_KEYS = ("hpwl_after", "overflow_after", "completed_iterations", "accepted_iterations", "gradient_evaluations",
               "value_evaluations", "rejected_candidates", "momentum_restarts", "best_checkpoint_iteration")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _job(seed, cells, iterations=50, **extra):
    d = design.served_design(seed, cells)
    d.center = hb.random_placement(d, seed)
    side = design.served_bins(cells)
    kw = dict(target_density=0.8, bins_x=side, bins_y=side, iterations=iterations, feasible_refinement_iterations=1,
              seed=seed)
    kw.update(extra)
    return d, abi.default_options(**kw)


def _mix():
    return [
        _job(300, 1500),
        _job(301, 4000, bins_x=64, bins_y=64),
        _job(302, 2500, bins_x=64, bins_y=128, dct_precision=abi.MP_DCT_BF16X3),
        _job(303, 9000, density_order_reuse_rows=0.0),
        _job(304, 1200, bins_x=64, bins_y=64, dct_precision=abi.MP_DCT_BF16X3, iterations=80),
        _job(305, 20000, bins_x=128, bins_y=128),
        _job(306, 6000, bins_x=96, bins_y=80),
    ]


def _same_as_alone(d, o, c, r):
    c_alone, r_alone, _ = hb.global_place(d, o)
    assert np.array_equal(c, c_alone), np.abs(c - c_alone).max()
    for k in RESULT_KEYS:
        assert r[k] == r_alone[k], (k, r[k], r_alone[k])


def test_batch_is_bit_equal_to_single_runs():
    jobs = _mix()
    say(f"[gpu] batched placement of {len(jobs)} jobs (dense, half-length, bf16x3, fresh-order)")
    out = hb.global_place_batch([d for d, _ in jobs], [o for _, o in jobs])
    for (d, o), (c, r) in zip(jobs, out):
        say(f"  job {d.module_count} modules, {o.bins_x}x{o.bins_y} bins: {r['completed_iterations']} iterations, "
            f"HPWL {r['hpwl_after']:.1f}")
        assert r["completed_iterations"] > 0 and r["hpwl_after"] < r["hpwl_before"]
        _same_as_alone(d, o, c, r)


def test_batch_of_one_and_a_failing_job():
    """A job with invalid options fails alone; the others still finish, bit-equal."""
    good = _job(310, 2000)
    bad_d, bad_o = _job(311, 2000)
    bad_o.target_density = 1.5
    say("[gpu] batch with one invalid job")
    with pytest.raises(ValueError):
        hb.global_place_batch([good[0], bad_d], [good[1], bad_o])
    say("[gpu] batch of one")
    (c, r), = hb.global_place_batch([good[0]], [good[1]])
    _same_as_alone(good[0], good[1], c, r)


def test_service_batched_mode_is_bit_equal_to_single_runs():
    jobs = _mix() + [_job(320 + k, c) for k, c in enumerate((3000, 1800, 5000))]
    say(f"[gpu] batched service: {len(jobs)} jobs, 4 slots")
    server = hb.PlacementServer(max_active_jobs=4, memory_budget_bytes=64 << 30, mode="batched")
    ids = [server.submit(d, o) for d, o in jobs]
    got = [None] * len(ids)

    def collect(i):
        got[i] = server.wait(ids[i], jobs[i][0].module_count)

    threads = [threading.Thread(target=collect, args=(i,)) for i in range(len(ids))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    server.close()
    spans = [(r["start_seconds"], r["finish_seconds"]) for _, r in got]
    # jobs ran together, and jobs joined while others were running (continuous batching)
    assert any(a0 < b1 and b0 < a1 for i, (a0, a1) in enumerate(spans) for (b0, b1) in spans[i + 1:])
    assert max(s for s, _ in spans) > min(f for _, f in spans) - 1e-3
    for (d, o), (c, rep) in zip(jobs, got):
        _same_as_alone(d, o, c, rep["result"])
        # the job's arrays were carved from one block sized by the footprint estimate
        assert rep["estimated_device_bytes"] >= rep["result"]["reserved_device_bytes"]


def test_configured_256_block_batch_is_bit_equal_to_single_runs():
    """The batching configuration at its size: 256 blocks of 500-50k cells (design.batch_job_specs,
    the bench's blocks), 100 iterations each, every other 64 x 64 block with split-BF16 products,
    all in one batched placement; each block must equal its own run bit for bit."""
    specs = design.batch_job_specs(256)
    jobs = []
    for k, (seed, cells, side) in enumerate(specs):
        d = design.served_design(seed, cells)
        d.center = hb.random_placement(d, seed)
        prec = abi.MP_DCT_BF16X3 if (side % 64 == 0 and k % 2 == 1) else abi.MP_DCT_FP32
        o = abi.default_options(target_density=0.8, bins_x=side, bins_y=side, iterations=100,
                                feasible_refinement_iterations=2 ** 31 - 1, seed=seed, dct_precision=prec)
        jobs.append((d, o))
    say(f"[gpu] configured batch: {len(jobs)} blocks, {sum(c for _, c, _ in specs)} cells, "
        f"grids {sorted(set(s for _, _, s in specs))}")
    out = hb.global_place_batch([d for d, _ in jobs], [o for _, o in jobs])
    say("  batch done; comparing with single runs")
    for (d, o), (c, r) in zip(jobs, out):
        assert r["completed_iterations"] == 100
        _same_as_alone(d, o, c, r)
