"""The placement service (mp_gp_server_*): admission against the memory budget, concurrent jobs,
and the same result for every job alone and in the mix (bit for bit: every evaluation is
deterministic and a job's kernels only touch its own buffers)."""

import sys
import threading

import numpy as np
import pytest

from myplacement_amd import abi, design
from myplacement_amd import backend as hb

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _job(seed, cells, iterations=60):
    d = design.served_design(seed, cells)
    d.center = hb.random_placement(d, seed)
    side = design.served_bins(cells)
    o = abi.default_options(target_density=0.8, bins_x=side, bins_y=side, iterations=iterations,
                            feasible_refinement_iterations=1, seed=seed)
    return d, o


def test_results_are_the_same_alone_and_in_the_mix():
    jobs = [_job(100 + k, c) for k, c in enumerate((1500, 4000, 2500, 9000, 1200))]
    say("[gpu] served mix of 5 jobs, 3 slots")
    server = hb.PlacementServer(max_active_jobs=3, memory_budget_bytes=64 << 30)
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
    # at least two jobs ran at the same time
    spans = [(r["start_seconds"], r["finish_seconds"]) for _, r in got]
    assert any(a0 < b1 and b0 < a1 for i, (a0, a1) in enumerate(spans) for (b0, b1) in spans[i + 1:])
    for (d, o), (c_mix, r_mix) in zip(jobs, got):
        say(f"  job {d.module_count} modules: alone vs mix")
        c_alone, r_alone, _ = hb.global_place(d, o)
        assert np.array_equal(c_mix, c_alone)
        for k in ("hpwl_after", "overflow_after", "completed_iterations", "accepted_iterations", "gradient_evaluations"):
            assert r_mix["result"][k] == r_alone[k], k
        assert r_mix["estimated_device_bytes"] >= r_mix["result"]["reserved_device_bytes"]


@pytest.mark.parametrize("mode", ["streams", "batched"])
def test_admission_waits_for_memory(mode):
    a, b = _job(7, 3000), _job(8, 3000)
    est = max(hb.estimate_device_bytes(a[0], a[1]), hb.estimate_device_bytes(b[0], b[1]))
    say(f"[gpu] admission under a budget that fits one job ({mode})")
    server = hb.PlacementServer(max_active_jobs=4, memory_budget_bytes=int(1.5 * est), mode=mode)
    ja, jb = server.submit(*a), server.submit(*b)
    _, ra = server.wait(ja, a[0].module_count)
    _, rb = server.wait(jb, b[0].module_count)
    assert rb["start_seconds"] >= ra["finish_seconds"] - 1e-6
    with pytest.raises(ValueError):
        server.submit(*_job(9, 200000))
    server.close()


@pytest.mark.parametrize("mode", ["batched", "streams"])
def test_served_sizes_1k_to_200k_are_the_same_alone_and_in_the_mix(mode):
    """Jobs across the served stream's whole size range (1k-200k cells, 32 x 32 to 256 x 256 bins),
    each run until design overflow <= 0.10 or 280 iterations as in the stream: every result equal
    alone and in the mix."""
    jobs = []
    for k, (cells, t) in enumerate(((1000, 0.7), (200000, 0.9), (40000, 0.8), (8000, 0.75), (120000, 0.85),
                                    (2500, 0.9))):
        d, o = _job(400 + k, cells, iterations=280)
        o.target_density = t
        jobs.append((d, o))
    say(f"[gpu] served sizes 1k-200k, {mode} mode, 4 slots")
    server = hb.PlacementServer(max_active_jobs=4, memory_budget_bytes=200 << 30, mode=mode)
    ids = [server.submit(d, o) for d, o in jobs]
    got = [server.wait(j, d.module_count) for j, (d, _) in zip(ids, jobs)]
    server.close()
    for (d, o), (c_mix, r_mix) in zip(jobs, got):
        c_alone, r_alone, _ = hb.global_place(d, o)
        say(f"  job {d.module_count} modules, {o.bins_x}x{o.bins_y} bins: {r_alone['completed_iterations']} "
            f"iterations, overflow {r_alone['overflow_after']:.3f}")
        assert np.array_equal(c_mix, c_alone)
        for k in ("hpwl_after", "overflow_after", "completed_iterations", "accepted_iterations", "gradient_evaluations"):
            assert r_mix["result"][k] == r_alone[k], k
