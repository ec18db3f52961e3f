"""Composition checks of the adaptive controller along a run's history (test helper, not a test module).

The scalar laws themselves (smoothingFromOverflow, updateDensityPenalty, considerCheckpoint) are
pinned by known-answer tests in test_oracle.py; these checks verify, row by row, that a run --
the oracle's or the device's -- feeds them the right metrics in the right places
(runAdaptiveGlobalPlacement, AdaptiveGlobalPlacer.cpp:652-953):

  objective       each row's objective is smooth wirelength + (penalty used) x energy
  smoothing       row k's smoothing is smoothingFromOverflow of row k-1's optimizer overflow against
                  the initial overflow (:741-742, :456-462); row 1 uses the initial smoothing
  penalty         after an accepted row, updateDensityPenalty(previous penalty, reference penalty,
                  the state before the step = the previous row, the accepted state = this row)
                  (:894-897); after a rejected row, x increase clamped to [lower, maximum option]
                  when the optimizer overflow exceeds its bound, else unchanged (:916-921)
  checkpoint      considerCheckpoint over the start and the accepted rows gives each row's new-best
                  flag and the restored iteration (:713-714, :906-907, :930-935)
  restarts        the restart flags sum to the restart count; no row after a rejected one restarts;
                  a row whose two predecessors both grew HPWL beyond the threshold restarts (:747-800)
  displacement    accepted moves stay within maximum_movement_in_bins x the smaller bin side (:796-800)
"""

import ctypes as C

import numpy as np

from myplacement_amd import abi
from oracle import bindings as ob


def check_history_composition(o: abi.MpOptions, result: dict, hist: list, bin_w: float, bin_h: float) -> dict:
    lib = ob.oracle()
    n = len(hist)
    assert n == result["completed_iterations"] and n >= 3
    pen = [h["penalty"] for h in hist]
    ref = pen[0]
    maxbin, minbin = max(bin_w, bin_h), min(bin_w, bin_h)
    lo = max(o.minimum_density_penalty, 0.02 * ref)
    hi = max(o.maximum_density_penalty, 10.0 * lo)
    stats = dict(accepted=0, rejected=0, penalty_checks=0, smoothing_checks=0, restarts=0)

    # objective
    for h in hist:
        assert h["objective"] == h["smooth_wirelength"] + h["penalty"] * h["density_energy"], h["iteration"]

    # smoothing: row 1 at the initial smoothing; later rows from the previous row's optimizer overflow
    assert hist[0]["smoothing"] == o.initial_smoothing * maxbin
    ovf0 = None
    for k in range(1, n):
        ratio = (hist[k]["smoothing"] / maxbin - o.final_smoothing) / (o.initial_smoothing - o.final_smoothing)
        if 1e-6 < ratio < 1.0 - 1e-6 and hist[k - 1]["overflow"] > 1e-6:
            ovf0 = hist[k - 1]["overflow"] / ratio ** (1.0 / o.smoothing_overflow_exponent)
            break
    assert ovf0 is not None, "no row anneals the smoothing"
    for k in range(1, n):
        want = lib.orc_smoothing(C.byref(o), maxbin, hist[k - 1]["overflow"], ovf0)
        assert abs(hist[k]["smoothing"] - want) <= 1e-9 * want, (k + 1, hist[k]["smoothing"], want)
        stats["smoothing_checks"] += 1

    # penalty
    for k in range(n - 1):
        h = hist[k]
        assert lo <= pen[k] <= hi, (k + 1, pen[k], lo, hi)
        if h["accepted"]:
            stats["accepted"] += 1
            if k == 0:
                continue  # the state before row 1 is the initial evaluation, not in the history
            b = hist[k - 1]
            want = lib.orc_update_penalty(C.byref(o), pen[k], ref, b["hpwl"], b["overflow"], h["hpwl"], h["overflow"],
                                          h["design_overflow"])
        else:
            stats["rejected"] += 1
            if h["overflow"] > o.maximum_bin_overflow:
                want = min(max(pen[k] * o.density_penalty_increase, lo), o.maximum_density_penalty)
            else:
                want = pen[k]
        assert pen[k + 1] == want, (k + 2, pen[k + 1], want, h["accepted"])
        stats["penalty_checks"] += 1

    # checkpoint: the start (iteration 0) and every accepted row
    its = [0] + [h["iteration"] for h in hist if h["accepted"]]
    hp = np.array([result["hpwl_before"]] + [h["hpwl"] for h in hist if h["accepted"]])
    ov = np.array([result["overflow_before"]] + [h["design_overflow"] for h in hist if h["accepted"]])
    flags = np.zeros(len(its), dtype=np.int32)
    feas = C.c_int32()
    chosen = lib.orc_checkpoint_sequence(C.byref(o), len(its), hp.ctypes.data_as(C.POINTER(C.c_double)),
                                         ov.ctypes.data_as(C.POINTER(C.c_double)), flags.ctypes.data_as(C.c_void_p),
                                         C.byref(feas))
    assert its[chosen] == result["best_checkpoint_iteration"], (its[chosen], result["best_checkpoint_iteration"])
    accepted_rows = [h for h in hist if h["accepted"]]
    for f, h in zip(flags[1:], accepted_rows):
        assert bool(f) == bool(h["best_checkpoint"]), h["iteration"]
    assert all(not h["best_checkpoint"] for h in hist if not h["accepted"])
    assert result["best_checkpoint_hpwl"] == hp[chosen]

    # restarts
    assert sum(h["momentum_restarted"] for h in hist) == result["momentum_restarts"]
    stats["restarts"] = result["momentum_restarts"]
    assert not hist[0]["momentum_restarted"]
    streak = 0
    for k, h in enumerate(hist):
        if streak >= 2:
            assert h["momentum_restarted"], (k + 1, streak)
        if k > 0 and not hist[k - 1]["accepted"]:
            assert not h["momentum_restarted"], k + 1
        if h["momentum_restarted"]:
            streak = 0
        if h["accepted"]:
            before = hist[k - 1]["hpwl"] if k > 0 else result["hpwl_before"]
            growth = (h["hpwl"] - before) / max(before, 1.0)
            streak = streak + 1 if growth > o.hpwl_growth_restart_threshold else 0
        else:
            streak = 0

    # displacement bound of accepted moves
    # (the move is measured as |base - new| after both were rounded at the positions' magnitude)
    cap = o.maximum_movement_in_bins * minbin * (1.0 + 1e-9)
    for h in hist:
        if h["accepted"]:
            assert 0.0 < h["maximum_displacement"] <= cap, (h["iteration"], h["maximum_displacement"], cap)
        else:
            assert h["step_size"] == 0.0 and h["maximum_displacement"] == 0.0
    return stats
