"""GPU parity: the HIP backend against the FP64 oracle on the same inputs.

Tolerances (stated here, checked below):
  HPWL, density metrics        relative 1e-10 (FP64 on both sides; only summation order differs)
  smooth wirelength            relative 2e-6  (exp / log in FP32 per pin, sums per net in FP32)
  electrostatic energy         relative 2e-5  (FP32 DCT products on MFMA; energy taken in the DCT basis)
  wire gradient                normwise relative 2e-5, elementwise 1e-4 of the largest entry
  density gradient / field     normwise relative 1e-5 (FP32 products: measured ~1e-6)
  end-to-end adaptive run      first iterations row by row at 1e-5; final HPWL within 0.5 % and final
                               design overflow within 0.005 as sixteen-run ensemble means
"""

import os
import sys

import numpy as np
import pytest

from myplacement_amd import abi, design
from myplacement_amd import backend as hb

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ob():
    from oracle import bindings
    return bindings


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def nrel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _setup(d, t, bins, seed=2026):
    bx, by = bins if isinstance(bins, tuple) else (bins, bins)
    d.center = hb.random_placement(d, seed)
    o = abi.default_options(target_density=t, bins_x=bx, bins_y=by)
    return d, o


def with_wide_nets(d, degrees, seed=5):
    """`d` with extra unit-weight nets of the given degrees appended, pins on distinct modules drawn
    uniformly, offsets uniform within the module: nets at the edges of the wirelength kernel's
    degree classes (8 / 9, 32 / 33 pins), at its register limit (512 / 513 pins) and far beyond it."""
    rng = np.random.default_rng(seed)
    M = d.module_count
    mods = [rng.choice(M, size=k, replace=False) for k in degrees]
    pm = np.concatenate([d.pin_module] + [m.astype(np.int32) for m in mods])
    extra = np.concatenate(mods)
    off = np.stack([(rng.random(extra.shape[0]) - 0.5) * d.width[extra],
                    (rng.random(extra.shape[0]) - 0.5) * d.height[extra]], axis=1)
    start = np.concatenate([d.net_pin_start, d.net_pin_start[-1] + np.cumsum(degrees)])
    return design.Design(width=d.width, height=d.height, center=d.center, orientation=d.orientation,
                         is_terminal=d.is_terminal, is_fixed=d.is_fixed,
                         net_weight=np.concatenate([d.net_weight, np.ones(len(degrees))]), net_pin_start=start,
                         pin_module=pm, pin_offset=np.concatenate([d.pin_offset, off]), rows=d.rows)


WIDE_DEGREES = [8, 9, 16, 17, 31, 32, 32, 33, 64, 65, 511, 512, 513, 1024, 2000]


def _cases():
    fx = os.path.join(ROOT, "tests", "golden", "bookshelf")
    return [
        ("tiny-8x8", lambda: (design.read_bookshelf(os.path.join(fx, "tiny.aux")), 0.85, 8)),
        ("mixed-8x8", lambda: (design.read_bookshelf(os.path.join(fx, "mixed.aux")), 0.85, 8)),
        ("plain-64", lambda: (design.plain_design(), 0.9, 64)),
        ("plain-48", lambda: (design.plain_design(seed=7), 0.9, 48)),
        ("plain-50x30", lambda: (design.plain_design(seed=8), 0.9, (50, 30))),
        ("plain-30x52", lambda: (design.plain_design(seed=9), 0.9, (30, 52))),
        ("plain-98x70", lambda: (design.plain_design(seed=11), 0.9, (98, 70))),
        ("wide-nets-64", lambda: (with_wide_nets(design.plain_design(seed=10), WIDE_DEGREES), 0.9, 64)),
        ("mid-128", lambda: (design.large_design(seed=3, cells=20000, nets=21000, row_count=200, sites=400,
                                                 macros=6), 0.8, 128)),
    ]


@pytest.mark.parametrize("name,make", _cases(), ids=[c[0] for c in _cases()])
def test_evaluate_matches_oracle(ob, name, make):
    d, t, bins = make()
    d, o = _setup(d, t, bins)
    bx, by = o.bins_x, o.bins_y
    orc = ob.OracleSession(d, o)
    L = orc.layout()
    say(f"[gpu] evaluate parity {name}: P={orc.particle_count} bins={bx}x{by}")
    be = hb.HipPlacementBackend.create(L)
    xy0 = orc.arrays()["positions"]
    rng = np.random.default_rng(1)
    region = L["region"]
    bw = (region[2] - region[0]) / bx
    for trial, gamma_bins in enumerate((0.75, 0.3, 0.12)):
        xy = xy0 + rng.normal(0.0, 2.0 * bw, size=xy0.shape) * (trial > 0)
        gamma = gamma_bins * max(bw, (region[3] - region[1]) / by)
        for grad in (True, False):
            r = orc.evaluate(xy, gamma, grad)
            h = be.evaluate(xy, gamma, grad)
            assert rel(h.hpwl, r["hpwl"]) < 1e-10
            assert rel(h.smooth_wirelength, r["smooth_wirelength"]) < 2e-6
            say(f"  gamma {gamma_bins} grad={grad}: smooth {rel(h.smooth_wirelength, r['smooth_wirelength']):.1e} "
                f"energy {rel(h.electrostatic_energy, r['electrostatic_energy']):.1e}")
            assert rel(h.electrostatic_energy, r["electrostatic_energy"]) < 2e-5, (h.electrostatic_energy, r["electrostatic_energy"])
            for key in ("optimizer_density", "design_density"):
                for m, v in r[key].items():
                    assert abs(h.__dict__[key][m] - v) <= 1e-10 * max(1.0, abs(v)), (key, m, h.__dict__[key][m], v)
            if grad:
                say(f"    wire gradient {nrel(h.wire_gradient, r['wire_gradient']):.1e} "
                    f"density gradient {nrel(h.density_gradient, r['density_gradient']):.1e}")
                assert nrel(h.wire_gradient, r["wire_gradient"]) < 2e-5
                assert np.abs(h.wire_gradient - r["wire_gradient"]).max() <= 1e-4 * np.abs(r["wire_gradient"]).max()
                assert nrel(h.density_gradient, r["density_gradient"]) < 1e-5, nrel(h.density_gradient, r["density_gradient"])
            else:
                assert h.wire_gradient is None
    be.close()


@pytest.mark.parametrize("bx,by", [(64, 64), (128, 64), (48, 48), (50, 30), (30, 52), (98, 70), (128, 128)])
def test_fields_match_oracle(ob, bx, by):
    """Source, potential and field grids: half-length products (both sides multiples of 64,
    square and not), the dense products (48 x 48) and the dense products on grids whose width is not
    a multiple of four (the bin kernel's per-bin path, padded non-square products, with padding of
    14 and 26 cells at 50 x 30 and 98 x 70)."""
    d = design.plain_design()
    d.center = hb.random_placement(d, 2026)
    o = abi.default_options(target_density=0.9, bins_x=bx, bins_y=by)
    orc = ob.OracleSession(d, o)
    L = orc.layout()
    say(f"[gpu] field grids plain {bx}x{by}")
    be = hb.HipPlacementBackend.create(L)
    xy = orc.arrays()["positions"]
    gamma = 0.75 * 360 / 64
    s_ref = orc.evaluate(xy, gamma, True)
    s = be.evaluate(xy, gamma, True)
    assert rel(s.electrostatic_energy, s_ref["electrostatic_energy"]) < 2e-5
    g_ref = orc.grids()
    g = be.debug_grids(bx * by)
    assert nrel(g["source"], g_ref["source"]) < 1e-6
    for k in ("potential", "field_x", "field_y"):
        say(f"  {k}: normwise relative {nrel(g[k], g_ref[k]):.1e}")
        assert nrel(g[k], g_ref[k]) < 1e-5, (k, nrel(g[k], g_ref[k]))
    be.close()


def test_evaluate_is_bit_reproducible(ob):
    """No floating-point atomics anywhere: two evaluations at the same positions agree bit for bit
    (the density is accumulated in fixed point, every reduction has a fixed order)."""
    d, o = _setup(design.plain_design(), 0.9, 64)
    orc = ob.OracleSession(d, o)
    be = hb.HipPlacementBackend.create(orc.layout())
    xy = orc.arrays()["positions"]
    say("[gpu] determinism plain-64")
    a = be.evaluate(xy, 3.0, True)
    b = be.evaluate(xy, 3.0, True)
    assert a.hpwl == b.hpwl and a.smooth_wirelength == b.smooth_wirelength
    assert a.electrostatic_energy == b.electrostatic_energy
    assert a.optimizer_density == b.optimizer_density and a.design_density == b.design_density
    assert np.array_equal(a.wire_gradient, b.wire_gradient)
    assert np.array_equal(a.density_gradient, b.density_gradient)
    be.close()


def test_bad_arguments_raise():
    d, o = _setup(design.plain_design(), 0.9, 64)
    pp = hb.PreparedDesign(d, o)
    be = hb.HipPlacementBackend.create(pp.layout())
    say("[gpu] argument checks")
    xy = pp.initial_positions()
    with pytest.raises(ValueError):
        be.evaluate(xy, 0.0, True)
    with pytest.raises(ValueError):
        be.evaluate(xy, float("nan"), False)
    with pytest.raises(ValueError):
        be.evaluate(xy[:-1], 1.0, False)
    assert be.device() == 0
    assert be.reserved_bytes() > 0
    be.close()


def test_adaptive_run_follows_oracle_then_matches_ensemble(ob):
    """The adaptive controller amplifies last-bit differences: the oracle itself, started from
    positions perturbed by 1e-9, ends up to ~0.7 % apart in HPWL and ~0.008 in overflow on this
    design.  So the end state is compared as an ensemble (sixteen runs per side from the same
    1e-9-perturbed starts): mean final HPWL within 0.5 %, mean final overflow within 0.005.  The
    early trajectory, before any discrete decision can flip, is compared row by row."""
    d, o = _setup(design.plain_design(), 0.9, 64)
    o.feasible_refinement_iterations = 2 ** 31 - 1
    say("[gpu] adaptive run plain-64, 280 iterations, unperturbed")
    c_h, r_h, hist_h = hb.global_place(d, o)
    c_o, r_o, hist_o = ob.global_place(d, o)
    assert r_h["completed_iterations"] == 280 and len(hist_h) == 280
    assert rel(r_h["hpwl_before"], r_o["hpwl_before"]) < 1e-12
    assert r_h["accepted_iterations"] > 0 and r_h["restored_best_checkpoint"] == 1
    for row in hist_h:
        assert all(np.isfinite(v) for v in row.values())
    for k in range(5):
        assert rel(hist_h[k]["hpwl"], hist_o[k]["hpwl"]) < 1e-5, k
        assert abs(hist_h[k]["overflow"] - hist_o[k]["overflow"]) < 1e-5, k
    rng = np.random.default_rng(0)
    mov = d.movable_ids()
    hh, ho = [], []
    for k in range(16):
        c = d.center.copy()
        c[mov] += rng.normal(0.0, 1e-9, size=c[mov].shape)
        say(f"  ensemble member {k}")
        _, a, _ = hb.global_place(d, o, center=c)
        _, b, _ = ob.global_place(d, o, center=c)
        hh.append((a["hpwl_after"], a["overflow_after"]))
        ho.append((b["hpwl_after"], b["overflow_after"]))
    hh, ho = np.array(hh), np.array(ho)
    say("  hip mean", hh.mean(0), "std", hh.std(0), "oracle mean", ho.mean(0), "std", ho.std(0))
    assert abs(hh[:, 0].mean() / ho[:, 0].mean() - 1.0) < 5e-3
    assert abs(hh[:, 1].mean() - ho[:, 1].mean()) < 5e-3


@pytest.mark.parametrize("start", ["inside", "outside"])
def test_reused_row_order_is_bit_identical_to_rebuilding_it(start):
    """The density deposit may reuse a particle row order built at earlier positions, and the
    current and look-ahead evaluations run as one batched chain.  Neither may change a single bit:
    the fixed-point density sum does not depend on the order, and every other reduction has a fixed
    order.  So whole runs with the order rebuilt for every evaluation
    (reuse limit 0, evaluations one after another), the default limit, and a limit so large that
    the order is never rebuilt must agree exactly.  The "outside" start places some cells beyond
    the core, where the movement bounds do not hold until the first accepted step."""
    d, o = _setup(design.plain_design(), 0.9, 64)
    o.iterations = 120
    center = d.center.copy()
    if start == "outside":
        core = d.core_region()
        mov = d.movable_ids()
        center[mov[:40], 0] = core[2] + 5.0
        center[mov[40:80], 1] = core[1] - 3.0
    runs = []
    for rows in (0.0, 2.5, 1e6):
        o.density_order_reuse_rows = rows
        say(f"[gpu] row-order reuse {start} start, limit {rows}")
        c, r, hist = hb.global_place(d, o, center=center)
        runs.append((c, r, np.array([[row[k] for k in abi.HISTORY_COLUMNS] for row in hist])))
    for c, r, h in runs[1:]:
        assert r["completed_iterations"] == runs[0][1]["completed_iterations"]
        assert np.array_equal(h, runs[0][2])
        assert np.array_equal(c, runs[0][0])
    assert runs[0][1]["accepted_iterations"] > 10


def test_session_steps_and_history():
    d, o = _setup(design.plain_design(), 0.9, 64)
    o.iterations = 30
    pp = hb.PreparedDesign(d, o)
    say("[gpu] session stepping plain-64")
    s = hb.PlacementSession(pp, o)
    assert s.step(10) == 10
    assert s.step(100) == 20 and s.stopped
    c, r = s.finish()
    assert r["completed_iterations"] == 30 and len(s.history()) == 30
    assert r["gradient_evaluations"] >= 32
    core = d.core_region()
    mov = d.movable_ids()
    assert np.all(c[mov, 0] - d.width[mov] / 2 >= core[0] - 1e-9)
    assert np.all(c[mov, 0] + d.width[mov] / 2 <= core[2] + 1e-9)
    s.close()


def test_session_renumbering_is_invisible_outside():
    """The session renumbers particles on the device along a Z-order curve; positions come back in
    the prepared order, and its backend view evaluates the renumbered particles consistently with a
    backend built from the prepared layout."""
    d, o = _setup(design.plain_design(), 0.9, 64)
    pp = hb.PreparedDesign(d, o)
    say("[gpu] session renumbering plain-64")
    s = hb.PlacementSession(pp, o)
    order = s.particle_order()
    n_mov = int(pp.info.movable_count)
    assert sorted(order.tolist()) == list(range(len(order)))
    assert set(order[:n_mov].tolist()) == set(range(n_mov))
    xy = s.positions()
    assert np.array_equal(xy, pp.initial_positions())
    ref = hb.HipPlacementBackend.create(pp.layout())
    a = ref.evaluate(xy, 3.0, True)
    b = s.backend().evaluate(xy[order], 3.0, True)
    assert rel(b.hpwl, a.hpwl) < 1e-12
    assert rel(b.electrostatic_energy, a.electrostatic_energy) < 1e-6
    assert b.optimizer_density == a.optimizer_density
    assert nrel(b.wire_gradient, a.wire_gradient[order]) < 1e-6
    ref.close()
    s.close()


def test_exceeded_device_budget_is_unavailable_with_a_reason():
    """DeviceMemoryBudget (.cu:60-80, :1156-1216): a backend whose allocations exceed
    maximum_device_bytes is not created; the factory reports "unavailable" with the reason, and a
    budget that fits creates it with reservedBytes() within the budget."""
    d, o = _setup(design.plain_design(), 0.9, 64)
    pp = hb.PreparedDesign(d, o)
    L = pp.layout()
    say("[gpu] budget injection")
    be, reason = hb.try_create_hip_placement_backend(L, maximum_device_bytes=1 << 20)
    assert be is None
    assert "budget exceeded" in reason and "limit 1048576 bytes" in reason, reason
    be, reason = hb.try_create_hip_placement_backend(L, maximum_device_bytes=64 << 20)
    assert be is not None and reason == ""
    assert 0 < be.reserved_bytes() <= 64 << 20
    be.close()
    o.maximum_device_bytes = 1 << 20
    # the explicit hip backend throws when it cannot be created (AdaptiveGlobalPlacer.cpp:275-277)
    with pytest.raises(RuntimeError, match="budget exceeded"):
        hb.global_place(d, o)


@pytest.mark.parametrize("case", ["plain-default", "plain-aggressive", "mid-quadratic"])
def test_device_adaptive_run_composes_the_pinned_laws(ob, case):
    """The device controller's whole run, row by row, against the oracle's pinned laws
    (tests/controller_checks.py): objective, smoothing, penalty, checkpoint, restart and
    displacement composition of AdaptiveGlobalPlacer.cpp:652-953."""
    from controller_checks import check_history_composition
    if case.startswith("plain"):
        d, o = _setup(design.plain_design(), 0.9, 64)
        if case == "plain-aggressive":
            o.initial_movement_in_bins = 1.0
            o.maximum_movement_in_bins = 1.0
            o.armijo_coefficient = 0.5
            o.objective_increase_for_density = 0.0
    else:
        d = design.large_design(seed=3, cells=5000, nets=5250, row_count=80, sites=300, macros=3)
        d.center, _ = hb.quadratic_placement(d)
        o = abi.default_options(target_density=0.8, bins_x=32, bins_y=32, iterations=150)
    o.feasible_refinement_iterations = 2 ** 31 - 1
    say(f"[gpu] controller composition on the device: {case}")
    _, r, hist = hb.global_place(d, o)
    core = d.core_region()
    stats = check_history_composition(o, r, hist, (core[2] - core[0]) / o.bins_x, (core[3] - core[1]) / o.bins_y)
    say(f"  {stats}")
    assert stats["accepted"] > 10
    if case == "plain-aggressive":
        assert r["rejected_candidates"] > 0 and r["momentum_restarts"] > 0


@pytest.mark.parametrize("case", ["plain-64", "mid-128-quadratic"])
def test_graph_captured_iterations_equal_direct_launches(case):
    """capture_graphs = 1 launches each iteration's chain up to its read-back as one HIP graph (one
    executable graph per chain shape, updated in place): the runs must be bit-identical to direct
    launches, whatever shapes the chain takes (momentum look-ahead, row-order rebuilds)."""
    if case == "plain-64":
        d, o = _setup(design.plain_design(), 0.9, 64)
    else:
        d = design.large_design(seed=3, cells=20000, nets=21000, row_count=200, sites=400, macros=6)
        d.center, _ = hb.quadratic_placement(d)
        o = abi.default_options(target_density=0.8, bins_x=128, bins_y=128)
    o.iterations = 150
    o.feasible_refinement_iterations = 2 ** 31 - 1
    runs = []
    for cg in (0, 1):
        o.capture_graphs = cg
        say(f"[gpu] {case}: capture_graphs={cg}")
        c, r, hist = hb.global_place(d, o)
        runs.append((c, r, np.array([[row[k] for k in abi.HISTORY_COLUMNS] for row in hist])))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][2], runs[1][2])
    assert runs[0][1]["gradient_evaluations"] == runs[1][1]["gradient_evaluations"]


@pytest.mark.parametrize("case", ["plain-64", "plain-64-exponent-2", "plain-50x30-fresh-order", "mid-128-quadratic"])
def test_density_taken_over_equals_recomputed(case):
    """An iteration's gradient evaluation at the candidate accepted the iteration before takes that
    candidate's density, potential and energy (recompute_density = 0, the default) instead of
    computing them again (recompute_density = 1): smoothing and penalty, the only inputs that changed,
    enter neither, so the runs must be bit-identical.  Cases: the half-length DCT products (64); the
    same with smoothing exponent 2, where no wirelength goes ahead, so a momentum iteration's launch
    of the bin pass and the wirelength carries two wirelength evaluations and one bin pass; the
    dense products of a grid that is not a multiple of 64 with the row order rebuilt for every
    evaluation (the look-ahead then evaluated on its own); and a quadratic start at 128 bins."""
    if case == "plain-64":
        d, o = _setup(design.plain_design(), 0.9, 64)
    elif case == "plain-64-exponent-2":
        d, o = _setup(design.plain_design(), 0.9, 64)
        o.smoothing_overflow_exponent = 2.0
    elif case == "plain-50x30-fresh-order":
        d, o = _setup(design.plain_design(), 0.9, (50, 30))
        o.density_order_reuse_rows = 0.0
    else:
        d = design.large_design(seed=3, cells=20000, nets=21000, row_count=200, sites=400, macros=6)
        d.center, _ = hb.quadratic_placement(d)
        o = abi.default_options(target_density=0.8, bins_x=128, bins_y=128)
    o.iterations = 150
    o.feasible_refinement_iterations = 2 ** 31 - 1
    runs = []
    for rd in (1, 0):
        o.recompute_density = rd
        say(f"[gpu] {case}: recompute_density={rd}")
        c, r, hist = hb.global_place(d, o)
        runs.append((c, r, np.array([[row[k] for k in abi.HISTORY_COLUMNS] for row in hist])))
    assert runs[0][1]["accepted_iterations"] > 20
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][2], runs[1][2])
    for k in ("gradient_evaluations", "value_evaluations", "momentum_restarts", "rejected_candidates"):
        assert runs[0][1][k] == runs[1][1][k], k
