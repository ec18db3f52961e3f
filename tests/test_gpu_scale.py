"""GPU parity at the configured shapes: the HIP backend against the FP64 oracle on the 210k-cell /
512 x 512 design (plain FP32 products) and on the 1.3M-cell / 1024 x 1024 design (split-BF16
products), evaluate by evaluate, along a session, and at the end of a whole run.

Tolerances (stated here, checked below):
  HPWL, density metrics       relative 1e-10 (FP64 on both sides)
  smooth wirelength           relative 2e-6
  energy                      relative 2e-5 FP32 products, 5e-5 split-BF16 products
  wire gradient               normwise relative 2e-5
  potential, field, density   normwise relative 1e-5 FP32 products, 5e-5 split-BF16 products
  gradient
  end state of a whole run    final HPWL within 0.5 % and final design overflow within 0.005 of the
                              oracle's run from the same start (one run per side: at 210k cells the
                              chaotic spread between runs is far below these bars)
"""

import sys

import numpy as np
import pytest

from myplacement_amd import abi, design
from myplacement_amd import backend as hb

pytestmark = pytest.mark.gpu


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


@pytest.fixture(scope="module")
def headline():
    d = design.large_design(seed=2026)
    quad, _ = hb.quadratic_placement(d)
    return d, quad


@pytest.fixture(scope="module")
def precision_design():
    return design.large_design(seed=2026, cells=1_300_000, nets=1_400_000, row_count=3000, sites=3000, macros=20,
                               macro_fraction=0.12)


def _check_evaluate(orc, be, xy, gamma, bins, energy_tol, field_tol, label):
    for grad in (True, False):
        r = orc.evaluate(xy, gamma, grad)
        h = be.evaluate(xy, gamma, grad)
        e_err = rel(h.electrostatic_energy, r["electrostatic_energy"])
        say(f"  {label} grad={grad}: hpwl {rel(h.hpwl, r['hpwl']):.1e} smooth "
            f"{rel(h.smooth_wirelength, r['smooth_wirelength']):.1e} energy {e_err:.1e}")
        assert rel(h.hpwl, r["hpwl"]) < 1e-10
        assert rel(h.smooth_wirelength, r["smooth_wirelength"]) < 2e-6
        assert e_err < energy_tol, e_err
        for key in ("optimizer_density", "design_density"):
            for m, v in r[key].items():
                assert abs(h.__dict__[key][m] - v) <= 1e-10 * max(1.0, abs(v)), (key, m, h.__dict__[key][m], v)
        if grad:
            wn = nrel(h.wire_gradient, r["wire_gradient"])
            dn = nrel(h.density_gradient, r["density_gradient"])
            g_ref = orc.grids()
            g = be.debug_grids(bins * bins)
            errs = {k: nrel(g[k], g_ref[k]) for k in ("potential", "field_x", "field_y")}
            say(f"    wire gradient {wn:.1e} density gradient {dn:.1e} "
                + " ".join(f"{k} {v:.1e}" for k, v in errs.items()))
            assert wn < 2e-5
            assert dn < field_tol, dn
            assert nrel(g["source"], g_ref["source"]) < 1e-6
            for k, v in errs.items():
                assert v < field_tol, (k, v)


@pytest.mark.parametrize("start", ["quadratic", "random"])
def test_headline_evaluate_matches_oracle(ob, headline, start):
    """210k cells, 12 macros, 220k nets (~0.9M pins, nets up to 500 pins), 512 x 512 bins."""
    d, quad = headline
    d.center = quad if start == "quadratic" else hb.random_placement(d, 2026)
    o = abi.default_options(target_density=0.8, bins_x=512, bins_y=512)
    orc = ob.OracleSession(d, o)
    L = orc.layout()
    say(f"[gpu] headline evaluate parity, {start} start: P={orc.particle_count} pins={L['pin_particles'].shape[0]}")
    be = hb.HipPlacementBackend.create(L)
    xy = orc.arrays()["positions"]
    region = L["region"]
    bw = max(region[2] - region[0], region[3] - region[1]) / 512
    for gamma_bins in (0.75, 0.2):
        _check_evaluate(orc, be, xy, gamma_bins * bw, 512, 2e-5, 1e-5, f"gamma {gamma_bins} bins")
    be.close()


@pytest.mark.parametrize("precision", [abi.MP_DCT_BF16X3, abi.MP_DCT_FP32], ids=["bf16x3", "fp32"])
def test_precision_design_evaluate_matches_oracle(ob, precision_design, precision):
    """1.3M cells, 20 macros, 1.4M nets (~5.7M pins), 1024 x 1024 bins, from a random start."""
    d = precision_design
    d.center = hb.random_placement(d, 2026)
    o = abi.default_options(target_density=0.7, bins_x=1024, bins_y=1024)
    orc = ob.OracleSession(d, o)
    L = orc.layout()
    say(f"[gpu] 1.3M-cell evaluate parity ({'bf16x3' if precision else 'fp32'}): P={orc.particle_count}")
    be = hb.HipPlacementBackend.create(L, dct_precision=precision)
    xy = orc.arrays()["positions"]
    region = L["region"]
    bw = max(region[2] - region[0], region[3] - region[1]) / 1024
    tol = (5e-5, 5e-5) if precision == abi.MP_DCT_BF16X3 else (2e-5, 1e-5)
    _check_evaluate(orc, be, xy, 0.5 * bw, 1024, tol[0], tol[1], "gamma 0.5 bins")
    be.close()


def test_precision_session_grids_every_20_iterations(ob, precision_design):
    """The precision form's own check: a split-BF16 session on the 1.3M-cell design from the
    quadratic start; every 20 of its 280 iterations the potential and field at the session's
    positions are compared with the oracle's FP64 DCT at the same positions."""
    d = precision_design
    d.center, _ = hb.quadratic_placement(d)
    o = abi.default_options(target_density=0.7, bins_x=1024, bins_y=1024, dct_precision=abi.MP_DCT_BF16X3)
    o.feasible_refinement_iterations = 2 ** 31 - 1
    prep = hb.PreparedDesign(d, o)
    sess = hb.PlacementSession(prep, o)
    order = sess.particle_order()
    be = sess.backend()
    orc = ob.OracleSession(d, o)
    region = orc.layout()["region"]
    gamma = 0.5 * max(region[2] - region[0], region[3] - region[1]) / 1024
    say("[gpu] 1.3M-cell bf16x3 session, grids against the FP64 DCT every 20 iterations")
    worst = 0.0
    for block in range(14):
        sess.step(20)
        pos = sess.positions()
        s_ref = orc.evaluate(pos, gamma, True)
        s = be.evaluate(pos[order], gamma, True)
        g_ref = orc.grids()
        g = be.debug_grids(1024 * 1024)
        errs = {k: nrel(g[k], g_ref[k]) for k in ("potential", "field_x", "field_y")}
        e_err = rel(s.electrostatic_energy, s_ref["electrostatic_energy"])
        say(f"  iteration {20 * (block + 1)}: energy {e_err:.1e}, " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
        assert e_err < 5e-5
        for k, v in errs.items():
            assert v < 5e-5, (block, k, v)
        worst = max(worst, *errs.values())
    sess.close()
    say(f"  worst normwise grid error {worst:.1e}")


@pytest.fixture(scope="module")
def headline_oracle_run(ob, headline):
    """The oracle's whole 280-iteration adaptive run of the 210k-cell design from the random start
    (early stop off), shared by the device runs below."""
    d, _ = headline
    d.center = hb.random_placement(d, 2026)
    o = abi.default_options(target_density=0.8, bins_x=512, bins_y=512)
    o.feasible_refinement_iterations = 2 ** 31 - 1
    say("[gpu] headline random start: the oracle's 280 iterations (about a minute on one core)")
    _, b, hist_b = ob.global_place(d, o)
    say(f"  oracle: HPWL {b['hpwl_after']:.1f} overflow {b['overflow_after']:.4f} in {b['loop_seconds']:.1f}s")
    return d.center.copy(), o, b, hist_b


@pytest.mark.parametrize("precision", [abi.MP_DCT_FP32, abi.MP_DCT_BF16X3], ids=["fp32", "bf16x3"])
def test_headline_random_start_end_state_matches_oracle(ob, headline, headline_oracle_run, precision):
    """The whole 280-iteration adaptive run of the 210k-cell design from the random start, on the
    device (FP32 products, and the split-BF16 precision form) and in the oracle: final HPWL within
    0.5 %, final overflow within 0.005.  At this size one run per side suffices: the chaotic spread
    between runs from perturbed starts, ~0.5 % on the 2k-cell design, averages out over 210k cells."""
    from controller_checks import check_history_composition
    d, _ = headline
    center, o0, b, hist_b = headline_oracle_run
    o = abi.MpOptions.from_buffer_copy(o0)
    o.dct_precision = precision
    _, a, hist_a = hb.global_place(d, o, center=center)
    say(f"  device ({'bf16x3' if precision else 'fp32'}): HPWL {a['hpwl_after']:.1f} ({a['hpwl_after'] / b['hpwl_after'] - 1:+.3%}) "
        f"overflow {a['overflow_after']:.4f} ({a['overflow_after'] - b['overflow_after']:+.4f}) in {a['loop_seconds']:.3f}s")
    assert a["completed_iterations"] == b["completed_iterations"] == 280
    assert rel(a["hpwl_before"], b["hpwl_before"]) < 1e-12
    assert abs(a["hpwl_after"] / b["hpwl_after"] - 1.0) < 5e-3
    assert abs(a["overflow_after"] - b["overflow_after"]) < 5e-3
    # both runs compose the pinned controller laws row by row
    core = d.core_region()
    for r, hist in ((a, hist_a), (b, hist_b)):
        check_history_composition(o, r, hist, (core[2] - core[0]) / 512, (core[3] - core[1]) / 512)
