"""GPU detailed placement (four-cell window reordering) against the oracle's restatement of the
reference's device pass, and the legalizer hand-off after global placement.

Bar: bit-exact.  The window scores are FP64 sums in a fixed order with contraction off, the best
ordering is an argmin with the lowest index on ties, and every accept decision is re-scored on
the host in FP64: the HIP pass and the oracle must end on the same placement, bit for bit, with
the same counters.  The legalizer is the reference's own Abacus legalizer (oracle/_ref, built from
its source), run unchanged on the CPU between global and detailed placement.
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


@pytest.fixture(scope="module")
def ref(ob):
    if ob.reference() is None:
        pytest.skip("the reference checker oracle/_ref is not built here")
    return ob


COUNTERS = ("evaluated_windows", "evaluated_permutations", "accepted_operations", "completed_passes")


def same_as_oracle(ob, d, options, center=None, orientation=None, row_orientation=None):
    c, o, r = hb.detailed_place(d, options, center, orientation, row_orientation)
    oc, oo, orr = ob.detailed_place(d, options, center, orientation, row_orientation)
    assert np.array_equal(c, oc), np.abs(c - oc).max()
    assert np.array_equal(o, oo)
    for k in COUNTERS:
        assert r[k] == orr[k], (k, r[k], orr[k])
    assert r["hpwl_before"] == orr["hpwl_before"] and r["hpwl_after"] == orr["hpwl_after"]
    return c, o, r


def test_reordering_fixture_on_the_device(ob):
    """tests/DetailedPlacementTests.cpp:55-70 on the device: 24 orderings, one accepted reorder,
    HPWL 20 -> 10."""
    from dp_designs import reordering_fixture
    d = reordering_fixture()
    say("[gpu] detailed placement: reference fixture")
    _, _, r = same_as_oracle(ob, d, abi.default_dp_options(passes=1))
    assert r["accepted_operations"] == 1 and r["evaluated_permutations"] == 24
    assert r["hpwl_after"] == 10.0 and r["device_used"] == 0


@pytest.mark.parametrize("seed,max_degree", [(1, 64), (2, 64), (3, 6), (4, 64)])
def test_window_reordering_bit_equal_to_oracle(ob, seed, max_degree):
    """Abutted rows with random nets (windows share nets, so the pass-start scores and the
    re-scoring differ); a small degree limit sends most windows down the exhaustive host path."""
    from dp_designs import abutted_design
    d = abutted_design(seed=seed, rows=30, sites=300)
    say(f"[gpu] detailed placement: abutted design seed {seed}, M={d.module_count}, max degree {max_degree}")
    _, _, r = same_as_oracle(ob, d, abi.default_dp_options(passes=3, maximum_net_degree=max_degree))
    assert r["accepted_operations"] > 0 and r["hpwl_after"] < r["hpwl_before"]


def test_row_orientation_taken_by_reordered_cells(ob):
    from dp_designs import abutted_design
    d = abutted_design(seed=5, rows=10, sites=200)
    rows = np.array([5 if k % 2 else 0 for k in range(d.rows.shape[0])], dtype=np.int32)  # FS on odd rows
    say("[gpu] detailed placement: row orientations")
    _, o, r = same_as_oracle(ob, d, abi.default_dp_options(), row_orientation=rows)
    assert r["accepted_operations"] > 0 and np.any(o == 5)


def test_options_the_device_pass_declines(ob):
    from dp_designs import reordering_fixture
    d = reordering_fixture()
    with pytest.raises(ValueError):
        hb.detailed_place(d, abi.default_dp_options(passes=0))
    with pytest.raises(ValueError):
        hb.detailed_place(d, abi.default_dp_options(improvement_epsilon=float("nan")))
    with pytest.raises(hb.BackendUnavailable):
        hb.detailed_place(d, abi.default_dp_options(window_size=3))
    with pytest.raises(hb.BackendUnavailable):
        hb.detailed_place(d, abi.default_dp_options(method=abi.MP_DP_ADJACENT_SWAP))
    c, o, r = hb.detailed_place(d, abi.default_dp_options(method=abi.MP_DP_NONE))
    assert np.array_equal(c, d.center) and r["hpwl_after"] == r["hpwl_before"] == 20.0 and r["device_used"] == -1


def test_global_then_legalize_then_detailed(ref):
    """The flow after the GPU global placement: the reference's Abacus legalizer, unchanged on the
    CPU, then the device window reordering.  On the 20k-cell design: legal after both steps, the
    detailed pass bit-equal to the oracle's and never worse than its start.  (250 rows x 500 sites:
    the cells fill 68 % of the free core.)"""
    ob = ref
    d = design.large_design(seed=3, cells=20000, nets=21000, row_count=250, sites=500, macros=6)
    d.center = hb.random_placement(d, 3)
    o = abi.default_options(target_density=0.8, bins_x=128, bins_y=128)
    say("[gpu] GP -> legalize -> detailed placement, 20k cells")
    c, r, _ = hb.global_place(d, o)
    lc, lo, rep = ob.ref_legalize(d, c)
    say(f"  GP HPWL {r['hpwl_after']:.1f} overflow {r['overflow_after']:.4f}; legalized: legal {rep['legal']}, "
        f"total displacement {rep['standard_cell_total_displacement']:.1f}")
    assert rep["legal"] == 1.0 and rep["unplaced_standard_cells"] == 0.0
    h_legal, m_legal = ob.design_metrics(d, lc, 128, 128, 0.8)
    dc, do, dr = same_as_oracle(ob, d, abi.default_dp_options(), lc, lo)
    say(f"  legalized HPWL {h_legal:.1f} overflow {m_legal['normalized_overflow']:.4f}; detailed HPWL "
        f"{dr['hpwl_after']:.1f} ({dr['accepted_operations']} reorders, {dr['elapsed_seconds']:.3f} s)")
    assert abs(dr["hpwl_before"] - h_legal) < 1e-9 * h_legal
    assert dr["hpwl_after"] < dr["hpwl_before"]
    assert ob.ref_legality(d, dc, do)["legal"] == 1.0


def test_legalized_quality_matches_oracle_flow(ref):
    """Post-legalization HPWL and overflow, the final quality check after GP: the HIP global
    placement and the FP64 oracle's, each followed by the reference's legalizer, as sixteen-run
    ensembles from 1e-9-perturbed starts (the adaptive controller amplifies last-bit differences):
    mean legalized HPWL within 0.5 %, mean legalized overflow within 0.005."""
    ob = ref
    d = design.plain_design()
    d.center = hb.random_placement(d, 2026)
    o = abi.default_options(target_density=0.9, bins_x=64, bins_y=64)
    rng = np.random.default_rng(0)
    mov = d.movable_ids()
    say("[gpu] legalized quality, plain-64, 16-run ensembles")
    hq, oq = [], []
    for k in range(16):
        c0 = d.center.copy()
        c0[mov] += rng.normal(0.0, 1e-9, size=c0[mov].shape)
        for place, out in ((hb.global_place, hq), (ob.global_place, oq)):
            c, _, _ = place(d, o, center=c0)
            lc, _, rep = ob.ref_legalize(d, c)
            assert rep["legal"] == 1.0
            h, m = ob.design_metrics(d, lc, 64, 64, 0.9)
            out.append((h, m["normalized_overflow"]))
    hq, oq = np.array(hq), np.array(oq)
    say("  hip legalized mean", hq.mean(0), "oracle legalized mean", oq.mean(0))
    assert abs(hq[:, 0].mean() / oq[:, 0].mean() - 1.0) < 5e-3
    assert abs(hq[:, 1].mean() - oq[:, 1].mean()) < 5e-3
