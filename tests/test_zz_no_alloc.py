"""The no-allocation detailed placement (mp_dp_run_noalloc) against mp_dp_run through the C ABI:
same centres, orientations and counters, bit for bit, on two calls in a row (the second on a
smaller design, so it runs over what the first left in the shared scratch), and after them a
design too large for the scratch is refused before anything runs."""

import sys

import numpy as np
import pytest

from myplacement_amd import abi, design
from myplacement_amd import backend as hb

pytestmark = pytest.mark.gpu

COUNTERS = ("evaluated_windows", "evaluated_permutations", "accepted_operations", "completed_passes", "hpwl_before",
            "hpwl_after", "device_used")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(d, options, row_orientation=None):
    a = hb.detailed_place(d, options, row_orientation=row_orientation)
    b = hb.detailed_place_noalloc(d, options, row_orientation=row_orientation)
    assert np.array_equal(a[0], b[0]), np.abs(a[0] - b[0]).max()
    assert np.array_equal(a[1], b[1])
    for k in COUNTERS:
        assert a[2][k] == b[2][k], (k, a[2][k], b[2][k])
    assert 0 < b[2]["reserved_device_bytes"] <= hb.noalloc_scratch_bytes()
    return b[2]


def test_noalloc_detailed_placement_equals_the_allocating_one():
    from dp_designs import abutted_design
    big = abutted_design(seed=11, rows=60, sites=600)
    small = abutted_design(seed=12, rows=10, sites=200)
    rows = np.array([5 if k % 2 else 0 for k in range(small.rows.shape[0])], dtype=np.int32)
    say(f"[gpu] no-alloc detailed placement, first call: M={big.module_count}")
    r1 = _same(big, abi.default_dp_options(passes=3))
    say(f"  {r1['accepted_operations']} reorders, scratch used {r1['reserved_device_bytes']} bytes")
    say(f"[gpu] no-alloc detailed placement, second call (smaller): M={small.module_count}")
    r2 = _same(small, abi.default_dp_options(passes=2), row_orientation=rows)
    say(f"  {r2['accepted_operations']} reorders, scratch used {r2['reserved_device_bytes']} bytes")
    assert r1["accepted_operations"] > 0 and r2["accepted_operations"] > 0
    assert r2["reserved_device_bytes"] < r1["reserved_device_bytes"]


def test_noalloc_refuses_a_design_beyond_its_scratch():
    d = design.large_design(seed=1, cells=1_300_000, nets=1_400_000, row_count=3000, sites=3000, macros=20)
    say("[gpu] no-alloc detailed placement: 1.3M-cell design must be refused")
    with pytest.raises(ValueError, match="too large for the no-allocation"):
        hb.detailed_place_noalloc(d, abi.default_dp_options())
