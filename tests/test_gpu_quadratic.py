"""GPU parity of the quadratic initial placement (InitialPlacer.cpp:266-402) against the oracle.

Both sides assemble the same bound-to-bound system in the same order and run the same BiCGSTAB
recurrences; only the order of the FP64 dot-product sums differs, so the solutions agree to the
stated tolerance (absolute, in placement units) rather than bit for bit.
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


CASES = [
    ("plain", lambda: design.plain_design()),
    ("mid", lambda: design.large_design(seed=3, cells=20000, nets=21000, row_count=200, sites=400, macros=6)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_converged_quadratic_placement_matches_oracle(ob, name, make):
    """Solves run to a relative residual of 1e-12: the two results agree to 1e-6 placement units."""
    d = make()
    o = abi.default_qp_options(outer_iterations=2, solver_iterations=4000, tolerance=1e-12)
    say(f"[gpu] converged quadratic placement {name}: {d.module_count} modules")
    c_h, r_h = hb.quadratic_placement(d, o)
    c_o, r_o = ob.quadratic_placement(d, o)
    say(f"  hip {r_h}\n  oracle {r_o}\n  max |diff| {np.abs(c_h - c_o).max():.3e}")
    assert r_h["iterations"] == r_o["iterations"] == 2
    assert max(r_h["relative_residual"]) <= 1e-12 and max(r_o["relative_residual"]) <= 1e-12
    assert np.abs(c_h - c_o).max() < 1e-6
    fixed = (d.is_fixed != 0) | (d.is_terminal != 0)
    assert np.array_equal(c_h[fixed], d.center[fixed])


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_default_quadratic_placement_matches_oracle(ob, name, make):
    """Default options (5 outer iterations, at most 100 BiCGSTAB iterations per solve, tolerance
    1e-5).  Where every solve reaches the tolerance, each side stops where its own residual crosses
    it, so the iterates differ by what the tolerance leaves open: HPWL after placement within 1e-5,
    positions within 0.5 units.  Where the iteration limit stops the solves first (the mid design),
    the unconverged Krylov iterates depend on the rounding of every dot product, as they would
    between the reference's vectorised sums and any other order: HPWL after placement within 0.2 %
    and final residuals within a factor 3 of the oracle's."""
    d = make()
    say(f"[gpu] default quadratic placement {name}: {d.module_count} modules")
    c_h, r_h = hb.quadratic_placement(d)
    c_o, r_o = ob.quadratic_placement(d)
    say(f"  hip {r_h}\n  oracle {r_o}\n  max |diff| {np.abs(c_h - c_o).max():.3e}")
    assert r_h["iterations"] == r_o["iterations"] == 5
    converged = max(r_o["relative_residual"]) <= 1e-5 and max(r_h["relative_residual"]) <= 1e-5
    if converged:
        assert abs(r_h["hpwl_after"] / r_o["hpwl_after"] - 1.0) < 1e-5
        assert np.abs(c_h - c_o).max() < 0.5
    else:
        assert r_h["solver_iterations"] == r_o["solver_iterations"] == 1000
        assert abs(r_h["hpwl_after"] / r_o["hpwl_after"] - 1.0) < 2e-3
        for a in range(2):
            assert r_h["relative_residual"][a] < 3.0 * r_o["relative_residual"][a]


def test_quadratic_one_cell_closed_form_on_device():
    m = 2
    d = design.Design(width=np.ones(m), height=np.ones(m), center=np.array([[10.0, 10.0], [80.0, 30.0]]),
                      orientation=np.zeros(m, dtype=np.int32), is_terminal=np.array([0, 1], dtype=np.uint8),
                      is_fixed=np.array([0, 1], dtype=np.uint8), net_weight=np.ones(1),
                      net_pin_start=np.array([0, 2]), pin_module=np.array([0, 1], dtype=np.int32),
                      pin_offset=np.zeros((2, 2)), rows=np.array([[0.0, 40.0, 0.0, 1.0, 100.0]]))
    say("[gpu] quadratic placement one cell")
    c, r = hb.quadratic_placement(d, abi.default_qp_options(outer_iterations=1, tolerance=1e-14))
    x0, y0 = 50.0, 20.0
    wx, wy = 2.0 / abs(x0 - 80.0), 2.0 / abs(y0 - 30.0)
    assert abs(c[0, 0] - (wx * 80.0 + 1e-4 * x0) / (wx + 1e-4)) < 1e-10
    assert abs(c[0, 1] - (wy * 30.0 + 1e-4 * y0) / (wy + 1e-4)) < 1e-10


def test_headline_quadratic_start_is_the_committed_one():
    """The 210k-cell design's quadratic start, which bench.py times from and hands to the CPU
    baseline, reproduces the committed bytes (tests/golden/cfg2_quadratic_start.json)."""
    import hashlib
    import json
    import os
    ref = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg2_quadratic_start.json")))
    d = design.large_design(seed=2026)
    c, r = hb.quadratic_placement(d)
    assert hashlib.sha256(np.ascontiguousarray(c, dtype="<f8").tobytes()).hexdigest() == ref["centers_sha256"]
    assert r["hpwl_after"] == ref["hpwl_after"] and r["solver_iterations"] == ref["solver_iterations"]
