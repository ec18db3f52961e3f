"""The MyPlacement-side adapter (myplacement_amd/integration/HipPlacementBackend.cpp) compiled
against the reference's own headers and sources (oracle/_ref/libmpadapter.so) and driven as
GlobalPlacementSession drives a CudaPlacementBackend (AdaptiveGlobalPlacer.cpp:114-126, :271-331).

CPU: the factory returns nullptr with a reason when no device is usable (the "auto falls back to the
CPU" path of :275-277), and the adapter's flattening of the reference's PlacementDatabase equals,
bit for bit, the layout the product prepares from the same design and the oracle's.
GPU: evaluate through the adapter matches the FP64 oracle at the parity tolerances of
test_gpu_parity.py, the interface's error behaviour holds, and an exceeded device budget gives
nullptr plus the budget reason.
"""

import os
import sys

import numpy as np
import pytest

from myplacement_amd import abi, design
from myplacement_amd import backend as hb
from oracle import bindings as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOOK = os.path.join(ROOT, "tests", "golden", "bookshelf")

LAYOUT_KEYS = ("particle_width", "particle_height", "particle_charge_area", "particle_layer", "net_offsets",
               "net_weights", "pin_particles", "pin_offsets_or_static", "fixed_area", "dark_area")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def adapter():
    if ob.adapter() is None:
        pytest.skip("the reference-side adapter oracle/_ref/libmpadapter.so is not built here")
    return ob.adapter()


def _gpu_present() -> bool:
    # the kernel driver's device node, not torch: torch brings its own HIP runtime, which must not
    # be loaded beside the one the product library links
    return os.path.exists("/dev/kfd")


def _aux_cases(tmp_path_factory):
    out = [("tiny", os.path.join(BOOK, "tiny.aux"), 8, 0.85), ("mixed", os.path.join(BOOK, "mixed.aux"), 8, 0.85)]
    d = design.plain_design(seed=31)
    d.center = ob.random_placement(d, 31)
    out.append(("plain", design.write_bookshelf(d, str(tmp_path_factory.mktemp("plain")), "plain"), 64, 0.9))
    d = design.large_design(seed=32, cells=12000, nets=12600, row_count=150, sites=300, macros=4)
    d.center = ob.random_placement(d, 32)
    out.append(("macro-12k", design.write_bookshelf(d, str(tmp_path_factory.mktemp("m12k")), "m12k"), 64, 0.8))
    return out


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    return _aux_cases(tmp_path_factory)


def test_adapter_layout_equals_the_product_and_oracle_layouts(adapter, cases):
    for name, aux, bins, t in cases:
        run = ob.AdapterRun(aux, bins, t, device=-1)  # no device needed: the flattening comes first
        a = run.layout()
        d = design.read_bookshelf(aux)
        o = abi.default_options(target_density=t, bins_x=bins, bins_y=bins)
        p = hb.PreparedDesign(d, o).layout()
        q = ob.OracleSession(d, o).layout()
        for k in LAYOUT_KEYS:
            assert np.array_equal(a[k], p[k]), (name, k)
            assert np.array_equal(a[k], q[k]), (name, k)
        assert a["region"] == p["region"] == q["region"] and (a["bins_x"], a["bins_y"]) == (bins, bins)
        run.close()


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device path")
def test_factory_without_a_device_returns_null_with_a_reason(adapter):
    run = ob.AdapterRun(os.path.join(BOOK, "tiny.aux"), 8, 0.85, device=0)
    assert run.status == 3 and not run.created
    assert "HIP device discovery failed" in run.reason, run.reason
    assert run.device() == -1 and run.reserved_bytes() == 0
    run.close()


def test_factory_with_an_unknown_device_returns_null_with_a_reason(adapter):
    run = ob.AdapterRun(os.path.join(BOOK, "mixed.aux"), 8, 0.85, device=-1)
    assert run.status == 3 and not run.created
    assert "HIP" in run.reason and ("does not exist" in run.reason or "discovery failed" in run.reason), run.reason
    run.close()


@pytest.mark.gpu
def test_adapter_evaluate_matches_oracle(adapter, cases):
    for name, aux, bins, t in cases:
        run = ob.AdapterRun(aux, bins, t, device=0)
        assert run.created, run.reason
        assert run.device() == 0 and run.reserved_bytes() > 0
        d = design.read_bookshelf(aux)
        orc = ob.OracleSession(d, abi.default_options(target_density=t, bins_x=bins, bins_y=bins))
        xy0 = run.positions()
        assert np.array_equal(xy0, orc.arrays()["positions"])
        region = run.layout()["region"]
        bw = max(region[2] - region[0], region[3] - region[1]) / bins
        rng = np.random.default_rng(2)
        say(f"[gpu] adapter evaluate {name}: P={xy0.shape[0]}")
        for trial, gb in enumerate((0.75, 0.25)):
            xy = xy0 + rng.normal(0.0, bw, size=xy0.shape) * trial
            for grad in (True, False):
                h = run.evaluate(xy, gb * bw, grad)
                r = orc.evaluate(xy, gb * bw, grad)
                assert abs(h["hpwl"] - r["hpwl"]) <= 1e-10 * r["hpwl"]
                assert abs(h["smooth_wirelength"] - r["smooth_wirelength"]) <= 2e-6 * r["smooth_wirelength"]
                assert abs(h["electrostatic_energy"] - r["electrostatic_energy"]) <= 2e-5 * r["electrostatic_energy"]
                for key in ("optimizer_density", "design_density"):
                    for m, v in r[key].items():
                        assert abs(h[key][m] - v) <= 1e-10 * max(1.0, abs(v)), (name, key, m)
                if grad:
                    for g in ("wire_gradient", "density_gradient"):
                        err = np.linalg.norm(h[g] - r[g]) / max(np.linalg.norm(r[g]), 1e-300)
                        assert err < (2e-5 if g == "wire_gradient" else 1e-5), (name, g, err)
        # the interface's argument errors (.cu:923-928): std::invalid_argument
        with pytest.raises(ValueError):
            run.evaluate(xy0[:-1], bw, True)
        with pytest.raises(ValueError):
            run.evaluate(xy0, 0.0, True)
        run.close()


@pytest.mark.gpu
def test_adapter_factory_reports_an_exceeded_budget(adapter):
    run = ob.AdapterRun(os.path.join(BOOK, "mixed.aux"), 8, 0.85, device=0, maximum_device_bytes=4096)
    assert run.status == 3 and not run.created
    assert "budget exceeded" in run.reason, run.reason
    run.close()
