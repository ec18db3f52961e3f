"""The FP64 oracle, pinned before it is trusted.

Parity status of each oracle function (oracle/gp_oracle.cpp), and what pins it:
  REDFT10 / REDFT01 / RODFT01        scipy.fft golden vectors (tests/golden/neumann_golden.npz);
                                     direct O(N^2) and FFT forms agree               -> pinned
  Neumann solve (ElectrostaticField) analytic cos-cos modes of tests/DensityFieldTests.cpp:12-100
                                     at 2e-12; uniform source -> zero (:102-111); scipy golden fields -> pinned
  parity sampling                    zero normal field on all four edges (:53-66); bin-centre
                                     sampling returns the grid value                 -> pinned
  DensityMap                         2x2 known answer (tests/DensityModelTests.cpp:10-52); the
                                     reference's own Metrics.cpp on random rectangles; mixed
                                     normalisation area 540 (PlacementFlowTests.cpp:305-310)  -> pinned
  static density, fillers            bit-equal to the reference's GlobalPlacementSupport.cpp    -> pinned
  design HPWL / overflow             the reference's calculateHpwl / calculateDensity           -> pinned
  smooth wirelength + gradient       two-pin closed form, gamma -> 0 limit (HPWL), central
                                     finite differences                              -> pinned
  electrostatic energy               equals the DCT-basis (Parseval) sum computed with scipy    -> pinned
  random initial placement           mt19937_64 10000th output (C++ standard), and the product's
                                     independent implementation bit-equal             -> pinned
  quadratic initial placement        one-cell closed form; one outer iteration against a dense
                                     system built here from the net list and solved by numpy
                                     (LAPACK); BiCGSTAB on a random SPD system against numpy;
                                     the reference's own invariants (PlacementFlowTests.cpp:121-129)
                                                                                      -> pinned
  controller laws                    estimateStep = 1/L on quadratics and within the inverse
                                     extreme curvatures; Armijo acceptance iff tL <= 2(1-c) on a
                                     quadratic line search; smoothing end points; updateDensityPenalty
                                     and considerCheckpoint on known-answer tables (every branch,
                                     both clamps, the tie rules)                      -> pinned
  adaptive loop composition          every row of whole runs (default and the reference's aggressive
                                     options, random and quadratic starts) re-derived from the pinned
                                     laws: objective, smoothing, penalty, checkpoint, restart and
                                     displacement composition; the non-accepted branch at a symmetric
                                     fixed point (exact penalty sequence, clamp at the maximum option,
                                     no move, no candidate)                           -> pinned
  adaptive trajectory                the reference ships no trajectory; the momentum coefficient
                                     min(0.92, age / (age + 3)) and the look-ahead restart test
                                     (:751-773) are covered only through the invariants of
                                     PlacementFlowTests.cpp:153-216                   -> parity unpinned
  window-reordering detailed pass    the reference's fixture (DetailedPlacementTests.cpp:55-70);
                                     bit-equal to the reference's own CPU placer (built from
                                     DetailedPlacer.cpp) when no two windows of a pass share a net;
                                     the same windows and orderings visited on a legalized
                                     placement                                        -> pinned
"""

import ctypes as C
import os

import numpy as np
import pytest
from scipy import fft as sf

from myplacement_amd import abi, design
from oracle import bindings as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
BOOK = os.path.join(GOLD, "bookshelf")


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLD, "neumann_golden.npz"))


@pytest.fixture(scope="module")
def ref():
    lib = ob.reference()
    if lib is None:
        pytest.skip("the reference checker oracle/_ref is not built here")
    return lib


# ---- transforms ------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [12, 16, 64, 100])
@pytest.mark.parametrize("method", [1, 0])
def test_transforms_match_scipy_golden(golden, n, method):
    ob.set_method(method)
    try:
        x = golden[f"x{n}"]
        for kind, key in ((0, "redft10"), (1, "redft01"), (2, "rodft01")):
            ref_v = golden[f"{key}_{n}"]
            got = ob.transform_1d(kind, x)
            assert np.abs(got - ref_v).max() <= 1e-13 * np.abs(ref_v).max()
    finally:
        ob.set_method(0)


def test_fft_and_direct_forms_agree():
    rng = np.random.default_rng(5)
    for n in (4, 8, 32, 256, 1024):
        x = rng.standard_normal(n)
        for kind in (0, 1, 2):
            ob.set_method(1)
            a = ob.transform_1d(kind, x)
            ob.set_method(2)
            b = ob.transform_1d(kind, x)
            ob.set_method(0)
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_transform_planted_error_is_caught(golden):
    # the golden check is sharp enough to see a single flipped sign in the sine transform
    x = golden["x16"]
    got = ob.transform_1d(2, x)
    wrong = got.copy()
    wrong[1::2] *= -1
    assert np.abs(wrong - golden["rodft01_16"]).max() > 1e-3


# ---- Neumann field (DensityFieldTests.cpp) -----------------------------------------------------

REGION = (-3.0, 2.0, 5.0, 8.0)


def _grid(cols, rows):
    bw, bh = (REGION[2] - REGION[0]) / cols, (REGION[3] - REGION[1]) / rows
    x = REGION[0] + (np.arange(cols) + 0.5) * bw
    y = REGION[1] + (np.arange(rows) + 0.5) * bh
    return np.meshgrid(x - REGION[0], y - REGION[1])


@pytest.mark.parametrize("mx,my", [(2, 3), (0, 3), (2, 0)])
def test_analytic_neumann_modes(mx, my):
    cols, rows = 16, 12
    X, Y = _grid(cols, rows)
    kx, ky = np.pi * mx / 8.0, np.pi * my / 6.0
    den = kx * kx + ky * ky
    src = np.cos(kx * X) * np.cos(ky * Y)
    pot, fx, fy = ob.field_solve(src, REGION)
    assert np.abs(pot - src / den).max() < 2e-12
    assert np.abs(fx - kx * np.sin(kx * X) * np.cos(ky * Y) / den).max() < 2e-12
    assert np.abs(fy - ky * np.cos(kx * X) * np.sin(ky * Y) / den).max() < 2e-12


def test_uniform_source_gives_zero():
    pot, fx, fy = ob.field_solve(np.full((12, 16), 7.0), REGION)
    for g in (pot, fx, fy):
        assert np.abs(g).max() < 2e-12


def test_zero_normal_field_on_edges():
    cols, rows = 16, 12
    X, Y = _grid(cols, rows)
    src = np.cos(np.pi * 2 / 8.0 * X) * np.cos(np.pi * 3 / 6.0 * Y)
    _, fx, fy = ob.field_solve(src, REGION)
    bw, bh = 8.0 / cols, 6.0 / rows
    ys = REGION[1] + (np.arange(rows) + 0.5) * bh
    xs = REGION[0] + (np.arange(cols) + 0.5) * bw
    left = ob.sample_field(fx, fy, REGION, np.stack([np.full(rows, REGION[0]), ys], 1))
    right = ob.sample_field(fx, fy, REGION, np.stack([np.full(rows, REGION[2]), ys], 1))
    bottom = ob.sample_field(fx, fy, REGION, np.stack([xs, np.full(cols, REGION[1])], 1))
    top = ob.sample_field(fx, fy, REGION, np.stack([xs, np.full(cols, REGION[3])], 1))
    assert np.abs(left[:, 0]).max() < 2e-12 and np.abs(right[:, 0]).max() < 2e-12
    assert np.abs(bottom[:, 1]).max() < 2e-12 and np.abs(top[:, 1]).max() < 2e-12
    # at bin centres the bilinear sample is the grid value itself
    Xc, Yc = np.meshgrid(xs, ys)
    s = ob.sample_field(fx, fy, REGION, np.stack([Xc.ravel(), Yc.ravel()], 1))
    assert np.abs(s[:, 0] - fx.ravel()).max() < 1e-13
    assert np.abs(s[:, 1] - fy.ravel()).max() < 1e-13


@pytest.mark.parametrize("key", ["field_12x16", "field_64x64", "field_48x80"])
def test_field_golden(golden, key):
    pot, fx, fy = ob.field_solve(golden[key + "_in"], golden[key + "_region"])
    for name, got in (("potential", pot), ("fx", fx), ("fy", fy)):
        ref_v = golden[f"{key}_{name}"]
        assert np.abs(got - ref_v).max() <= 1e-12 * np.abs(ref_v).max()


def test_oversized_grid_rejected():
    with pytest.raises(ValueError):
        ob.field_solve(np.zeros((1024, 1025)), REGION)
    with pytest.raises(ValueError):
        ob.density_map(REGION, 1025, 1024, 0.85, [], [], [])


# ---- density model ---------------------------------------------------------------------------

KAT_RECTS = [(0, 0, 2, 5), (0, 5, 2, 10), (2, 0, 5, 5), (2, 5, 5, 10)]
KAT_LAYERS = [1, 2, 0, 3]  # movable, macro, fixed, filler


def test_density_two_by_two_known_answer():
    w, wo, bins = ob.density_map((0, 0, 10, 10), 2, 2, 0.5, [(0, 0, 5, 10)], KAT_RECTS, KAT_LAYERS)
    assert wo["dark_area"] == pytest.approx(50.0, abs=1e-12)
    assert wo["placeable_area"] == pytest.approx(50.0, abs=1e-12)
    assert wo["normalization_area"] == pytest.approx(15.0, abs=1e-12)
    assert w["normalization_area"] == pytest.approx(15.0, abs=1e-12)
    assert wo["total_charge_area"] == pytest.approx(47.5, abs=1e-12)
    assert w["total_charge_area"] == pytest.approx(55.0, abs=1e-12)
    assert wo["total_overflow_area"] == pytest.approx(5.0, abs=1e-12)
    assert wo["normalized_overflow"] == pytest.approx(1.0 / 3.0, abs=1e-12)
    assert bins[1, 4] == pytest.approx(25.0, abs=1e-12)  # bin (1, 0) dark area


def test_overlapping_rows_rejected():
    with pytest.raises(ValueError):
        ob.density_map((0, 0, 10, 10), 2, 2, 0.5, [(0, 0, 10, 10), (0, 0, 10, 10)], [], [])


def test_density_map_matches_reference_code(ref):
    rng = np.random.default_rng(11)
    for trial in range(5):
        cols, rows = int(rng.integers(3, 40)), int(rng.integers(3, 40))
        region = (-5.0, 3.0, 95.0, 60.0)
        rects, layers = [], []
        for _ in range(300):
            cx, cy = rng.uniform(-10, 100), rng.uniform(0, 65)
            w, h = rng.uniform(0, 9), rng.uniform(0, 9)
            rects.append((cx - w / 2, cy - h / 2, cx + w / 2, cy + h / 2))
            layers.append(int(rng.integers(0, 5)))
        regions = [(-5.0, 3.0 + 5.7 * k, 95.0, 3.0 + 5.7 * (k + 1)) for k in range(10)]
        a = ob.density_map(region, cols, rows, 0.8, regions, rects, layers)
        b = ob.density_map(region, cols, rows, 0.8, regions, rects, layers, lib=ref)
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2], b[2])


def _fixture(name):
    return design.read_bookshelf(os.path.join(BOOK, f"{name}.aux"))


def test_mixed_normalization_area_540():
    d = _fixture("mixed")
    _, m = ob.design_metrics(d, d.center, 8, 8, 0.85)
    assert m["normalization_area"] == pytest.approx(540.0, abs=1e-9)


def test_tiny_fixture_counts():
    d = _fixture("tiny")
    assert d.module_count == 4 and d.net_count == 2 and d.pin_count == 5
    assert int((d.is_fixed | d.is_terminal).sum()) == 1 and d.movable_ids().shape[0] == 3
    core = d.core_region()
    assert core[2] - core[0] == pytest.approx(100.0) and core[3] - core[1] == pytest.approx(40.0)
    hpwl, _ = ob.design_metrics(d, d.center, 8, 8, 0.85)
    assert hpwl > 0


def _reference_parse(ref, aux):
    """The database the reference's own BookshelfParser::parseAux builds from `aux`, as arrays."""
    h = C.c_void_p()
    assert ref.ref_parse_aux(aux.encode(), C.byref(h)) == 0, ref.ref_last_error()
    sizes = np.zeros(7, dtype=np.int64)
    ref.ref_db_sizes(h, sizes.ctypes.data_as(C.POINTER(C.c_int64)))
    M, N, I, R = (int(v) for v in sizes[:4])
    out = dict(width=np.zeros(M), height=np.zeros(M), center=np.zeros((M, 2)), orientation=np.zeros(M, np.int32),
               is_terminal=np.zeros(M, np.uint8), is_fixed=np.zeros(M, np.uint8), net_weight=np.zeros(N),
               net_pin_start=np.zeros(N + 1, np.int64), pin_module=np.zeros(I, np.int32),
               pin_offset=np.zeros((I, 2)), rows=np.zeros((R, 5)))
    core = np.zeros(4)
    args = [out[k].ctypes.data_as(C.c_void_p) for k in out] + [core.ctypes.data_as(C.c_void_p)]
    assert ref.ref_db_export(h, *args) == 0
    ref.ref_db_destroy(h)
    return out, tuple(core)


def test_bookshelf_reader_matches_reference_parser(ref):
    for name in ("tiny", "mixed"):
        d = _fixture(name)
        out, core = _reference_parse(ref, os.path.join(BOOK, f"{name}.aux"))
        for k, v in out.items():
            assert np.array_equal(getattr(d, k), v), (name, k)
        assert core == d.core_region()


NETLIST_FIELDS = ("width", "height", "orientation", "is_terminal", "net_weight", "net_pin_start", "pin_module",
                  "pin_offset", "rows")


@pytest.mark.parametrize("which", ["tiny", "mixed", "plain-random", "served-12k", "large-30k-quadratic-like"])
def test_bookshelf_writer_round_trips_through_reference_parser(ref, tmp_path, which):
    """design.write_bookshelf -> the reference's BookshelfParser (oracle/_ref) and read_bookshelf:
    both read back the same arrays bit for bit, and the netlist, rows, weights and centres are the
    written design's own (centres are written as lower-left corners that give them back exactly;
    terminals come back fixed, as refreshDerivedData makes them)."""
    if which in ("tiny", "mixed"):
        d = _fixture(which)
    elif which == "plain-random":
        d = _synthetic_with_random_start()
    elif which == "served-12k":
        d = design.served_design(77, 12000)
        d.center = ob.random_placement(d, 77)
        d.net_weight = np.random.default_rng(3).uniform(0.5, 2.0, d.net_count)
        d.orientation[: d.module_count // 3] = np.arange(d.module_count // 3) % 8
    else:
        d = design.large_design(seed=9, cells=30000, nets=31000, row_count=300, sites=600, macros=6)
        rng = np.random.default_rng(9)
        mov = d.movable_ids()
        core = d.core_region()
        d.center[mov, 0] = rng.normal((core[0] + core[2]) / 2, (core[2] - core[0]) / 10, mov.shape[0])
        d.center[mov, 1] = rng.normal((core[1] + core[3]) / 2, (core[3] - core[1]) / 10, mov.shape[0])
    aux = design.write_bookshelf(d, str(tmp_path), "rt")
    out, core = _reference_parse(ref, aux)
    mine = design.read_bookshelf(aux)
    for k, v in out.items():
        assert np.array_equal(getattr(mine, k), v), (which, k)
    assert core == mine.core_region() == d.core_region()
    for k in NETLIST_FIELDS:
        assert np.array_equal(getattr(d, k), out[k]), (which, k)
    assert np.array_equal(out["is_fixed"], (d.is_fixed | d.is_terminal).astype(np.uint8))
    assert np.array_equal(out["center"], d.center), (which, np.abs(out["center"] - d.center).max())


def _synthetic_with_random_start(seed=2026):
    d = design.plain_design(seed=seed)
    d.center = ob.random_placement(d, seed)
    return d


def test_fillers_and_static_grids_bit_equal_reference(ref):
    for d, t, bins in ((_synthetic_with_random_start(), 0.9, 64), (_fixture("mixed"), 0.85, 8),
                       (design.large_design(seed=4, cells=20000, nets=21000, row_count=200, sites=400, macros=6), 0.8, 128)):
        if d.module_count > 10000:
            d.center = ob.random_placement(d, 9)
        o = abi.default_options(target_density=t, bins_x=bins, bins_y=bins, maximum_fillers=50000)
        s = ob.OracleSession(d, o)
        L = s.layout()
        A = s.arrays()
        cnt = C.c_int64()
        cap = 60000
        F = np.zeros((cap, 4))
        fixed, dark = np.zeros(bins * bins), np.zeros(bins * bins)
        dc = d.to_c()
        assert ref.ref_fillers(C.byref(dc), bins, bins, t, 50000, 2026, fixed.ctypes.data_as(C.c_void_p),
                               dark.ctypes.data_as(C.c_void_p), C.byref(cnt), F.ctypes.data_as(C.c_void_p), cap) == 0
        n = cnt.value
        assert n == s.info.filler_count
        assert np.array_equal(F[:n, :2], A["filler_centers"])
        assert np.all(F[:n, 2] == s.info.filler_width) and np.all(F[:n, 3] == s.info.filler_height)
        assert np.array_equal(fixed, L["fixed_area"]) and np.array_equal(dark, L["dark_area"])


def test_design_metrics_match_reference(ref):
    d = _synthetic_with_random_start()
    hp, m = ob.design_metrics(d, d.center, 64, 64, 0.9)
    hp_r = C.c_double()
    m_r = abi.MpDensityMetrics()
    dc = d.to_c()
    assert ref.ref_design_metrics(C.byref(dc), 64, 64, 0.9, C.byref(hp_r), C.byref(m_r)) == 0
    assert hp == hp_r.value
    assert m == m_r.as_dict()


# ---- smooth wirelength -----------------------------------------------------------------------

def _two_module_design(ax, ay, bx, by):
    return design.Design(
        width=[2.0, 2.0], height=[10.0, 10.0], center=[[ax, ay], [bx, by]], orientation=[0, 0],
        is_terminal=[0, 0], is_fixed=[0, 0], net_weight=[1.5], net_pin_start=[0, 2], pin_module=[0, 1],
        pin_offset=[[0.0, 0.0], [0.0, 0.0]], rows=[[10.0 * r, 10.0, 0.0, 1.0, 100.0] for r in range(10)])


def test_two_pin_closed_form_and_limit():
    a, b = (20.0, 33.0), (61.5, 48.25)
    d = _two_module_design(*a, *b)
    o = abi.default_options(bins_x=8, bins_y=8, maximum_fillers=0)
    s = ob.OracleSession(d, o)
    xy = np.array([a, b])
    for gamma in (0.5, 3.0, 20.0):
        r = s.evaluate(xy, gamma, True)
        def lse(u, v):
            return gamma * np.logaddexp(u / gamma, v / gamma) + gamma * np.logaddexp(-u / gamma, -v / gamma)
        closed = 1.5 * (lse(a[0], b[0]) + lse(a[1], b[1]))
        assert r["smooth_wirelength"] == pytest.approx(closed, rel=1e-13)
        assert r["hpwl"] == pytest.approx(1.5 * ((b[0] - a[0]) + (b[1] - a[1])), rel=1e-15)
        # gradient of pin 0 along x: w * (softmax weight of a in max - softmax weight of a in min)
        ex = 1.5 * (1 / (1 + np.exp((b[0] - a[0]) / gamma)) - 1 / (1 + np.exp((a[0] - b[0]) / gamma)))
        assert r["wire_gradient"][0, 0] == pytest.approx(ex, rel=1e-12)
    r = s.evaluate(xy, 1e-3, False)
    assert r["smooth_wirelength"] == pytest.approx(r["hpwl"], rel=1e-12)


def test_wire_gradient_finite_differences():
    d = _synthetic_with_random_start()
    o = abi.default_options(target_density=0.9)
    s = ob.OracleSession(d, o)
    xy = s.arrays()["positions"]
    gamma = 4.0
    g = s.evaluate(xy, gamma, True)["wire_gradient"]
    rng = np.random.default_rng(3)
    h = 1e-4
    for p in rng.choice(int(s.info.movable_count), 12, replace=False):
        for axis in (0, 1):
            up, dn = xy.copy(), xy.copy()
            up[p, axis] += h
            dn[p, axis] -= h
            fd = (s.evaluate(up, gamma, False)["smooth_wirelength"] - s.evaluate(dn, gamma, False)["smooth_wirelength"]) / (2 * h)
            assert fd == pytest.approx(g[p, axis], abs=1e-6 * max(1.0, abs(g[p, axis])))


# ---- energy ----------------------------------------------------------------------------------

def test_energy_equals_dct_basis_sum():
    d = _synthetic_with_random_start()
    o = abi.default_options(target_density=0.9)
    s = ob.OracleSession(d, o)
    xy = s.arrays()["positions"]
    r = s.evaluate(xy, 3.0, False)
    src = s.grids()["source"].reshape(64, 64)
    core = d.core_region()
    W, H = core[2] - core[0], core[3] - core[1]
    chat = sf.dct(sf.dct(src, type=2, axis=1), type=2, axis=0)
    wx = np.pi * np.arange(64) / W
    wy = np.pi * np.arange(64) / H
    den = wy[:, None] ** 2 + wx[None, :] ** 2
    den[0, 0] = np.inf
    phihat = chat / (4 * 64 * 64) / den
    c = np.where(np.arange(64) == 0, 1.0, 2.0)
    weight = c[:, None] * c[None, :] / 4
    energy = 0.5 * (W / 64) * (H / 64) * np.sum(weight * chat * phihat)
    assert r["electrostatic_energy"] == pytest.approx(energy, rel=1e-10)


# ---- random placement / controller -----------------------------------------------------------

def test_random_placement_inside_core_and_deterministic():
    d = design.plain_design()
    a = ob.random_placement(d, 2026)
    b = ob.random_placement(d, 2026)
    assert np.array_equal(a, b)
    core = d.core_region()
    mov = d.movable_ids()
    assert np.all(a[mov, 0] - d.width[mov] / 2 >= core[0]) and np.all(a[mov, 0] + d.width[mov] / 2 <= core[2])
    fixed = np.setdiff1d(np.arange(d.module_count), mov)
    assert np.array_equal(a[fixed], d.center[fixed])


def test_adaptive_run_invariants_tiny():
    d = _fixture("tiny")
    o = abi.default_options(iterations=24, bins_x=8, bins_y=8, maximum_fillers=100)
    c, r, hist = ob.global_place(d, o)
    assert r["completed_iterations"] == 24 and len(hist) == 24
    assert r["accepted_iterations"] > 0 and r["restored_best_checkpoint"] == 1
    assert r["best_checkpoint_iteration"] <= r["completed_iterations"]
    for row in hist:
        assert all(np.isfinite(v) for v in row.values())
        assert row["step_size"] >= 0 and row["maximum_displacement"] >= 0
    core = d.core_region()
    mov = d.movable_ids()
    assert np.all(c[mov, 0] - d.width[mov] / 2 >= core[0] - 1e-9)


def test_aggressive_controller_backtracks_and_restarts():
    d = _fixture("tiny")
    d.center = ob.random_placement(d, 2026)
    o = abi.default_options(iterations=24, bins_x=8, bins_y=8, maximum_fillers=100, initial_movement_in_bins=1.0,
                            maximum_movement_in_bins=1.0, armijo_coefficient=0.5, objective_increase_for_density=0.0)
    _, r, _ = ob.global_place(d, o)
    assert r["rejected_candidates"] > 0
    assert r["momentum_restarts"] > 0


def test_invalid_controller_option_rejected():
    d = _fixture("tiny")
    o = abi.default_options(iterations=1, bins_x=8, bins_y=8, feasible_refinement_iterations=0)
    with pytest.raises(ValueError):
        ob.global_place(d, o)


# ---- quadratic initial placement (InitialPlacer.cpp:266-402) ---------------------------------------

def _tiny_design(cells_xy, terminals_xy, nets, W=100.0, H=40.0, offsets=None):
    """Unit cells and 1x1 terminals on one 40-unit-high row region, nets as lists of module ids."""
    m = len(cells_xy) + len(terminals_xy)
    pins = [p for n in nets for p in n]
    start = np.concatenate([[0], np.cumsum([len(n) for n in nets])])
    off = np.zeros((len(pins), 2)) if offsets is None else np.asarray(offsets, dtype=np.float64)
    return design.Design(
        width=np.ones(m), height=np.ones(m), center=np.array(list(cells_xy) + list(terminals_xy), dtype=np.float64),
        orientation=np.zeros(m, dtype=np.int32),
        is_terminal=np.array([0] * len(cells_xy) + [1] * len(terminals_xy), dtype=np.uint8),
        is_fixed=np.array([0] * len(cells_xy) + [1] * len(terminals_xy), dtype=np.uint8),
        net_weight=np.ones(len(nets)), net_pin_start=start, pin_module=np.array(pins, dtype=np.int32),
        pin_offset=off, rows=np.array([[0.0, H, 0.0, 1.0, W]]))


def test_quadratic_one_cell_closed_form():
    """One cell tied to one terminal: each outer iteration solves (w + a) x = w X + a x_prev with the
    bound-to-bound weight w = 2 / max(|x_prev - X|, 1) (clique weight 2w/(d-1) = 2) and anchor a = 1e-4."""
    d = _tiny_design([(10.0, 10.0)], [(80.0, 30.0)], [[0, 1]])
    o = abi.default_qp_options(outer_iterations=1, tolerance=1e-14)
    c, r = ob.quadratic_placement(d, o)
    x0, y0 = 50.0, 20.0  # core centre
    wx, wy = 2.0 / max(abs(x0 - 80.0), 1.0), 2.0 / max(abs(y0 - 30.0), 1.0)
    ex = (wx * 80.0 + 1e-4 * x0) / (wx + 1e-4)
    ey = (wy * 30.0 + 1e-4 * y0) / (wy + 1e-4)
    assert abs(c[0, 0] - ex) < 1e-10 and abs(c[0, 1] - ey) < 1e-10
    assert r["iterations"] == 1 and np.all(c[1] == [80.0, 30.0])


def _dense_system(d, cx, cy, limit=96, min_dist=1.0):
    """The clique / star bound-to-bound system written densely from the net list."""
    mov = d.movable_ids()
    idx = {int(m): k for k, m in enumerate(mov)}
    n = len(mov)
    Ax, Ay, bx, by = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    pos = lambda p: (cx[d.pin_module[p]] + d.pin_offset[p, 0], cy[d.pin_module[p]] + d.pin_offset[p, 1])
    for k in range(d.net_count):
        pins = list(range(d.net_pin_start[k], d.net_pin_start[k + 1]))
        deg = len(pins)
        if deg < 2:
            continue
        if deg <= limit:
            pairs, base = [(a, b) for i, a in enumerate(pins) for b in pins[i + 1:]], 2.0 * d.net_weight[k] / (deg - 1)
        else:
            pairs, base = [(pins[0], b) for b in pins[1:]], d.net_weight[k] / (deg - 1)
        for a, b in pairs:
            ma, mb = d.pin_module[a], d.pin_module[b]
            if ma == mb or (ma not in idx and mb not in idx):
                continue
            (ax_, ay_), (bx_, by_) = pos(a), pos(b)
            for A, rhs, pa, pb, oa, ob_, ca, cb in ((Ax, bx, ax_, bx_, d.pin_offset[a, 0], d.pin_offset[b, 0], cx, cx),
                                                    (Ay, by, ay_, by_, d.pin_offset[a, 1], d.pin_offset[b, 1], cy, cy)):
                w = base / max(abs(pa - pb), min_dist)
                # energy w (xa + oa - xb - ob)^2 with fixed terms moved to the right-hand side
                if ma in idx and mb in idx:
                    i, j = idx[ma], idx[mb]
                    A[i, i] += w; A[j, j] += w; A[i, j] -= w; A[j, i] -= w
                    rhs[i] -= w * (oa - ob_); rhs[j] += w * (oa - ob_)
                elif ma in idx:
                    i = idx[ma]
                    A[i, i] += w
                    rhs[i] += w * (cb[mb] + ob_ - oa)
                else:
                    j = idx[mb]
                    A[j, j] += w
                    rhs[j] += w * (ca[ma] + oa - ob_)
    for k, m in enumerate(mov):
        Ax[k, k] += 1e-4; Ay[k, k] += 1e-4
        bx[k] += 1e-4 * cx[m]; by[k] += 1e-4 * cy[m]
    return Ax, Ay, bx, by


def test_quadratic_one_iteration_matches_dense_solve():
    d = design.plain_design(cells=120, nets=150, row_count=10, sites=100, terminals=16)
    o = abi.default_qp_options(outer_iterations=1, solver_iterations=2000, tolerance=1e-13)
    c, r = ob.quadratic_placement(d, o)
    mov = d.movable_ids()
    core = d.core_region()
    cx, cy = d.center[:, 0].copy(), d.center[:, 1].copy()
    cx[mov], cy[mov] = (core[0] + core[2]) / 2, (core[1] + core[3]) / 2
    Ax, Ay, bx, by = _dense_system(d, cx, cy)
    sx, sy = np.linalg.solve(Ax, bx), np.linalg.solve(Ay, by)
    w = d.width[mov] / 2
    h = d.height[mov] / 2
    ex = np.clip(sx, core[0] + w, core[2] - w)
    ey = np.clip(sy, core[1] + h, core[3] - h)
    assert r["relative_residual"][0] < 1e-12 and r["relative_residual"][1] < 1e-12
    assert np.abs(c[mov, 0] - ex).max() < 1e-6 and np.abs(c[mov, 1] - ey).max() < 1e-6


def test_quadratic_invariants_like_the_reference():
    """PlacementFlowTests.cpp:121-129: three outer iterations on tiny.aux keep every movable module in
    the core with a finite HPWL; the plain design's HPWL is below that of the centre start."""
    d = design.read_bookshelf(os.path.join(BOOK, "tiny.aux"))
    c, r = ob.quadratic_placement(d, abi.default_qp_options(outer_iterations=3))
    core = d.core_region()
    mov = d.movable_ids()
    assert r["iterations"] == 3 and np.isfinite(r["hpwl_after"])
    assert np.all(c[mov, 0] - d.width[mov] / 2 >= core[0] - 1e-9)
    assert np.all(c[mov, 0] + d.width[mov] / 2 <= core[2] + 1e-9)
    assert np.all(c[mov, 1] - d.height[mov] / 2 >= core[1] - 1e-9)
    assert np.all(c[mov, 1] + d.height[mov] / 2 <= core[3] + 1e-9)
    p = design.plain_design()
    cp, rp = ob.quadratic_placement(p)
    centre = p.center.copy()
    pm = p.movable_ids()
    pc = p.core_region()
    centre[pm] = [(pc[0] + pc[2]) / 2, (pc[1] + pc[3]) / 2]
    h_centre, _ = ob.design_metrics(p, centre, 8, 8, 0.9)
    assert rp["hpwl_after"] < h_centre
    with pytest.raises(ValueError):
        ob.quadratic_placement(p, abi.default_qp_options(minimum_distance=0.0))


# ---- detailed placement: the device window-reordering pass ----------------------------------

def test_window_reordering_fixture_like_the_reference():
    """tests/DetailedPlacementTests.cpp:55-70: one pass, one window, 24 orderings, one accepted
    reorder, HPWL 20 -> 10; the reference's own CPU placer gives the same."""
    from dp_designs import reordering_fixture
    d = reordering_fixture()
    _, _, r = ob.detailed_place(d, abi.default_dp_options(passes=1))
    assert r["accepted_operations"] == 1 and r["evaluated_permutations"] == 24 and r["evaluated_windows"] == 1
    assert r["hpwl_before"] == 20.0 and abs(r["hpwl_after"] - 10.0) < 1e-12
    if ob.reference() is not None:
        _, _, rr = ob.ref_detailed_place(d, passes=1)
        assert rr["hpwl_after"] == r["hpwl_after"] and rr["accepted_operations"] == 1


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_window_reordering_matches_reference_placer_when_windows_share_no_net(ref, seed):
    """With every net inside one phase-0 window plus a fixed terminal, scoring the windows at the
    start of the pass (the device pass) and one after another (the reference's CPU placer,
    DetailedPlacer.cpp:289-335, built from its own source) must give the same placement bit for
    bit: segments, windows, orderings, the re-scoring and the apply rule all coincide.  (The
    terminal anchors each net: without it, orderings that only translate a net score equal up to
    rounding, and the device pass, which takes the lowest score, and the CPU placer, which keeps the
    first ordering better by more than epsilon, may pick different ones of equal HPWL.)"""
    from dp_designs import abutted_design
    d = abutted_design(seed=seed, private=True, terminal_share=1.0)
    c, o, r = ob.detailed_place(d, abi.default_dp_options(passes=1))
    rc, ro, rr = ob.ref_detailed_place(d, passes=1)
    assert r["accepted_operations"] > 0
    assert np.array_equal(c, rc) and np.array_equal(o, ro)
    for k in ("evaluated_windows", "evaluated_permutations", "accepted_operations"):
        assert r[k] == rr[k], k
    assert r["hpwl_after"] == rr["hpwl_after"]


def test_window_reordering_on_a_legalized_placement(ref):
    """Oracle GP -> the reference's Abacus legalizer -> the device pass: still legal, HPWL not
    worse, the same windows and orderings visited as the reference's CPU placer (whose sequential
    scoring may accept a few different reorders), and a final HPWL within 0.1 % of it."""
    d = design.plain_design(seed=2026)
    d.center = ob.random_placement(d, 2026)
    c, _, _ = ob.global_place(d, abi.default_options(target_density=0.9, bins_x=64, bins_y=64, iterations=120))
    lc, lo, rep = ob.ref_legalize(d, c)
    assert rep["legal"] == 1.0
    oc, oo, r = ob.detailed_place(d, abi.default_dp_options(), lc, lo)
    _, _, rr = ob.ref_detailed_place(d, lc, lo)
    assert ob.ref_legality(d, oc, oo)["legal"] == 1.0
    assert r["hpwl_after"] <= r["hpwl_before"] and r["accepted_operations"] > 0
    assert r["evaluated_windows"] == rr["evaluated_windows"]
    assert r["evaluated_permutations"] == rr["evaluated_permutations"]
    assert abs(r["hpwl_after"] / rr["hpwl_after"] - 1.0) < 1e-3
    h, _ = ob.design_metrics(d, oc, 8, 8, 0.9)
    assert abs(h - r["hpwl_after"]) < 1e-9 * h


def test_window_reordering_invalid_options_rejected():
    from dp_designs import reordering_fixture
    d = reordering_fixture()
    for bad in (dict(passes=0), dict(maximum_net_degree=1), dict(improvement_epsilon=-1.0),
                dict(improvement_epsilon=float("nan")), dict(method=7)):
        with pytest.raises(ValueError):
            ob.detailed_place(d, abi.default_dp_options(**bad))


# ---- controller laws on objectives whose answers are closed forms -------------------------------

def _step(pos, pre, prev_pos, prev_grad, has_prev=True, fallback=0.3):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (pos, pre, prev_pos, prev_grad)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    return ob.oracle().orc_estimate_step(*ptrs, len(arrs[0]) // 2, 1 if has_prev else 0, fallback)


def test_estimate_step_is_the_inverse_curvature_of_a_quadratic():
    """estimateStep (AdaptiveGlobalPlacer.cpp:470-492) on f = L/2 |x|^2, whose preconditioned gradient
    is L x: the secant curvature is exactly L, so the step is 1/L while that lies within [fallback/4,
    4 fallback], and the nearer bound otherwise.  A dropped square root (1/L^2) or an inverted ratio
    (L) fails."""
    rng = np.random.default_rng(5)
    x, xp = rng.normal(size=(50, 2)), rng.normal(size=(50, 2))
    for L in (2.5, 1.0, 3.7):
        assert _step(x, L * x, xp, L * xp) == pytest.approx(1.0 / L, rel=1e-12)
    assert _step(x, 100.0 * x, xp, 100.0 * xp, fallback=0.3) == pytest.approx(0.075, rel=1e-15)  # clamped below
    assert _step(x, 0.01 * x, xp, 0.01 * xp, fallback=0.3) == pytest.approx(1.2, rel=1e-15)      # clamped above
    # no previous state, or no movement since it: the fallback
    assert _step(x, 2.0 * x, xp, 2.0 * xp, has_prev=False) == 0.3
    assert _step(x, 2.0 * x, x, 3.0 * x) == 0.3


def test_estimate_step_lies_between_the_inverse_extreme_curvatures():
    """On f = 1/2 x^T H x with H = diag(a, b) per particle the secant step is sqrt(|dx|^2 / |H dx|^2):
    exactly 1/a (1/b) for motion along x (y) alone, and between 1/b and 1/a for any other motion
    (a Rayleigh-quotient bound)."""
    a, b = 1.0, 4.0
    H = np.array([a, b])
    rng = np.random.default_rng(6)
    xp = rng.normal(size=(40, 2))
    along_x = xp + np.column_stack([rng.normal(size=40), np.zeros(40)])
    along_y = xp + np.column_stack([np.zeros(40), rng.normal(size=40)])
    assert _step(along_x, along_x * H, xp, xp * H) == pytest.approx(1.0 / a, rel=1e-12)
    assert _step(along_y, along_y * H, xp, xp * H) == pytest.approx(1.0 / b, rel=1e-12)
    for _ in range(5):
        x = xp + rng.normal(size=(40, 2))
        s = _step(x, x * H, xp, xp * H)
        assert 1.0 / b < s < 1.0 / a


def test_smoothing_anneals_between_its_end_points():
    """smoothingFromOverflow (:456-462): the initial smoothing (in bins) at or above the initial
    overflow, the final one at zero overflow, and the overflow ratio raised to the exponent between."""
    lib = ob.oracle()
    o = abi.default_options(initial_smoothing=0.75, final_smoothing=0.12, smoothing_overflow_exponent=1.0)
    mb = 3.0
    assert lib.orc_smoothing(C.byref(o), mb, 0.4, 0.4) == pytest.approx(0.75 * mb, rel=1e-15)
    assert lib.orc_smoothing(C.byref(o), mb, 0.9, 0.4) == pytest.approx(0.75 * mb, rel=1e-15)
    assert lib.orc_smoothing(C.byref(o), mb, 0.0, 0.4) == pytest.approx(0.12 * mb, rel=1e-15)
    assert lib.orc_smoothing(C.byref(o), mb, 0.2, 0.4) == pytest.approx((0.12 + 0.75) / 2 * mb, rel=1e-14)
    o.smoothing_overflow_exponent = 2.0
    assert lib.orc_smoothing(C.byref(o), mb, 0.2, 0.4) == pytest.approx((0.12 + 0.63 / 4) * mb, rel=1e-14)
    values = [lib.orc_smoothing(C.byref(o), mb, v, 0.4) for v in np.linspace(0.0, 0.4, 9)]
    assert all(u < w for u, w in zip(values, values[1:]))


def test_armijo_acceptance_on_a_quadratic_line_search():
    """acceptsCandidate (:521-534) on f(x) = L/2 x^2 from x = 1 with the gradient step y = 1 - tL:
    the predicted decrease is g (x - y) = t L^2, and Armijo's condition f(y) <= f(x) - c t L^2 holds
    exactly when tL <= 2 (1 - c).  With the overflow worse, the stability and density-trade rules stay
    out; with the overflow unchanged, any candidate no worse than the base is accepted (stability);
    with the overflow lower by more than the stall gain, a bounded objective increase is accepted."""
    lib = ob.oracle()
    c, L = 0.25, 1.0
    o = abi.default_options(armijo_coefficient=c, density_stall_threshold=0.002, objective_increase_for_density=0.1)
    base = 0.5 * L

    def accepts(tL, cand_overflow):
        t = tL / L
        return lib.orc_accepts(C.byref(o), base, 0.3, 0.5 * L * (1.0 - tL) ** 2, cand_overflow, t * L * L)

    limit = 2.0 * (1.0 - c)
    for tL in (0.2, 1.0, limit - 0.01):
        assert accepts(tL, 0.4) == 1
    for tL in (limit + 0.01, 1.9):
        assert accepts(tL, 0.4) == 0
    assert accepts(1.9, 0.3) == 1    # f(y) = 0.405 <= 0.5 and the overflow did not grow
    assert accepts(2.04, 0.3) == 0   # f(y) = 0.5408 > 0.5
    assert accepts(2.04, 0.29) == 1  # overflow down by 0.01 > the stall gain, objective within 10 %
    assert accepts(2.1, 0.29) == 0   # f(y) = 0.605: beyond the 10 % allowance (0.55)


# ---- the rest of the controller: penalty law, checkpoint rules, and their composition ----------

def test_density_penalty_law_known_answers():
    """updateDensityPenalty (AdaptiveGlobalPlacer.cpp:536-557), each branch and both clamps, with
    the default options (bound 0.1, critical above 0.2, stall 0.002, HPWL growth 2 %, x1.05 / x0.88):
    a swapped branch order, a wrong comparison or the wrong clamp bound fails one of these."""
    lib = ob.oracle()
    o = abi.default_options()

    def upd(pen, ref, b_hpwl, b_ovf, a_hpwl, a_ovf, a_design, opts=o):
        return lib.orc_update_penalty(C.byref(opts), pen, ref, b_hpwl, b_ovf, a_hpwl, a_ovf, a_design)

    # critical (0.5 > 0.2) and stalled (improvement 0.001 < 0.002): increase
    assert upd(10.0, 10.0, 100.0, 0.501, 100.0, 0.5, 0.5) == 10.0 * 1.05
    # critical but improving, design overflow high: unchanged
    assert upd(10.0, 10.0, 100.0, 0.6, 100.0, 0.5, 0.5) == 10.0
    # not critical (0.15), stalled, HPWL grew 3 % > 2 %: decrease
    assert upd(10.0, 10.0, 100.0, 0.151, 103.0, 0.15, 0.15) == 10.0 * 0.88
    # not critical, stalled, HPWL grew 1 %, overflow above the 0.1 bound: increase
    assert upd(10.0, 10.0, 100.0, 0.151, 101.0, 0.15, 0.15) == 10.0 * 1.05
    # exactly at the critical threshold (0.2 is not above it): the HPWL-growth decrease applies
    assert upd(10.0, 10.0, 100.0, 0.2005, 103.0, 0.2, 0.2) == 10.0 * 0.88
    # improving into the feasible range (design overflow <= 0.1): decrease
    assert upd(10.0, 10.0, 100.0, 0.2, 100.0, 0.05, 0.05) == 10.0 * 0.88
    # improving, overflow 0.15 > 0.1: unchanged
    assert upd(10.0, 10.0, 100.0, 0.3, 100.0, 0.15, 0.15) == 10.0
    # HPWL growth is measured against max(HPWL, 1): 0.5 -> 0.515 is 1.5 % of 1 (not 3 % of 0.5), so
    # no decrease; the overflow above its bound then increases the penalty
    assert upd(10.0, 10.0, 0.5, 0.151, 0.515, 0.15, 0.15) == 10.0 * 1.05
    # lower clamp max(minimum, 0.02 reference): 0.21 x 0.88 = 0.1848 -> 0.2
    assert upd(0.21, 10.0, 100.0, 0.2, 100.0, 0.05, 0.05) == 0.2
    assert upd(1e-4, 1e-3, 100.0, 0.2, 100.0, 0.05, 0.05) == 1e-4  # max(1e-4, 2e-5)
    # upper clamp max(maximum, 10 x lower)
    tight = abi.default_options(maximum_density_penalty=11.0)
    assert upd(10.5, 10.0, 100.0, 0.501, 100.0, 0.5, 0.5, tight) == 11.0
    wide = abi.default_options(minimum_density_penalty=5.0, maximum_density_penalty=6.0)
    assert upd(40.0, 10.0, 100.0, 0.501, 100.0, 0.5, 0.5, wide) == 42.0  # 10 x lower = 50 > 6
    assert upd(49.0, 10.0, 100.0, 0.501, 100.0, 0.5, 0.5, wide) == 50.0


def _checkpoints(hpwl, ovf, **kw):
    o = abi.default_options(**kw)
    hp = np.ascontiguousarray(hpwl, dtype=np.float64)
    ov = np.ascontiguousarray(ovf, dtype=np.float64)
    flags = np.zeros(len(hp), dtype=np.int32)
    feas = C.c_int32()
    chosen = ob.oracle().orc_checkpoint_sequence(C.byref(o), len(hp), hp.ctypes.data_as(C.POINTER(C.c_double)),
                                                 ov.ctypes.data_as(C.POINTER(C.c_double)),
                                                 flags.ctypes.data_as(C.c_void_p), C.byref(feas))
    return chosen, feas.value, flags.tolist()


def test_checkpoint_selection_rules():
    """considerCheckpoint (:563-596) and the restore choice (:930-935) on constructed sequences:
    feasible (design overflow <= 0.1, the bound included) beats any infeasible record whatever its
    HPWL; among feasible records the lowest HPWL, a record within 1e-10 counting only if its HPWL is
    lower; before any feasible record the lowest overflow, within 1e-12 the lower HPWL; once a
    feasible record exists the fallback no longer moves."""
    # never feasible: overflow 0.4 at iteration 1, then a 1e-13-higher overflow with lower HPWL
    # ties and wins; a lower overflow by less than 1e-12 with higher HPWL does not
    assert _checkpoints([100, 120, 110, 90, 130], [0.5, 0.4, 0.4 + 1e-13, 0.45, 0.4]) == (2, 0, [1, 1, 1, 0, 0])
    # feasible from iteration 1; the infeasible 70 never wins; 0.1 itself is feasible; a 5e-11 lower
    # HPWL is a tie that still replaces (lower HPWL), 5e-10 lower replaces outright
    seq_h = [80, 120, 70, 125, 119, 119 - 5e-11, 119 - 5e-11 - 5e-10, 200]
    seq_o = [0.5, 0.09, 0.3, 0.08, 0.1, 0.05, 0.07, 0.0]
    assert _checkpoints(seq_h, seq_o) == (6, 1, [1, 1, 0, 0, 1, 1, 1, 0])
    # an equal HPWL does not replace the earlier record
    assert _checkpoints([100, 100, 100], [0.05, 0.05, 0.01]) == (0, 1, [1, 0, 0])
    # a tighter bound makes all but the last record infeasible: the fallback follows the overflow down
    # (0.5, 0.09, 0.08, 0.05) until the feasible record takes over
    assert _checkpoints(seq_h, seq_o, maximum_bin_overflow=0.04) == (7, 1, [1, 1, 0, 1, 0, 1, 0, 1])
    assert _checkpoints([], []) == (-1, 0, [])


def _two_terminal_fixed_point():
    """One movable 20 x 10 standard cell at the centre of a 320 x 320 core, midway between two fixed
    1 x 1 terminals at (156, 160) and (164, 160), one two-pin net to each; no fillers."""
    rows = np.array([(r * 10.0, 10.0, 0.0, 1.0, 320.0) for r in range(32)])
    term = np.array([0, 1, 1], np.uint8)
    return design.Design(width=np.array([20.0, 1.0, 1.0]), height=np.array([10.0, 1.0, 1.0]),
                         center=np.array([[160.0, 160.0], [156.0, 160.0], [164.0, 160.0]]),
                         orientation=np.zeros(3, np.int32), is_terminal=term, is_fixed=term.copy(),
                         net_weight=np.ones(2), net_pin_start=np.array([0, 2, 4]),
                         pin_module=np.array([0, 1, 0, 2], np.int32), pin_offset=np.zeros((4, 2)), rows=rows)


def test_non_accepted_iterations_at_a_symmetric_fixed_point():
    """At a point where the gradient vanishes by symmetry (the cell midway between its two
    terminals at the centre of the core: every wirelength term cancels exactly and the field is
    zero there) every proposal predicts no reduction, so no iteration is accepted (:818-822, :909):
    the cell never moves, no candidate is evaluated, momentum never starts, and since the overflow
    (0.602 at t = 0.2) stays above its bound the penalty grows by exactly x1.05 per row, clamped at
    maximum_density_penalty itself -- not at max(maximum, 10 x lower), the accepted branch's bound
    (the options make the two differ: lower 0.5, 10 x lower 5, maximum 1.2)."""
    d = _two_terminal_fixed_point()
    o = abi.default_options(target_density=0.2, bins_x=32, bins_y=32, maximum_fillers=0, iterations=12,
                            minimum_density_penalty=0.5, maximum_density_penalty=1.2)
    o.feasible_refinement_iterations = 2 ** 31 - 1
    c, r, hist = ob.global_place(d, o)
    assert r["completed_iterations"] == 12 and r["accepted_iterations"] == 0
    assert r["value_evaluations"] == 0 and r["gradient_evaluations"] == 2 + 11
    assert r["momentum_restarts"] == 0 and r["rejected_candidates"] == 0
    assert np.array_equal(c, d.center)
    pen = [h["penalty"] for h in hist]
    assert pen[0] == 1.0  # no density gradient at all: the reference penalty falls back to 1
    expect = [1.0]
    for _ in range(11):
        expect.append(min(max(expect[-1] * 1.05, 0.5), 1.2))
    assert pen == expect and pen[-1] == 1.2
    for h in hist:  # 0.6 from the cell, 0.002 from the terminals' fixed area
        assert h["overflow"] == pytest.approx(0.602, abs=1e-12) and not h["accepted"] and h["step_size"] == 0.0


@pytest.mark.parametrize("case", ["plain-default", "plain-aggressive", "mid-quadratic"])
def test_adaptive_run_composes_the_pinned_laws(case):
    """The oracle's whole adaptive run, row by row, against the pinned laws (see controller_checks):
    objective, smoothing, penalty, checkpoint, restart and displacement composition.  The
    aggressive options are the reference's own controller test (PlacementFlowTests.cpp:197-216):
    larger moves and a strict Armijo coefficient, so backtracking and momentum restarts occur."""
    from controller_checks import check_history_composition
    if case.startswith("plain"):
        d = design.plain_design()
        d.center = ob.random_placement(d, 2026)
        o = abi.default_options(target_density=0.9, bins_x=64, bins_y=64)
        if case == "plain-aggressive":
            o.initial_movement_in_bins = 1.0
            o.maximum_movement_in_bins = 1.0
            o.armijo_coefficient = 0.5
            o.objective_increase_for_density = 0.0
    else:
        d = design.large_design(seed=3, cells=5000, nets=5250, row_count=80, sites=300, macros=3)
        d.center, _ = ob.quadratic_placement(d)
        o = abi.default_options(target_density=0.8, bins_x=32, bins_y=32, iterations=150)
    o.feasible_refinement_iterations = 2 ** 31 - 1
    _, r, hist = ob.global_place(d, o)
    core = d.core_region()
    stats = check_history_composition(o, r, hist, (core[2] - core[0]) / o.bins_x, (core[3] - core[1]) / o.bins_y)
    assert stats["accepted"] > 10 and stats["penalty_checks"] == len(hist) - 1 - 1 * bool(hist[0]["accepted"])
    if case == "plain-aggressive":
        assert r["rejected_candidates"] > 0 and r["momentum_restarts"] > 0
