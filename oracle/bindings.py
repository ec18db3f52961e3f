"""ctypes bindings of the CPU checker libraries.

TEST INFRASTRUCTURE.  Imported only by tests/, __graft_entry__.smoke() and the CPU-baseline leg
of bench.py.  liboracle.so is the FP64 restatement (gp_oracle.cpp); _ref/libmpref.so is the
reference's own density/filler/parser code (ref_shim.cpp), present only where it was built;
_ref/libmpadapter.so is the MyPlacement-side adapter over the product library, compiled against
the reference's headers and sources (adapter_shim.cpp), likewise.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from myplacement_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmpref.so")
ADAPTER_SO = os.path.join(HERE, "_ref", "libmpadapter.so")

_dp = C.POINTER(C.c_double)
_oracle = None
_ref = None
_adapter = None


class OracleError(RuntimeError):
    pass


def _check(status: int, lib) -> None:
    if status != 0:
        msg = lib.orc_last_error().decode() if hasattr(lib, "orc_last_error") else lib.ref_last_error().decode()
        if status == abi.MP_GP_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise OracleError(msg)


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise OracleError(f"{ORACLE_SO} is not built (run __graft_entry__.build())")
        lib = C.CDLL(ORACLE_SO)
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_session_create.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions), C.POINTER(C.c_void_p)]
        lib.orc_session_destroy.argtypes = [C.c_void_p]
        lib.orc_session_info.argtypes = [C.c_void_p, C.POINTER(abi.MpPreparedInfo)]
        lib.orc_session_layout.argtypes = [C.c_void_p, C.POINTER(abi.MpStatic)]
        lib.orc_session_static_grids.argtypes = [C.c_void_p, _dp, _dp]
        lib.orc_session_arrays.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_session_evaluate.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int, C.POINTER(abi.MpScalars),
                                             C.c_void_p, C.c_void_p]
        lib.orc_session_grids.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_global_place.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions), _dp,
                                         C.POINTER(abi.MpResult), C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        lib.orc_random_placement.argtypes = [C.POINTER(abi.MpDesign), C.c_uint64, _dp]
        lib.orc_quadratic_placement.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpQpOptions), _dp,
                                                C.POINTER(abi.MpQpResult)]
        lib.orc_design_metrics.argtypes = [C.POINTER(abi.MpDesign), C.c_int, C.c_int, C.c_double, _dp,
                                           C.POINTER(abi.MpDensityMetrics)]
        lib.orc_transform_1d.argtypes = [C.c_int, C.c_int, _dp, _dp]
        lib.orc_transform_2d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        lib.orc_field_solve.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        lib.orc_sample_field.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int64, _dp, _dp]
        lib.orc_density_map.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_int64, _dp, C.c_int64, _dp,
                                        C.POINTER(C.c_int32), C.POINTER(abi.MpDensityMetrics),
                                        C.POINTER(abi.MpDensityMetrics), C.c_void_p]
        lib.orc_set_method.argtypes = [C.c_int]
        lib.orc_estimate_step.argtypes = [_dp, _dp, _dp, _dp, C.c_int64, C.c_int, C.c_double]
        lib.orc_estimate_step.restype = C.c_double
        lib.orc_smoothing.argtypes = [C.POINTER(abi.MpOptions), C.c_double, C.c_double, C.c_double]
        lib.orc_smoothing.restype = C.c_double
        lib.orc_accepts.argtypes = [C.POINTER(abi.MpOptions), C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double]
        lib.orc_update_penalty.argtypes = [C.POINTER(abi.MpOptions)] + [C.c_double] * 7
        lib.orc_update_penalty.restype = C.c_double
        lib.orc_checkpoint_sequence.argtypes = [C.POINTER(abi.MpOptions), C.c_int64, _dp, _dp, C.c_void_p, C.c_void_p]
        lib.orc_checkpoint_sequence.restype = C.c_int32
        lib.orc_detailed_place.argtypes = [C.POINTER(abi.MpDesign), C.c_void_p, C.POINTER(abi.MpDpOptions), _dp,
                                           C.c_void_p, C.POINTER(abi.MpDpResult)]
        _oracle = lib
    return _oracle


def reference():
    """The reference's own code (None when oracle/_ref was not built here)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        lib = C.CDLL(REF_SO)
        lib.ref_last_error.restype = C.c_char_p
        lib.ref_density_map.argtypes = oracle().orc_density_map.argtypes
        lib.ref_fillers.argtypes = [C.POINTER(abi.MpDesign), C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64,
                                    C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.c_int64]
        lib.ref_deposit.argtypes = [C.POINTER(abi.MpDesign), C.c_int, C.c_int, C.c_double, C.c_int64, _dp,
                                    C.c_void_p, C.POINTER(abi.MpDensityMetrics), C.POINTER(abi.MpDensityMetrics)]
        lib.ref_design_metrics.argtypes = [C.POINTER(abi.MpDesign), C.c_int, C.c_int, C.c_double, _dp,
                                           C.POINTER(abi.MpDensityMetrics)]
        lib.ref_derived.argtypes = [C.POINTER(abi.MpDesign), C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, _dp, _dp]
        lib.ref_parse_aux.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        lib.ref_db_destroy.argtypes = [C.c_void_p]
        lib.ref_db_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.ref_db_export.argtypes = [C.c_void_p] + [C.c_void_p] * 12
        lib.ref_legalize.argtypes = [C.POINTER(abi.MpDesign), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp,
                                     C.c_void_p, _dp]
        lib.ref_legality.argtypes = [C.POINTER(abi.MpDesign), C.c_double, _dp]
        lib.ref_detailed_place.argtypes = [C.POINTER(abi.MpDesign), C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_double,
                                           _dp, C.c_void_p, _dp]
        _ref = lib
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ---------------------------------------------------------------------------------------------
# transforms / field

def transform_1d(kind: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    _check(oracle().orc_transform_1d(kind, x.shape[0], abi.dptr(x), abi.dptr(out)), oracle())
    return out


def transform_2d(kind_y: int, kind_x: int, grid: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(grid, dtype=np.float64)
    out = np.empty_like(g)
    _check(oracle().orc_transform_2d(g.shape[1], g.shape[0], kind_y, kind_x, abi.dptr(g), abi.dptr(out)), oracle())
    return out


def set_method(method: int) -> None:
    """0 auto (FFT for powers of two), 1 direct O(N^2), 2 FFT."""
    oracle().orc_set_method(method)


def field_solve(source: np.ndarray, region) -> tuple:
    s = np.ascontiguousarray(source, dtype=np.float64)
    rows, cols = s.shape
    reg = np.asarray(region, dtype=np.float64)
    pot, fx, fy = np.empty_like(s), np.empty_like(s), np.empty_like(s)
    _check(oracle().orc_field_solve(cols, rows, abi.dptr(reg), abi.dptr(s), abi.dptr(pot), abi.dptr(fx),
                                    abi.dptr(fy)), oracle())
    return pot, fx, fy


def sample_field(fx: np.ndarray, fy: np.ndarray, region, xy: np.ndarray) -> np.ndarray:
    fx = np.ascontiguousarray(fx, dtype=np.float64)
    fy = np.ascontiguousarray(fy, dtype=np.float64)
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    reg = np.asarray(region, dtype=np.float64)
    out = np.empty_like(xy)
    rows, cols = fx.shape
    _check(oracle().orc_sample_field(cols, rows, abi.dptr(reg), abi.dptr(fx), abi.dptr(fy), xy.shape[0],
                                     abi.dptr(xy), abi.dptr(out)), oracle())
    return out


def density_map(region, cols, rows, t, regions, rects, layers, lib=None):
    """DensityMap driver: returns (metrics_with, metrics_without, bins[G,5])."""
    lib = lib or oracle()
    reg = np.asarray(region, dtype=np.float64)
    regions = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1, 4))
    rects = np.ascontiguousarray(np.asarray(rects, dtype=np.float64).reshape(-1, 4))
    layers = np.ascontiguousarray(np.asarray(layers, dtype=np.int32))
    w, wo = abi.MpDensityMetrics(), abi.MpDensityMetrics()
    bins = np.zeros((cols * rows, 5))
    fn = lib.orc_density_map if hasattr(lib, "orc_density_map") else lib.ref_density_map
    _check(fn(abi.dptr(reg), cols, rows, t, regions.shape[0], abi.dptr(regions), rects.shape[0], abi.dptr(rects),
              layers.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(w), C.byref(wo), _p(bins)), lib)
    return w.as_dict(), wo.as_dict(), bins


# ---------------------------------------------------------------------------------------------
# design-level

def random_placement(design, seed: int = 2026) -> np.ndarray:
    c = design.center.copy()
    d = design.to_c()
    _check(oracle().orc_random_placement(C.byref(d), seed, abi.dptr(c)), oracle())
    return c


def quadratic_placement(design, options=None):
    """InitialPlacer quadraticPlacement (InitialPlacer.cpp:315-402) on the CPU; returns (centres, result)."""
    o = abi.default_qp_options() if options is None else options
    c = np.ascontiguousarray(design.center, dtype=np.float64).copy()
    d = design.to_c(c)
    r = abi.MpQpResult()
    _check(oracle().orc_quadratic_placement(C.byref(d), C.byref(o), abi.dptr(c), C.byref(r)), oracle())
    return c, r.as_dict()


def design_metrics(design, center, cols, rows, t) -> tuple:
    d = design.to_c(center)
    hpwl = C.c_double()
    m = abi.MpDensityMetrics()
    _check(oracle().orc_design_metrics(C.byref(d), cols, rows, t, C.byref(hpwl), C.byref(m)), oracle())
    return hpwl.value, m.as_dict()


class OracleSession:
    """GlobalPlacementSession on the CPU: static density, fillers, layout and evaluate."""

    def __init__(self, design, options: abi.MpOptions, center=None):
        self._design = design
        self._c = design.to_c(center)
        self._opts = options
        h = C.c_void_p()
        _check(oracle().orc_session_create(C.byref(self._c), C.byref(options), C.byref(h)), oracle())
        self.h = h
        info = abi.MpPreparedInfo()
        _check(oracle().orc_session_info(self.h, C.byref(info)), oracle())
        self.info = info

    def __del__(self):
        if getattr(self, "h", None):
            oracle().orc_session_destroy(self.h)
            self.h = None

    @property
    def particle_count(self) -> int:
        return int(self.info.particle_count)

    def layout(self) -> dict:
        L = abi.MpStatic()
        _check(oracle().orc_session_layout(self.h, C.byref(L)), oracle())
        P, N, I = L.particle_count, L.net_count, L.pin_count
        G = L.bins_x * L.bins_y

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)

        fixed, dark = np.empty(G), np.empty(G)
        _check(oracle().orc_session_static_grids(self.h, abi.dptr(fixed), abi.dptr(dark)), oracle())
        return dict(
            particle_width=arr(L.particle_width, P, np.float64), particle_height=arr(L.particle_height, P, np.float64),
            particle_charge_area=arr(L.particle_charge_area, P, np.float64),
            particle_layer=arr(L.particle_layer, P, np.int32), net_offsets=arr(L.net_offsets, N + 1, np.int32),
            net_weights=arr(L.net_weights, N, np.float64), pin_particles=arr(L.pin_particles, I, np.int32),
            pin_offsets_or_static=arr(L.pin_offsets_or_static, 2 * I, np.float64).reshape(-1, 2),
            bins_x=L.bins_x, bins_y=L.bins_y,
            region=(L.region_ll_x, L.region_ll_y, L.region_ur_x, L.region_ur_y), target_density=L.target_density,
            fixed_area=fixed, dark_area=dark,
        )

    def arrays(self) -> dict:
        M, F, P = self.info.movable_count, self.info.filler_count, self.info.particle_count
        mov = np.empty(M, dtype=np.int64)
        fc = np.empty((F, 2))
        pre = np.empty(P)
        pos = np.empty((P, 2))
        _check(oracle().orc_session_arrays(self.h, _p(mov), _p(fc), _p(pre), _p(pos)), oracle())
        return dict(movable_ids=mov, filler_centers=fc, preconditioner=pre, positions=pos)

    def evaluate(self, xy: np.ndarray, smoothing: float, want_grad: bool = True) -> dict:
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        s = abi.MpScalars()
        wg = np.zeros_like(xy) if want_grad else None
        dg = np.zeros_like(xy) if want_grad else None
        _check(oracle().orc_session_evaluate(self.h, abi.dptr(xy), smoothing, int(want_grad), C.byref(s), _p(wg),
                                             _p(dg)), oracle())
        return dict(hpwl=s.hpwl, smooth_wirelength=s.smooth_wirelength, electrostatic_energy=s.electrostatic_energy,
                    optimizer_density=s.optimizer_density.as_dict(), design_density=s.design_density.as_dict(),
                    wire_gradient=wg, density_gradient=dg)

    def grids(self) -> dict:
        G = int(self.info.bin_count)
        src, pot, fx, fy = (np.empty(G) for _ in range(4))
        _check(oracle().orc_session_grids(self.h, _p(src), _p(pot), _p(fx), _p(fy)), oracle())
        return dict(source=src, potential=pot, field_x=fx, field_y=fy)


def global_place(design, options: abi.MpOptions, center=None, history_capacity: int = 100000):
    c = np.ascontiguousarray(design.center if center is None else center, dtype=np.float64).copy()
    d = design.to_c(c)
    res = abi.MpResult()
    hist = (abi.MpHistoryRow * history_capacity)()
    n = C.c_int64()
    _check(oracle().orc_global_place(C.byref(d), C.byref(options), abi.dptr(c), C.byref(res), hist,
                                     history_capacity, C.byref(n)), oracle())
    rows = [hist[i].as_dict() for i in range(min(n.value, history_capacity))]
    return c, res.as_dict(), rows


# ---------------------------------------------------------------------------------------------
# after global placement: the reference's own legalizer and CPU detailed placer (oracle/_ref),
# and the restatement of the device window-reordering pass

LEGALITY_FIELDS = ("legal", "placed_standard_cells", "unplaced_standard_cells", "overlap_pairs",
                   "out_of_core_modules", "off_row_modules", "abacus_reverse_pass_selected",
                   "standard_cell_total_displacement", "standard_cell_total_squared_displacement",
                   "standard_cell_weighted_squared_displacement", "standard_cell_maximum_displacement",
                   "elapsed_seconds")


def _placed(design, center, orientation):
    c = np.ascontiguousarray(design.center if center is None else center, dtype=np.float64).reshape(-1, 2).copy()
    o = np.ascontiguousarray(design.orientation if orientation is None else orientation, dtype=np.int32).copy()
    return c, o


def detailed_place(design, options: abi.MpDpOptions, center=None, orientation=None, row_orientation=None):
    """The reference's device window-reordering pass restated on the CPU (gp_oracle.cpp
    detailedPlace); returns (centres, orientations, result dict)."""
    c, o = _placed(design, center, orientation)
    d = design.to_c(c, o)
    ro = None if row_orientation is None else np.ascontiguousarray(row_orientation, dtype=np.int32)
    out_c, out_o = c.copy(), o.copy()
    r = abi.MpDpResult()
    _check(oracle().orc_detailed_place(C.byref(d), _p(ro), C.byref(options), abi.dptr(out_c), _p(out_o), C.byref(r)),
           oracle())
    return out_c, out_o, r.as_dict()


def ref_legalize(design, center=None, orientation=None, strategy: int = 0, candidate_rows: int = 8,
                 macro_search_radius: int = 80, bidirectional: bool = True, epsilon: float = 1e-6):
    """Legalizer::legalize of the reference (Legalizer.cpp:707-738) with LegalizationOptions'
    defaults; returns (centres, orientations, report dict)."""
    lib = reference()
    if lib is None:
        raise OracleError("oracle/_ref was not built (the reference tree is absent)")
    c, o = _placed(design, center, orientation)
    d = design.to_c(c, o)
    out_c, out_o = c.copy(), o.copy()
    rep = np.zeros(12)
    _check(lib.ref_legalize(C.byref(d), strategy, candidate_rows, macro_search_radius, int(bidirectional), epsilon,
                            abi.dptr(out_c), _p(out_o), abi.dptr(rep)), lib)
    return out_c, out_o, dict(zip(LEGALITY_FIELDS, rep.tolist()))


def ref_legality(design, center=None, orientation=None, epsilon: float = 1e-6) -> dict:
    """Legalizer::check of the reference (Legalizer.cpp:740-805)."""
    lib = reference()
    if lib is None:
        raise OracleError("oracle/_ref was not built (the reference tree is absent)")
    c, o = _placed(design, center, orientation)
    d = design.to_c(c, o)
    rep = np.zeros(4)
    _check(lib.ref_legality(C.byref(d), epsilon, abi.dptr(rep)), lib)
    return dict(zip(("legal", "overlap_pairs", "out_of_core_modules", "off_row_modules"), rep.tolist()))


def ref_detailed_place(design, center=None, orientation=None, method: int = abi.MP_DP_WINDOW_REORDER, passes: int = 2,
                       window_size: int = 4, maximum_net_degree: int = 64, improvement_epsilon: float = 1e-9):
    """DetailedPlacer::run of the reference on the CPU (DetailedPlacer.cpp:289-335); returns
    (centres, orientations, result dict)."""
    lib = reference()
    if lib is None:
        raise OracleError("oracle/_ref was not built (the reference tree is absent)")
    c, o = _placed(design, center, orientation)
    d = design.to_c(c, o)
    out_c, out_o = c.copy(), o.copy()
    res = np.zeros(7)
    _check(lib.ref_detailed_place(C.byref(d), method, passes, window_size, maximum_net_degree, improvement_epsilon,
                                  abi.dptr(out_c), _p(out_o), abi.dptr(res)), lib)
    keys = ("hpwl_before", "hpwl_after", "evaluated_windows", "evaluated_permutations", "accepted_operations",
            "completed_passes", "elapsed_seconds")
    return out_c, out_o, dict(zip(keys, res.tolist()))


# ---------------------------------------------------------------------------------------------
# the reference-side adapter (myplacement_amd/integration/HipPlacementBackend.cpp) driven as
# GlobalPlacementSession drives a CudaPlacementBackend

def adapter():
    """The adapter library (None when oracle/_ref was not built here)."""
    global _adapter
    if _adapter is None:
        if not os.path.exists(ADAPTER_SO):
            return None
        lib = C.CDLL(ADAPTER_SO)
        V = C.c_void_p
        lib.adp_last_error.restype = C.c_char_p
        lib.adp_open.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.c_uint64,
                                 C.POINTER(V), C.c_char_p, C.c_size_t]
        lib.adp_layout.argtypes = [V, C.POINTER(abi.MpStatic)]
        lib.adp_positions.argtypes = [V, V]
        lib.adp_particle_count.argtypes = [V]
        lib.adp_particle_count.restype = C.c_int64
        lib.adp_evaluate.argtypes = [V, V, C.c_int64, C.c_double, C.c_int, C.POINTER(abi.MpScalars), V, V]
        lib.adp_device.argtypes = [V]
        lib.adp_reserved_bytes.argtypes = [V]
        lib.adp_reserved_bytes.restype = C.c_uint64
        lib.adp_close.argtypes = [V]
        _adapter = lib
    return _adapter


class AdapterRun:
    """A BookShelf design set up as GlobalPlacementSession does, with the backend from
    tryCreateHipPlacementBackend (or None and the factory's reason)."""

    def __init__(self, aux: str, bins: int, target_density: float, device: int = 0,
                 maximum_device_bytes: int = 40 << 30, maximum_fillers: int = 50000, seed: int = 2026):
        lib = adapter()
        h = C.c_void_p()
        reason = C.create_string_buffer(1024)
        self.status = lib.adp_open(aux.encode(), bins, bins, target_density, maximum_fillers, seed, device,
                                   maximum_device_bytes, C.byref(h), reason, 1024)
        if self.status in (1, 2):
            raise (ValueError if self.status == 1 else RuntimeError)(lib.adp_last_error().decode())
        self.h = h
        self.reason = reason.value.decode()
        self.created = self.status == 0

    def close(self):
        if getattr(self, "h", None):
            adapter().adp_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layout(self) -> dict:
        L = abi.MpStatic()
        adapter().adp_layout(self.h, C.byref(L))
        P, N, I = L.particle_count, L.net_count, L.pin_count
        G = L.bins_x * L.bins_y

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)

        return dict(
            particle_width=arr(L.particle_width, P, np.float64), particle_height=arr(L.particle_height, P, np.float64),
            particle_charge_area=arr(L.particle_charge_area, P, np.float64),
            particle_layer=arr(L.particle_layer, P, np.int32), net_offsets=arr(L.net_offsets, N + 1, np.int32),
            net_weights=arr(L.net_weights, N, np.float64), pin_particles=arr(L.pin_particles, I, np.int32),
            pin_offsets_or_static=arr(L.pin_offsets_or_static, 2 * I, np.float64).reshape(-1, 2),
            bins_x=L.bins_x, bins_y=L.bins_y,
            region=(L.region_ll_x, L.region_ll_y, L.region_ur_x, L.region_ur_y), target_density=L.target_density,
            fixed_area=arr(L.fixed_area, G, np.float64), dark_area=arr(L.dark_area, G, np.float64),
        )

    def positions(self) -> np.ndarray:
        xy = np.empty((int(adapter().adp_particle_count(self.h)), 2))
        adapter().adp_positions(self.h, xy.ctypes.data_as(C.c_void_p))
        return xy

    def evaluate(self, xy: np.ndarray, smoothing: float, want_grad: bool = True) -> dict:
        lib = adapter()
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        s = abi.MpScalars()
        wg = np.zeros_like(xy) if want_grad else None
        dg = np.zeros_like(xy) if want_grad else None
        st = lib.adp_evaluate(self.h, _p(xy), xy.shape[0], smoothing, int(want_grad), C.byref(s), _p(wg), _p(dg))
        if st:
            raise (ValueError if st == 1 else RuntimeError)(lib.adp_last_error().decode())
        return dict(hpwl=s.hpwl, smooth_wirelength=s.smooth_wirelength, electrostatic_energy=s.electrostatic_energy,
                    optimizer_density=s.optimizer_density.as_dict(), design_density=s.design_density.as_dict(),
                    wire_gradient=wg, density_gradient=dg)

    def device(self) -> int:
        return int(adapter().adp_device(self.h))

    def reserved_bytes(self) -> int:
        return int(adapter().adp_reserved_bytes(self.h))
