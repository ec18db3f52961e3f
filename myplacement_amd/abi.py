"""ctypes mirrors of the structs in include/mp_gp_backend.h.

The field order must match the header exactly; tests/test_abi.py checks the sizes against the
compiled library's own sizeof() report.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

MP_GP_OK = 0
MP_GP_INVALID_ARGUMENT = 1
MP_GP_RUNTIME_ERROR = 2
MP_GP_UNAVAILABLE = 3

MP_LAYER_STANDARD = 0
MP_LAYER_MACRO = 1
MP_LAYER_FILLER = 2

MP_DCT_FP32 = 0
MP_DCT_BF16X3 = 1

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class MpDesign(C.Structure):
    _fields_ = [
        ("module_count", C.c_int64),
        ("width", _dp),
        ("height", _dp),
        ("center", _dp),
        ("orientation", _i32p),
        ("is_terminal", _u8p),
        ("is_fixed", _u8p),
        ("net_count", C.c_int64),
        ("net_weight", _dp),
        ("net_pin_start", _i64p),
        ("pin_count", C.c_int64),
        ("pin_module", _i32p),
        ("pin_offset", _dp),
        ("row_count", C.c_int64),
        ("rows", _dp),
    ]


class MpStatic(C.Structure):
    _fields_ = [
        ("particle_count", C.c_int64),
        ("particle_width", _dp),
        ("particle_height", _dp),
        ("particle_charge_area", _dp),
        ("particle_layer", _i32p),
        ("net_count", C.c_int64),
        ("net_offsets", _i32p),
        ("net_weights", _dp),
        ("pin_count", C.c_int64),
        ("pin_particles", _i32p),
        ("pin_offsets_or_static", _dp),
        ("bins_x", C.c_int32),
        ("bins_y", C.c_int32),
        ("region_ll_x", C.c_double),
        ("region_ll_y", C.c_double),
        ("region_ur_x", C.c_double),
        ("region_ur_y", C.c_double),
        ("target_density", C.c_double),
        ("fixed_area", _dp),
        ("dark_area", _dp),
        ("device", C.c_int32),
        ("dct_precision", C.c_int32),
        ("maximum_device_bytes", C.c_uint64),
    ]


class MpDensityMetrics(C.Structure):
    _fields_ = [
        ("total_overflow_area", C.c_double),
        ("normalized_overflow", C.c_double),
        ("maximum_density", C.c_double),
        ("average_density", C.c_double),
        ("total_charge_area", C.c_double),
        ("normalization_area", C.c_double),
        ("placeable_area", C.c_double),
        ("dark_area", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class MpScalars(C.Structure):
    _fields_ = [
        ("hpwl", C.c_double),
        ("smooth_wirelength", C.c_double),
        ("electrostatic_energy", C.c_double),
        ("optimizer_density", MpDensityMetrics),
        ("design_density", MpDensityMetrics),
    ]


class MpOptions(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("bins_x", C.c_int32),
        ("bins_y", C.c_int32),
        ("target_density", C.c_double),
        ("maximum_bin_overflow", C.c_double),
        ("initial_smoothing", C.c_double),
        ("final_smoothing", C.c_double),
        ("maximum_movement_in_bins", C.c_double),
        ("maximum_backtracks", C.c_int32),
        ("backtracking_ratio", C.c_double),
        ("armijo_coefficient", C.c_double),
        ("maximum_momentum", C.c_double),
        ("initial_movement_in_bins", C.c_double),
        ("objective_increase_for_density", C.c_double),
        ("hpwl_growth_restart_threshold", C.c_double),
        ("density_stall_threshold", C.c_double),
        ("density_penalty_increase", C.c_double),
        ("density_penalty_decrease", C.c_double),
        ("minimum_density_penalty", C.c_double),
        ("maximum_density_penalty", C.c_double),
        ("smoothing_overflow_exponent", C.c_double),
        ("feasible_refinement_iterations", C.c_int32),
        ("maximum_fillers", C.c_uint64),
        ("seed", C.c_uint64),
        ("device", C.c_int32),
        ("dct_precision", C.c_int32),
        ("maximum_device_bytes", C.c_uint64),
        ("density_order_reuse_rows", C.c_double),
        ("capture_graphs", C.c_int32),
        ("recompute_density", C.c_int32),
    ]


HISTORY_COLUMNS = (
    "iteration", "hpwl", "overflow", "design_overflow", "smooth_wirelength", "density_energy", "objective",
    "penalty", "smoothing", "step_size", "maximum_displacement", "gradient_norm", "curvature",
    "rudy_energy", "rudy_proxy_overflow", "rudy_maximum_utilization", "rudy_weight",
    "backtracks", "momentum_restarted", "accepted", "best_checkpoint", "rudy_active",
)


class MpHistoryRow(C.Structure):
    _fields_ = [("iteration", C.c_int32)] + [
        (name, C.c_double) for name in HISTORY_COLUMNS[1:17]
    ] + [(name, C.c_int32) for name in HISTORY_COLUMNS[17:]]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name in HISTORY_COLUMNS}


class MpResult(C.Structure):
    _fields_ = [
        ("hpwl_before", C.c_double),
        ("hpwl_after", C.c_double),
        ("overflow_before", C.c_double),
        ("overflow_after", C.c_double),
        ("completed_iterations", C.c_int32),
        ("accepted_iterations", C.c_int32),
        ("momentum_restarts", C.c_int32),
        ("rejected_candidates", C.c_int32),
        ("best_checkpoint_iteration", C.c_int32),
        ("restored_best_checkpoint", C.c_int32),
        ("best_checkpoint_hpwl", C.c_double),
        ("best_checkpoint_overflow", C.c_double),
        ("elapsed_seconds", C.c_double),
        ("loop_seconds", C.c_double),
        ("gradient_evaluations", C.c_int64),
        ("value_evaluations", C.c_int64),
        ("particle_count", C.c_int64),
        ("filler_count", C.c_int64),
        ("device_used", C.c_int32),
        ("reserved_device_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class MpQpOptions(C.Structure):
    """mp_qp_options: InitialPlacementOptions, quadratic fields (InitialPlacer.hpp:21-25)."""
    _fields_ = [
        ("outer_iterations", C.c_int32),
        ("solver_iterations", C.c_int32),
        ("tolerance", C.c_double),
        ("minimum_distance", C.c_double),
        ("clique_degree_limit", C.c_uint64),
        ("device", C.c_int32),
    ]


def default_qp_options(**overrides) -> MpQpOptions:
    o = MpQpOptions(outer_iterations=5, solver_iterations=100, tolerance=1e-5, minimum_distance=1.0,
                    clique_degree_limit=96, device=0)
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


class MpQpResult(C.Structure):
    _fields_ = [
        ("hpwl_before", C.c_double),
        ("hpwl_after", C.c_double),
        ("elapsed_seconds", C.c_double),
        ("iterations", C.c_int32),
        ("solver_iterations", C.c_int32),
        ("relative_residual", C.c_double * 2),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "relative_residual"}
        d["relative_residual"] = (self.relative_residual[0], self.relative_residual[1])
        return d


MP_DP_NONE, MP_DP_ADJACENT_SWAP, MP_DP_WINDOW_REORDER = 0, 1, 2


class MpDpOptions(C.Structure):
    """mp_dp_options: DetailedPlacementOptions (DetailedPlacer.hpp:25-45)."""
    _fields_ = [
        ("method", C.c_int32),
        ("passes", C.c_int32),
        ("window_size", C.c_int32),
        ("maximum_net_degree", C.c_uint64),
        ("improvement_epsilon", C.c_double),
        ("device", C.c_int32),
        ("maximum_device_bytes", C.c_uint64),
    ]


def default_dp_options(**overrides) -> MpDpOptions:
    o = MpDpOptions(method=MP_DP_WINDOW_REORDER, passes=2, window_size=4, maximum_net_degree=64,
                    improvement_epsilon=1e-9, device=0, maximum_device_bytes=2 << 30)
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


class MpDpResult(C.Structure):
    """mp_dp_result: DetailedPlacementResult (DetailedPlacer.hpp:47-57)."""
    _fields_ = [
        ("hpwl_before", C.c_double),
        ("hpwl_after", C.c_double),
        ("elapsed_seconds", C.c_double),
        ("evaluated_windows", C.c_uint64),
        ("evaluated_permutations", C.c_uint64),
        ("accepted_operations", C.c_uint64),
        ("completed_passes", C.c_int32),
        ("device_used", C.c_int32),
        ("reserved_device_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


MP_GP_SERVE_STREAMS, MP_GP_SERVE_BATCHED = 0, 1


class MpServerOptions(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("max_active_jobs", C.c_int32),
        ("memory_budget_bytes", C.c_uint64),
        ("mode", C.c_int32),
        ("reserved", C.c_int32),
    ]


class MpJobReport(C.Structure):
    _fields_ = [
        ("submit_seconds", C.c_double),
        ("start_seconds", C.c_double),
        ("finish_seconds", C.c_double),
        ("queue_seconds", C.c_double),
        ("latency_seconds", C.c_double),
        ("estimated_device_bytes", C.c_uint64),
        ("result", MpResult),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "result"}
        d["result"] = self.result.as_dict()
        return d


class MpPreparedInfo(C.Structure):
    _fields_ = [
        ("movable_count", C.c_int64),
        ("filler_count", C.c_int64),
        ("particle_count", C.c_int64),
        ("net_count", C.c_int64),
        ("pin_count", C.c_int64),
        ("bin_count", C.c_int64),
        ("core_ll_x", C.c_double),
        ("core_ll_y", C.c_double),
        ("core_ur_x", C.c_double),
        ("core_ur_y", C.c_double),
        ("nominal_row_height", C.c_double),
        ("filler_width", C.c_double),
        ("filler_height", C.c_double),
    ]


def default_options(**overrides) -> MpOptions:
    """GlobalPlacementOptions defaults (GlobalPlacer.hpp:26-80)."""
    o = MpOptions(
        iterations=280, bins_x=64, bins_y=64, target_density=0.85, maximum_bin_overflow=0.10,
        initial_smoothing=0.75, final_smoothing=0.12, maximum_movement_in_bins=0.22, maximum_backtracks=6,
        backtracking_ratio=0.5, armijo_coefficient=1e-4, maximum_momentum=0.92, initial_movement_in_bins=0.12,
        objective_increase_for_density=0.002, hpwl_growth_restart_threshold=0.02, density_stall_threshold=0.002,
        density_penalty_increase=1.05, density_penalty_decrease=0.88, minimum_density_penalty=1e-4,
        maximum_density_penalty=1e6, smoothing_overflow_exponent=1.0, feasible_refinement_iterations=12,
        maximum_fillers=50000, seed=2026, device=0, dct_precision=MP_DCT_FP32,
        maximum_device_bytes=40 * 1024 ** 3, density_order_reuse_rows=2.5,
    )
    for key, value in overrides.items():
        setattr(o, key, value)
    return o


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def i32ptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32p)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64p)


def u8ptr(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)
