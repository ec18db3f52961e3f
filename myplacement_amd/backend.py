"""Python face of the HIP global-placement backend (ctypes over libmp_gp_hip.so).

Mirrors the reference's plugin boundary and its caller:
  HipPlacementBackend                 class CudaPlacementBackend   src/placement/CudaPlacementBackend.hpp:25-39
  try_create_hip_placement_backend    tryCreateCudaPlacementBackend               :45-48 (None + reason)
  hip_placement_backend_compiled      cudaPlacementBackendCompiled                :44
  PlacementEvaluation                 CudaPlacementEvaluation                     :15-23
  ComputeBackend / parse_compute_backend   include/myplacement/placement/GlobalPlacer.hpp:19,
                                      src/placement/GlobalPlacer.cpp:126-132 (with the `hip` value)
  global_place                        GlobalPlacer::run for the adaptive optimizer, GlobalPlacer.cpp:134-146

There is no CPU fallback here: without the built library every entry point raises.
"""

from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# MP_GP_LIBRARY names another build of the same library (the host-sanitized one the CPU tests use)
LIB_PATH = os.environ.get("MP_GP_LIBRARY") or os.path.join(HERE, "libmp_gp_hip.so")

_lib = None


class BackendUnavailable(RuntimeError):
    """The factory's recoverable failure (nullptr + reason in the reference)."""


def library():
    """The HIP backend library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m myplacement_amd.build`")
        lib = C.CDLL(LIB_PATH)
        V = C.c_void_p
        lib.mp_gp_compiled.restype = C.c_int
        lib.mp_gp_last_error.restype = C.c_char_p
        lib.mp_gp_create.argtypes = [C.POINTER(abi.MpStatic), C.POINTER(V)]
        lib.mp_gp_evaluate.argtypes = [V, V, C.c_double, C.c_int, C.POINTER(abi.MpScalars), V, V]
        lib.mp_gp_device.argtypes = [V]
        lib.mp_gp_reserved_bytes.argtypes = [V]
        lib.mp_gp_reserved_bytes.restype = C.c_uint64
        lib.mp_gp_destroy.argtypes = [V]
        lib.mp_gp_debug_grids.argtypes = [V, V, V, V, V]
        lib.mp_gp_default_options.argtypes = [C.POINTER(abi.MpOptions)]
        lib.mp_gp_prepare.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions), C.POINTER(V)]
        lib.mp_gp_prepared_destroy.argtypes = [V]
        lib.mp_gp_prepared_query.argtypes = [V, C.POINTER(abi.MpPreparedInfo)]
        lib.mp_gp_prepared_layout.argtypes = [V, C.POINTER(abi.MpStatic)]
        lib.mp_gp_prepared_arrays.argtypes = [V, V, V, V]
        lib.mp_gp_random_placement.argtypes = [C.POINTER(abi.MpDesign), C.c_uint64, V]
        lib.mp_gp_default_qp_options.argtypes = [C.POINTER(abi.MpQpOptions)]
        lib.mp_gp_default_server_options.argtypes = [C.POINTER(abi.MpServerOptions)]
        lib.mp_gp_estimate_device_bytes.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions)]
        lib.mp_gp_estimate_device_bytes.restype = C.c_uint64
        lib.mp_gp_server_create.argtypes = [C.POINTER(abi.MpServerOptions), C.POINTER(V)]
        lib.mp_gp_server_submit.argtypes = [V, C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions),
                                            C.POINTER(C.c_int64)]
        lib.mp_gp_server_wait.argtypes = [V, C.c_int64, V, C.POINTER(abi.MpJobReport)]
        lib.mp_gp_server_destroy.argtypes = [V]
        lib.mp_dp_default_options.argtypes = [C.POINTER(abi.MpDpOptions)]
        lib.mp_dp_run.argtypes = [C.POINTER(abi.MpDesign), V, C.POINTER(abi.MpDpOptions), V, V,
                                  C.POINTER(abi.MpDpResult)]
        lib.mp_dp_run_noalloc.argtypes = lib.mp_dp_run.argtypes
        lib.mp_dp_noalloc_scratch_bytes.restype = C.c_uint64
        lib.mp_gp_quadratic_placement.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpQpOptions), V,
                                                  C.POINTER(abi.MpQpResult)]
        lib.mp_gp_session_create.argtypes = [V, C.POINTER(abi.MpOptions), V, C.POINTER(V)]
        lib.mp_gp_session_step.argtypes = [V, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        lib.mp_gp_session_finish.argtypes = [V, V, C.POINTER(abi.MpResult)]
        lib.mp_gp_session_history.argtypes = [V, V, C.c_int64]
        lib.mp_gp_session_history.restype = C.c_int64
        lib.mp_gp_session_positions.argtypes = [V, V]
        lib.mp_gp_session_enable_kernel_timing.argtypes = [V, C.c_int]
        lib.mp_gp_session_enable_kernel_timing_every.argtypes = [V, C.c_int, C.c_int]
        lib.mp_gp_session_kernel_timing.argtypes = [V, C.c_char_p, C.c_size_t, V, V, V, C.c_int32]
        lib.mp_gp_session_kernel_timing.restype = C.c_int32
        lib.mp_gp_session_particle_order.argtypes = [V, V]
        lib.mp_gp_session_backend.argtypes = [V]
        lib.mp_gp_session_backend.restype = V
        lib.mp_gp_session_destroy.argtypes = [V]
        lib.mp_gp_global_place.argtypes = [C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions), V,
                                           C.POINTER(abi.MpResult), V, C.c_int64]
        lib.mp_gp_global_place_batch.argtypes = [C.c_int32, C.POINTER(abi.MpDesign), C.POINTER(abi.MpOptions),
                                                 C.POINTER(C.c_void_p), C.POINTER(abi.MpResult), C.POINTER(C.c_int32)]
        _lib = lib
    return _lib


def _raise(status: int) -> None:
    if status == abi.MP_GP_OK:
        return
    msg = library().mp_gp_last_error().decode()
    if status == abi.MP_GP_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == abi.MP_GP_UNAVAILABLE:
        raise BackendUnavailable(msg)
    raise RuntimeError(msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class ComputeBackend(enum.Enum):
    CPU = "cpu"
    CUDA = "cuda"
    AUTO = "auto"
    HIP = "hip"


def parse_compute_backend(value: str) -> ComputeBackend:
    """parseComputeBackend (GlobalPlacer.cpp:126-132) plus `hip` / `rocm`."""
    v = value.lower()
    if v == "cpu":
        return ComputeBackend.CPU
    if v in ("cuda", "gpu"):
        return ComputeBackend.CUDA
    if v == "auto":
        return ComputeBackend.AUTO
    if v in ("hip", "rocm"):
        return ComputeBackend.HIP
    raise ValueError(f"Unknown compute backend: {value}. Use cpu, cuda, hip, or auto.")


def hip_placement_backend_compiled() -> bool:
    return bool(library().mp_gp_compiled())


def default_options(**overrides) -> abi.MpOptions:
    o = abi.MpOptions()
    library().mp_gp_default_options(C.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


@dataclass
class PlacementEvaluation:
    """CudaPlacementEvaluation (CudaPlacementBackend.hpp:15-23)."""
    hpwl: float
    smooth_wirelength: float
    electrostatic_energy: float
    optimizer_density: dict
    design_density: dict
    wire_gradient: np.ndarray | None
    density_gradient: np.ndarray | None


def layout_struct(layout: dict, device: int = 0, dct_precision: int = abi.MP_DCT_FP32,
                  maximum_device_bytes: int = 40 * 1024 ** 3):
    """An MpStatic over numpy arrays in `layout` (keys as oracle.bindings.OracleSession.layout)."""
    keep = {
        "w": np.ascontiguousarray(layout["particle_width"], dtype=np.float64),
        "h": np.ascontiguousarray(layout["particle_height"], dtype=np.float64),
        "q": np.ascontiguousarray(layout["particle_charge_area"], dtype=np.float64),
        "l": np.ascontiguousarray(layout["particle_layer"], dtype=np.int32),
        "no": np.ascontiguousarray(layout["net_offsets"], dtype=np.int32),
        "nw": np.ascontiguousarray(layout["net_weights"], dtype=np.float64),
        "pp": np.ascontiguousarray(layout["pin_particles"], dtype=np.int32),
        "po": np.ascontiguousarray(np.asarray(layout["pin_offsets_or_static"], dtype=np.float64).reshape(-1)),
        "fx": np.ascontiguousarray(layout["fixed_area"], dtype=np.float64),
        "dk": np.ascontiguousarray(layout["dark_area"], dtype=np.float64),
    }
    reg = layout["region"]
    L = abi.MpStatic(
        particle_count=keep["w"].shape[0], particle_width=abi.dptr(keep["w"]), particle_height=abi.dptr(keep["h"]),
        particle_charge_area=abi.dptr(keep["q"]), particle_layer=abi.i32ptr(keep["l"]),
        net_count=keep["nw"].shape[0], net_offsets=abi.i32ptr(keep["no"]), net_weights=abi.dptr(keep["nw"]),
        pin_count=keep["pp"].shape[0], pin_particles=abi.i32ptr(keep["pp"]),
        pin_offsets_or_static=abi.dptr(keep["po"]), bins_x=layout["bins_x"], bins_y=layout["bins_y"],
        region_ll_x=reg[0], region_ll_y=reg[1], region_ur_x=reg[2], region_ur_y=reg[3],
        target_density=layout["target_density"], fixed_area=abi.dptr(keep["fx"]), dark_area=abi.dptr(keep["dk"]),
        device=device, dct_precision=dct_precision, maximum_device_bytes=maximum_device_bytes,
    )
    return L, keep


class HipPlacementBackend:
    """CudaPlacementBackend on HIP: evaluate(positions, smoothing, calculate_gradient)."""

    def __init__(self, handle, keep=None, owned: bool = True):
        self._h = handle
        self._keep = keep
        self._owned = owned
        self._particles = None

    @classmethod
    def create(cls, layout: dict, **kw) -> "HipPlacementBackend":
        L, keep = layout_struct(layout, **kw)
        h = C.c_void_p()
        _raise(library().mp_gp_create(C.byref(L), C.byref(h)))
        b = cls(h, keep)
        b._particles = int(L.particle_count)
        return b

    def close(self) -> None:
        if self._h and self._owned:
            library().mp_gp_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device(self) -> int:
        return int(library().mp_gp_device(self._h))

    def reserved_bytes(self) -> int:
        return int(library().mp_gp_reserved_bytes(self._h))

    def evaluate(self, positions: np.ndarray, smoothing: float, calculate_gradient: bool) -> PlacementEvaluation:
        xy = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
        if self._particles is not None and xy.shape[0] != self._particles:
            raise ValueError("HIP placement positions do not match the particle count.")
        s = abi.MpScalars()
        wg = np.empty_like(xy) if calculate_gradient else None
        dg = np.empty_like(xy) if calculate_gradient else None
        _raise(library().mp_gp_evaluate(self._h, _p(xy), float(smoothing), int(bool(calculate_gradient)),
                                        C.byref(s), _p(wg), _p(dg)))
        return PlacementEvaluation(s.hpwl, s.smooth_wirelength, s.electrostatic_energy,
                                   s.optimizer_density.as_dict(), s.design_density.as_dict(), wg, dg)

    def debug_grids(self, bins: int) -> dict:
        src, pot, fx, fy = (np.empty(bins) for _ in range(4))
        _raise(library().mp_gp_debug_grids(self._h, _p(src), _p(pot), _p(fx), _p(fy)))
        return dict(source=src, potential=pot, field_x=fx, field_y=fy)


def try_create_hip_placement_backend(layout: dict, **kw):
    """Returns (backend, "") or (None, reason) like the reference factory."""
    try:
        return HipPlacementBackend.create(layout, **kw), ""
    except BackendUnavailable as e:
        return None, str(e)


class PreparedDesign:
    """Static density, fillers, layout and preconditioner of a design (GlobalPlacementSession set-up)."""

    def __init__(self, design, options: abi.MpOptions, center=None):
        self._design = design
        self._c = design.to_c(center)
        self._opts = options
        h = C.c_void_p()
        _raise(library().mp_gp_prepare(C.byref(self._c), C.byref(options), C.byref(h)))
        self._h = h
        info = abi.MpPreparedInfo()
        _raise(library().mp_gp_prepared_query(self._h, C.byref(info)))
        self.info = info

    def __del__(self):
        if getattr(self, "_h", None):
            library().mp_gp_prepared_destroy(self._h)
            self._h = None

    def layout(self) -> dict:
        L = abi.MpStatic()
        L.maximum_device_bytes = 1
        _raise(library().mp_gp_prepared_layout(self._h, C.byref(L)))
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

    def arrays(self) -> dict:
        M, F, P = self.info.movable_count, self.info.filler_count, self.info.particle_count
        mov = np.empty(M, dtype=np.int64)
        fc = np.empty((F, 2))
        pre = np.empty(P)
        _raise(library().mp_gp_prepared_arrays(self._h, _p(mov), _p(fc), _p(pre)))
        return dict(movable_ids=mov, filler_centers=fc, preconditioner=pre)

    def initial_positions(self) -> np.ndarray:
        a = self.arrays()
        c = self._design.center
        return np.concatenate([c[a["movable_ids"]], a["filler_centers"]]).reshape(-1, 2)


class PlacementSession:
    """A device-resident adaptive GP run; step() advances whole outer iterations."""

    def __init__(self, prepared: PreparedDesign, options: abi.MpOptions):
        self._prepared = prepared
        self._opts = options
        h = C.c_void_p()
        _raise(library().mp_gp_session_create(prepared._h, C.byref(options), None, C.byref(h)))
        self._h = h
        self.stopped = False

    def close(self):
        if getattr(self, "_h", None):
            library().mp_gp_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, iterations: int) -> int:
        done, stopped = C.c_int32(), C.c_int32()
        _raise(library().mp_gp_session_step(self._h, iterations, C.byref(done), C.byref(stopped)))
        self.stopped = bool(stopped.value)
        return done.value

    def finish(self):
        c = self._prepared._design.center.copy()
        r = abi.MpResult()
        _raise(library().mp_gp_session_finish(self._h, _p(c), C.byref(r)))
        return c, r.as_dict()

    def history(self) -> list:
        n = library().mp_gp_session_history(self._h, None, 0)
        rows = (abi.MpHistoryRow * max(1, n))()
        n = library().mp_gp_session_history(self._h, rows, n)
        return [rows[i].as_dict() for i in range(n)]

    def positions(self) -> np.ndarray:
        xy = np.empty((self._prepared.info.particle_count, 2))
        _raise(library().mp_gp_session_positions(self._h, _p(xy)))
        return xy

    TIMING_CLASSES = ("k_wire", "k_density", "k_density_sort", "k_gemm", "k_gather", "controller", "k_wire_bins")

    def enable_kernel_timing(self, on: bool = True, classes=None, every: int = 1) -> None:
        """Times the named kernel classes (all when `classes` is None) with HIP events; with
        `every` > 1 only every every-th launch of each class."""
        mask = 0
        if on:
            mask = -1 if classes is None else sum(1 << self.TIMING_CLASSES.index(c) for c in classes)
        _raise(library().mp_gp_session_enable_kernel_timing_every(self._h, mask, int(every)))

    def kernel_timing(self) -> dict:
        names = C.create_string_buffer(256)
        tot = np.zeros(16)
        cnt = np.zeros(16, dtype=np.int64)
        work = np.zeros(16)
        n = library().mp_gp_session_kernel_timing(self._h, names, 256, _p(tot), _p(cnt), _p(work), 16)
        if n < 0:
            _raise(abi.MP_GP_RUNTIME_ERROR)
        keys = names.value.decode().split(",")
        return {k: dict(total_us=float(tot[i]), launches=int(cnt[i]), work=float(work[i])) for i, k in enumerate(keys[:n])}

    def particle_order(self) -> np.ndarray:
        """order[k] = prepared particle index of the session's device particle k."""
        o = np.zeros(int(self._prepared.info.particle_count), dtype=np.int64)
        _raise(library().mp_gp_session_particle_order(self._h, _p(o)))
        return o

    def backend(self) -> HipPlacementBackend:
        """The session's backend; it evaluates positions in the device order of particle_order()."""
        b = HipPlacementBackend(C.c_void_p(library().mp_gp_session_backend(self._h)), owned=False)
        b._particles = int(self._prepared.info.particle_count)
        return b


def random_placement(design, seed: int = 2026) -> np.ndarray:
    """InitialPlacer random placement (InitialPlacer.cpp:102-120)."""
    c = design.center.copy()
    d = design.to_c()
    _raise(library().mp_gp_random_placement(C.byref(d), seed, _p(c)))
    return c


def quadratic_placement(design, options: abi.MpQpOptions | None = None):
    """InitialPlacer quadraticPlacement (InitialPlacer.cpp:315-402) with the BiCGSTAB solves on the
    GPU; returns (centres, result dict)."""
    o = abi.default_qp_options() if options is None else options
    c = np.ascontiguousarray(design.center, dtype=np.float64).copy()
    d = design.to_c(c)
    r = abi.MpQpResult()
    _raise(library().mp_gp_quadratic_placement(C.byref(d), C.byref(o), _p(c), C.byref(r)))
    return c, r.as_dict()


def global_place(design, options: abi.MpOptions, center=None, history_capacity: int = 100000):
    """Adaptive GP on the HIP backend; returns (centres, result dict, history rows)."""
    c = np.ascontiguousarray(design.center if center is None else center, dtype=np.float64).copy()
    d = design.to_c(c)
    r = abi.MpResult()
    hist = (abi.MpHistoryRow * history_capacity)()
    _raise(library().mp_gp_global_place(C.byref(d), C.byref(options), _p(c), C.byref(r), hist, history_capacity))
    n = min(r.completed_iterations, history_capacity)
    return c, r.as_dict(), [hist[i].as_dict() for i in range(n)]


def global_place_batch(designs, options, centers=None):
    """mp_gp_global_place_batch: the adaptive GP of several independent designs in lockstep rounds,
    each stage one batched launch over all of them.  options: one MpOptions or one per design;
    centers: starting centres per design (default: each design's own).  Returns a list of
    (centres, result dict), one per design; raises the first failing job's error."""
    n = len(designs)
    opts = list(options) if isinstance(options, (list, tuple)) else [options] * n
    if len(opts) != n:
        raise ValueError("one options structure per design")
    cs = [np.ascontiguousarray(d.center if centers is None or centers[i] is None else centers[i],
                               dtype=np.float64).copy() for i, d in enumerate(designs)]
    ds = (abi.MpDesign * max(n, 1))()
    for i, d in enumerate(designs):
        ds[i] = d.to_c(cs[i])
    os_ = (abi.MpOptions * max(n, 1))(*opts)
    ptrs = (C.c_void_p * max(n, 1))(*[c.ctypes.data for c in cs])
    rs = (abi.MpResult * max(n, 1))()
    st = (C.c_int32 * max(n, 1))()
    _raise(library().mp_gp_global_place_batch(n, ds, os_, ptrs, rs, st))
    return [(cs[i], rs[i].as_dict()) for i in range(n)]


def detailed_place(design, options: abi.MpDpOptions | None = None, center=None, orientation=None,
                   row_orientation=None):
    """Window-reordering detailed placement on the device (tryRunCudaDetailedPlacement,
    CudaDetailedPlacementBackend.hpp:14-17) from a legal placement; returns (centres,
    orientations, result dict).  BackendUnavailable for a method or window size the device pass
    does not run, as the reference's pass declines them."""
    o = abi.default_dp_options() if options is None else options
    c = np.ascontiguousarray(design.center if center is None else center, dtype=np.float64).reshape(-1, 2).copy()
    ori = np.ascontiguousarray(design.orientation if orientation is None else orientation, dtype=np.int32).copy()
    d = design.to_c(c, ori)
    ro = None if row_orientation is None else np.ascontiguousarray(row_orientation, dtype=np.int32)
    out_c = c.copy()
    out_o = ori.copy()
    r = abi.MpDpResult()
    _raise(library().mp_dp_run(C.byref(d), _p(ro), C.byref(o), _p(out_c), _p(out_o), C.byref(r)))
    return out_c, out_o, r.as_dict()


def detailed_place_noalloc(design, options: abi.MpDpOptions | None = None, center=None, orientation=None,
                           row_orientation=None):
    """detailed_place through mp_dp_run_noalloc: the same pass with every device buffer carved from
    the library's fixed per-device scratch (no device allocation, null stream)."""
    o = abi.default_dp_options() if options is None else options
    c = np.ascontiguousarray(design.center if center is None else center, dtype=np.float64).reshape(-1, 2).copy()
    ori = np.ascontiguousarray(design.orientation if orientation is None else orientation, dtype=np.int32).copy()
    d = design.to_c(c, ori)
    ro = None if row_orientation is None else np.ascontiguousarray(row_orientation, dtype=np.int32)
    out_c = c.copy()
    out_o = ori.copy()
    r = abi.MpDpResult()
    _raise(library().mp_dp_run_noalloc(C.byref(d), _p(ro), C.byref(o), _p(out_c), _p(out_o), C.byref(r)))
    return out_c, out_o, r.as_dict()


def noalloc_scratch_bytes() -> int:
    return int(library().mp_dp_noalloc_scratch_bytes())


def estimate_device_bytes(design, options: abi.MpOptions) -> int:
    """Upper bound of the device bytes one global_place run of `design` reserves."""
    d = design.to_c()
    return int(library().mp_gp_estimate_device_bytes(C.byref(d), C.byref(options)))


class PlacementServer:
    """The placement service of mp_gp_server_*: jobs of mixed sizes admitted against a device
    memory budget into `max_active_jobs` slots; mode "streams" runs each on its own stream, mode
    "batched" runs all admitted jobs in lockstep rounds of batched launches.  `wait` releases the
    GIL while it blocks, so several Python threads may wait at once."""

    def __init__(self, max_active_jobs: int = 8, memory_budget_bytes: int = 200 << 30, device: int = 0,
                 mode: str = "streams"):
        o = abi.MpServerOptions()
        library().mp_gp_default_server_options(C.byref(o))
        o.max_active_jobs = max_active_jobs
        o.memory_budget_bytes = memory_budget_bytes
        o.device = device
        modes = {"streams": abi.MP_GP_SERVE_STREAMS, "batched": abi.MP_GP_SERVE_BATCHED}
        if mode not in modes:
            raise ValueError(f"unknown service mode {mode!r}")
        o.mode = modes[mode]
        h = C.c_void_p()
        _raise(library().mp_gp_server_create(C.byref(o), C.byref(h)))
        self._h = h

    def submit(self, design, options: abi.MpOptions) -> int:
        d = design.to_c()
        job = C.c_int64()
        _raise(library().mp_gp_server_submit(self._h, C.byref(d), C.byref(options), C.byref(job)))
        return job.value

    def wait(self, job: int, module_count: int):
        """Blocks until `job` is done; returns (centres [M, 2], report dict)."""
        c = np.zeros((module_count, 2))
        r = abi.MpJobReport()
        _raise(library().mp_gp_server_wait(self._h, job, _p(c), C.byref(r)))
        return c, r.as_dict()

    def close(self) -> None:
        if self._h:
            library().mp_gp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
