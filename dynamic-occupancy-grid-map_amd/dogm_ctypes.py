"""ctypes binding of libdogm_hip.so (the C-ABI declared in include/dogm/dogm_c.h).

Mirrors the reference's dogm::DOGM surface (dogm/include/dogm/dogm.h:20-199): ``Params``,
``update_grid`` (updateGrid), the public stage methods, and the read-backs.  There is no fallback:
if the HIP library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("DOGM_HIP_LIB", _PKG / "lib" / "libdogm_hip.so"))

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python dynamic-occupancy-grid-map_amd/build.py` "
        "(the GPU path has no CPU fallback)")

lib = C.CDLL(str(LIB_PATH))

DOGM_OK = 0
DOGM_ERR_INVALID_ARGUMENT = 1
DOGM_DEFAULT_SEED = 123456
CDF_MODES = {"exact": 0, "fp64": 1, "fp32": 2, "fp32c": 3}
EGO_MODES = {"reference": 0, "metric": 1}
RESAMPLING_MODES = {"multinomial": 0, "systematic": 1}


class dogm_params(C.Structure):
    _fields_ = [
        ("size", C.c_float),
        ("resolution", C.c_float),
        ("particle_count", C.c_int),
        ("new_born_particle_count", C.c_int),
        ("persistence_prob", C.c_float),
        ("stddev_process_noise_position", C.c_float),
        ("stddev_process_noise_velocity", C.c_float),
        ("birth_prob", C.c_float),
        ("stddev_velocity", C.c_float),
        ("init_max_velocity", C.c_float),
        ("freespace_discount", C.c_float),
    ]


class dogm_lidar_params(C.Structure):
    _fields_ = [("grid_size", C.c_int), ("beam_count", C.c_int), ("resolution", C.c_double),
                ("max_range", C.c_double), ("occ_mass", C.c_float), ("free_mass", C.c_float)]


class dogm_server_config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_int)]


class dogm_frame_result(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("vehicle_id", C.c_int), ("status", C.c_int), ("batch_size", C.c_int),
                ("pad", C.c_int), ("t_arrival", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


GRID_CELL_DTYPE = np.dtype([
    ("start_idx", "<i4"), ("end_idx", "<i4"), ("new_born_occ_mass", "<f4"), ("pers_occ_mass", "<f4"),
    ("free_mass", "<f4"), ("occ_mass", "<f4"), ("pred_occ_mass", "<f4"), ("mu_A", "<f4"), ("mu_UA", "<f4"),
    ("w_A", "<f4"), ("w_UA", "<f4"), ("mean_x_vel", "<f4"), ("mean_y_vel", "<f4"), ("var_x_vel", "<f4"),
    ("var_y_vel", "<f4"), ("covar_xy_vel", "<f4"),
])
MEAS_CELL_DTYPE = np.dtype([("free_mass", "<f4"), ("occ_mass", "<f4"), ("likelihood", "<f4"), ("p_A", "<f4")])
VELOCITY_CELL_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("data", GRID_CELL_DTYPE)])
assert GRID_CELL_DTYPE.itemsize == 64 and MEAS_CELL_DTYPE.itemsize == 16 and VELOCITY_CELL_DTYPE.itemsize == 72

CLUSTER_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("vx", "<f4"), ("vy", "<f4"), ("cells", "<i4")])
OBJECT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("vx", "<f4"), ("vy", "<f4")])
SOURCES = {"grid": 0, "measurement": 1}
EXTRACT_OCCUPANCY, EXTRACT_IMAGE, EXTRACT_VELOCITY = 1, 2, 4

_vp = C.c_void_p
_h = C.c_void_p
_i = C.c_int
_f = C.c_float


class dogm_tuning(C.Structure):
    _fields_ = [("sort", C.c_int), ("merge_slice", C.c_int), ("pred_block", C.c_int), ("seg_items", C.c_int),
                ("lanes", C.c_int)]


_SIGS = {
    "dogm_create": (_i, [C.POINTER(dogm_params), C.POINTER(_h)]),
    "dogm_create_ex": (_i, [C.POINTER(dogm_params), C.c_uint64, C.c_uint32, C.POINTER(_h)]),
    "dogm_destroy": (None, [_h]),
    "dogm_last_error": (C.c_char_p, [_h]),
    "dogm_status_string": (C.c_char_p, [_i]),
    "dogm_update_grid": (_i, [_h, _vp, _f, _f, _f, _f, _i]),
    "dogm_update_grid_noalloc": (_i, [_h, _vp, _f, _f, _f, _f, _i]),
    "dogm_update_measurement_grid": (_i, [_h, _vp, _i]),
    "dogm_update_pose": (_i, [_h, _f, _f, _f]),
    "dogm_initialize_particles": (_i, [_h]),
    "dogm_particle_prediction": (_i, [_h, _f]),
    "dogm_particle_assignment": (_i, [_h]),
    "dogm_grid_cell_occupancy_update": (_i, [_h, _f]),
    "dogm_update_persistent_particles": (_i, [_h]),
    "dogm_initialize_new_particles": (_i, [_h]),
    "dogm_statistical_moments": (_i, [_h]),
    "dogm_resampling": (_i, [_h]),
    "dogm_get_grid_cells": (_i, [_h, _vp]),
    "dogm_get_measurement_cells": (_i, [_h, _vp]),
    "dogm_get_particles": (_i, [_h, _vp, _vp, _vp, _vp]),
    "dogm_get_grid_size": (_i, [_h]),
    "dogm_get_particle_count": (_i, [_h]),
    "dogm_get_new_born_particle_count": (_i, [_h]),
    "dogm_get_resolution": (_f, [_h]),
    "dogm_get_position_x": (_f, [_h]),
    "dogm_get_position_y": (_f, [_h]),
    "dogm_get_yaw": (_f, [_h]),
    "dogm_get_birth_particles": (_i, [_h, _vp, _vp, _vp, _vp]),
    "dogm_get_weight_array": (_i, [_h, _vp]),
    "dogm_get_born_masses": (_i, [_h, _vp]),
    "dogm_get_resampled_indices": (_i, [_h, _vp]),
    "dogm_set_particles": (_i, [_h, _vp, _vp, _vp, _vp]),
    "dogm_set_birth_particles": (_i, [_h, _vp, _vp, _vp, _vp]),
    "dogm_set_grid_cells": (_i, [_h, _vp]),
    "dogm_set_weight_array": (_i, [_h, _vp]),
    "dogm_set_born_masses": (_i, [_h, _vp]),
    "dogm_get_occupancy": (_i, [_h, _i, _vp]),
    "dogm_get_occupancy_image": (_i, [_h, _i, _vp]),
    "dogm_get_cells_with_velocity": (_i, [_h, _f, _f, _vp, _i, C.POINTER(_i)]),
    "dogm_dbscan": (_i, [_vp, _i, _f, _i, _vp, C.POINTER(_i)]),
    "dogm_cluster_means": (_i, [_vp, _i, _vp, _i, _vp]),
    "dogm_evaluator_create": (_i, [_f, _f, _f, _i, C.POINTER(_h)]),
    "dogm_evaluator_destroy": (None, [_h]),
    "dogm_evaluator_add_step": (_i, [_h, _vp, _i, _vp, _i, _vp, _i, C.POINTER(_i)]),
    "dogm_evaluator_summary": (_i, [_h, _vp, _vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "dogm_lidar_create": (_i, [C.POINTER(dogm_lidar_params), _vp, _vp, C.POINTER(_h)]),
    "dogm_lidar_destroy": (None, [_h]),
    "dogm_lidar_last_error": (C.c_char_p, [_h]),
    "dogm_lidar_generate": (_i, [_h, _vp, _vp]),
    "dogm_lidar_generate_many": (_i, [_h, _i, _vp, _vp]),
    "dogm_device_malloc": (_i, [C.POINTER(_vp), C.c_size_t]),
    "dogm_device_free": (_i, [_vp]),
    "dogm_memcpy": (_i, [_vp, _vp, C.c_size_t, _i]),
    "dogm_device_copy": (_i, [_vp, _vp, C.c_size_t]),
    "dogm_device_synchronize": (_i, []),
    "dogm_kernel_count": (_i, []),
    "dogm_kernel_name": (C.c_char_p, [_i]),
    "dogm_set_kernel_timing": (_i, [_h, _i, _i]),
    "dogm_get_kernel_timing": (_i, [_h, _i, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "dogm_reset_kernel_timing": (_i, [_h]),
    "dogm_set_cdf_mode": (_i, [_h, _i]),
    "dogm_set_ego_mode": (_i, [_h, _i]),
    "dogm_set_resampling_mode": (_i, [_h, _i]),
    "dogm_set_graph": (_i, [_h, _i]),
    "dogm_batch_set_graph": (_i, [_h, _i]),
    "dogm_get_resampling_mode": (_i, [_h]),
    "dogm_batch_set_resampling_mode": (_i, [_h, _i, _i]),
    "dogm_get_ego_mode": (_i, [_h]),
    "dogm_batch_set_ego_mode": (_i, [_h, _i, _i]),
    "dogm_get_cdf_mode": (_i, [_h]),
    "dogm_get_epoch": (C.c_uint32, [_h]),
    "dogm_batch_set_cdf_mode": (_i, [_h, _i, _i]),
    "dogm_batch_create": (_i, [C.POINTER(_h)]),
    "dogm_batch_destroy": (None, [_h]),
    "dogm_batch_last_error": (C.c_char_p, [_h]),
    "dogm_vehicle_footprint": (_i, [C.POINTER(dogm_params), C.POINTER(C.c_uint64)]),
    "dogm_batch_add_vehicle": (_i, [_h, C.POINTER(dogm_params), C.c_uint64, C.c_uint32, C.POINTER(_i)]),
    "dogm_batch_remove_vehicle": (_i, [_h, _i]),
    "dogm_batch_vehicle_count": (_i, [_h]),
    "dogm_batch_device_bytes": (C.c_uint64, [_h]),
    "dogm_batch_measurement_buffer": (_i, [_h, _i, C.POINTER(_vp)]),
    "dogm_batch_update": (_i, [_h, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "dogm_batch_get_grid_cells": (_i, [_h, _i, _vp]),
    "dogm_batch_get_particles": (_i, [_h, _i, _vp, _vp, _vp, _vp]),
    "dogm_batch_get_resampled_indices": (_i, [_h, _i, _vp]),
    "dogm_batch_set_lanes": (_i, [_h, _i]),
    "dogm_batch_set_extraction": (_i, [_h, _i, _f, _f]),
    "dogm_batch_get_occupancy": (_i, [_h, _i, _vp]),
    "dogm_batch_get_occupancy_image": (_i, [_h, _i, _vp]),
    "dogm_batch_get_cells_with_velocity": (_i, [_h, _i, _vp, _i, C.POINTER(_i)]),
    "dogm_batch_set_kernel_timing": (_i, [_h, _i, _i]),
    "dogm_batch_get_kernel_timing": (_i, [_h, _i, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "dogm_batch_reset_kernel_timing": (_i, [_h]),
    "dogm_server_create": (_i, [C.POINTER(dogm_server_config), C.POINTER(_h)]),
    "dogm_server_destroy": (None, [_h]),
    "dogm_server_last_error": (C.c_char_p, [_h]),
    "dogm_server_clock": (C.c_double, [_h]),
    "dogm_server_admit": (_i, [_h, C.POINTER(dogm_params), C.c_uint64, C.c_uint32, C.POINTER(_i)]),
    "dogm_server_release": (_i, [_h, _i]),
    "dogm_server_resident_bytes": (C.c_uint64, [_h]),
    "dogm_server_submit": (_i, [_h, _i, _vp, _i, _f, _f, _f, _f, C.c_double, C.POINTER(C.c_uint64)]),
    "dogm_server_submit_scan": (_i, [_h, _i, _h, _vp, _f, _f, _f, _f, C.c_double, C.POINTER(C.c_uint64)]),
    "dogm_server_queued": (_i, [_h]),
    "dogm_server_step": (_i, [_h, C.c_double, C.POINTER(_i), C.POINTER(C.c_double)]),
    "dogm_server_start": (_i, [_h]),
    "dogm_server_stop": (_i, [_h]),
    "dogm_server_wait_idle": (_i, [_h]),
    "dogm_server_poll": (_i, [_h, C.POINTER(dogm_frame_result), _i, C.POINTER(_i)]),
    "dogm_server_get_grid_cells": (_i, [_h, _i, _vp]),
    "dogm_server_get_particles": (_i, [_h, _i, _vp, _vp, _vp, _vp]),
    "dogm_server_stats": (_i, [_h, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "dogm_server_set_extraction": (_i, [_h, _i, _f, _f]),
    "dogm_server_get_occupancy": (_i, [_h, _i, _vp]),
    "dogm_server_get_occupancy_image": (_i, [_h, _i, _vp]),
    "dogm_server_get_cells_with_velocity": (_i, [_h, _i, _vp, _i, C.POINTER(_i)]),
    "dogm_tuning_load": (_i, [C.c_char_p]),
    "dogm_tuning_override": (_i, [C.POINTER(dogm_tuning)]),
    "dogm_tuning_lookup": (_i, [_i, C.c_longlong, C.POINTER(dogm_tuning)]),
    "dogm_tuning_entries": (_i, []),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


class DogmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{lib.dogm_status_string(status).decode()}: {msg}")
        self.status = status


@dataclass
class Params:
    """dogm::DOGM::Params (dogm/include/dogm/dogm.h:26-60), same field order."""
    size: float = 50.0
    resolution: float = 0.2
    particle_count: int = 300_000
    new_born_particle_count: int = 30_000
    persistence_prob: float = 0.99
    stddev_process_noise_position: float = 0.1
    stddev_process_noise_velocity: float = 2.0
    birth_prob: float = 0.02
    stddev_velocity: float = 30.0
    init_max_velocity: float = 15.0
    freespace_discount: float = 0.01

    def to_c(self) -> dogm_params:
        return dogm_params(*[getattr(self, f.name) for f in fields(self)])


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class DeviceBuffer:
    """A device allocation made through the library (for device-resident measurement grids)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        st = lib.dogm_device_malloc(C.byref(p), nbytes)
        if st != DOGM_OK:
            raise DogmError(st, f"device_malloc({nbytes})")
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, host: np.ndarray) -> None:
        assert host.nbytes <= self.nbytes and host.flags.c_contiguous
        st = lib.dogm_memcpy(C.c_void_p(self.ptr), _ptr(host), host.nbytes, 0)
        if st != DOGM_OK:
            raise DogmError(st, "memcpy host->device")

    def free(self) -> None:
        if self.ptr:
            lib.dogm_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DOGM:
    """Python mirror of dogm::DOGM over the C-ABI."""

    def __init__(self, params: Params, seed: int = DOGM_DEFAULT_SEED, instance: int = 0):
        self.params = params
        h = C.c_void_p()
        self._cp = params.to_c()
        st = lib.dogm_create_ex(C.byref(self._cp), seed, instance, C.byref(h))
        if st != DOGM_OK:
            raise DogmError(st, "dogm_create_ex")
        self.h = h
        self.grid_size = lib.dogm_get_grid_size(h)
        self.grid_cell_count = self.grid_size * self.grid_size
        self.particle_count = lib.dogm_get_particle_count(h)
        self.new_born_particle_count = lib.dogm_get_new_born_particle_count(h)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.dogm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != DOGM_OK:
            raise DogmError(st, f"{what}: {lib.dogm_last_error(self.h).decode()}")

    # ---- per-frame update and stages -----------------------------------------------------------
    def update_grid(self, meas, x: float, y: float, yaw: float, dt: float) -> None:
        """meas: host structured array (MEAS_CELL_DTYPE) or a DeviceBuffer."""
        if isinstance(meas, DeviceBuffer):
            self._check(lib.dogm_update_grid(self.h, C.c_void_p(meas.ptr), x, y, yaw, dt, 1), "updateGrid")
        else:
            m = np.ascontiguousarray(meas, dtype=MEAS_CELL_DTYPE)
            self._check(lib.dogm_update_grid(self.h, _ptr(m), x, y, yaw, dt, 0), "updateGrid")

    def update_measurement_grid(self, meas) -> None:
        if isinstance(meas, DeviceBuffer):
            self._check(lib.dogm_update_measurement_grid(self.h, C.c_void_p(meas.ptr), 1), "updateMeasurementGrid")
        else:
            m = np.ascontiguousarray(meas, dtype=MEAS_CELL_DTYPE)
            self._check(lib.dogm_update_measurement_grid(self.h, _ptr(m), 0), "updateMeasurementGrid")

    def update_pose(self, x: float, y: float, yaw: float) -> None:
        self._check(lib.dogm_update_pose(self.h, x, y, yaw), "updatePose")

    def initialize_particles(self):
        self._check(lib.dogm_initialize_particles(self.h), "initializeParticles")

    def particle_prediction(self, dt: float):
        self._check(lib.dogm_particle_prediction(self.h, dt), "particlePrediction")

    def particle_assignment(self):
        self._check(lib.dogm_particle_assignment(self.h), "particleAssignment")

    def grid_cell_occupancy_update(self, dt: float):
        self._check(lib.dogm_grid_cell_occupancy_update(self.h, dt), "gridCellOccupancyUpdate")

    def update_persistent_particles(self):
        self._check(lib.dogm_update_persistent_particles(self.h), "updatePersistentParticles")

    def initialize_new_particles(self):
        self._check(lib.dogm_initialize_new_particles(self.h), "initializeNewParticles")

    def statistical_moments(self):
        self._check(lib.dogm_statistical_moments(self.h), "statisticalMoments")

    def resampling(self):
        self._check(lib.dogm_resampling(self.h), "resampling")

    # ---- read-backs ----------------------------------------------------------------------------
    def get_grid_cells(self) -> np.ndarray:
        out = np.empty(self.grid_cell_count, dtype=GRID_CELL_DTYPE)
        self._check(lib.dogm_get_grid_cells(self.h, _ptr(out)), "getGridCells")
        return out

    def get_measurement_cells(self) -> np.ndarray:
        out = np.empty(self.grid_cell_count, dtype=MEAS_CELL_DTYPE)
        self._check(lib.dogm_get_measurement_cells(self.h, _ptr(out)), "getMeasurementCells")
        return out

    # ---- result extraction (image_creation.cpp:19-128, on the device) ---------------------------
    def get_occupancy(self, source: str = "grid") -> np.ndarray:
        out = np.empty(self.grid_cell_count, np.float32)
        self._check(lib.dogm_get_occupancy(self.h, SOURCES[source], _ptr(out)), "getOccupancy")
        return out

    def get_occupancy_image(self, source: str = "grid") -> np.ndarray:
        out = np.empty(self.grid_cell_count, np.uint8)
        self._check(lib.dogm_get_occupancy_image(self.h, SOURCES[source], _ptr(out)), "getOccupancyImage")
        return out

    def get_cells_with_velocity(self, min_occupancy: float = 0.7, min_velocity: float = 4.0,
                                capacity: int | None = None) -> tuple[np.ndarray, int]:
        """(records, count): records of the passing cells (at most `capacity`), count of all."""
        return _velocity_cells(lambda out, cap, cnt: lib.dogm_get_cells_with_velocity(
            self.h, min_occupancy, min_velocity, out, cap, cnt), self.grid_cell_count, capacity, self._check)

    def _get_set(self, fn, n: int, what: str) -> dict:
        state = np.empty((n, 4), np.float32)
        idx = np.empty(n, np.int32)
        w = np.empty(n, np.float32)
        a = np.empty(n, np.uint8)
        self._check(fn(self.h, _ptr(state), _ptr(idx), _ptr(w), _ptr(a)), what)
        return {"state": state, "grid_cell_idx": idx, "weight": w, "associated": a}

    def get_particles(self) -> dict:
        return self._get_set(lib.dogm_get_particles, self.particle_count, "getParticles")

    def get_birth_particles(self) -> dict:
        return self._get_set(lib.dogm_get_birth_particles, self.new_born_particle_count, "getBirthParticles")

    def set_particles(self, state=None, grid_cell_idx=None, weight=None, associated=None) -> None:
        n = self.particle_count
        args = _as_particle_arrays(n, state, grid_cell_idx, weight, associated)
        self._check(lib.dogm_set_particles(self.h, *map(_ptr, args)), "setParticles")

    def set_birth_particles(self, state=None, grid_cell_idx=None, weight=None, associated=None) -> None:
        n = self.new_born_particle_count
        args = _as_particle_arrays(n, state, grid_cell_idx, weight, associated)
        self._check(lib.dogm_set_birth_particles(self.h, *map(_ptr, args)), "setBirthParticles")

    def set_grid_cells(self, cells: np.ndarray) -> None:
        c = np.ascontiguousarray(cells, dtype=GRID_CELL_DTYPE)
        assert c.shape == (self.grid_cell_count,)
        self._check(lib.dogm_set_grid_cells(self.h, _ptr(c)), "setGridCells")

    def get_weight_array(self) -> np.ndarray:
        out = np.empty(self.particle_count, np.float32)
        self._check(lib.dogm_get_weight_array(self.h, _ptr(out)), "getWeightArray")
        return out

    def set_weight_array(self, w) -> None:
        a = np.ascontiguousarray(w, dtype=np.float32)
        assert a.shape == (self.particle_count,)
        self._check(lib.dogm_set_weight_array(self.h, _ptr(a)), "setWeightArray")

    def get_born_masses(self) -> np.ndarray:
        out = np.empty(self.grid_cell_count, np.float32)
        self._check(lib.dogm_get_born_masses(self.h, _ptr(out)), "getBornMasses")
        return out

    def set_born_masses(self, m) -> None:
        a = np.ascontiguousarray(m, dtype=np.float32)
        assert a.shape == (self.grid_cell_count,)
        self._check(lib.dogm_set_born_masses(self.h, _ptr(a)), "setBornMasses")

    def get_resampled_indices(self) -> np.ndarray:
        out = np.empty(self.particle_count, np.int32)
        self._check(lib.dogm_get_resampled_indices(self.h, _ptr(out)), "getResampledIndices")
        return out

    def set_ego_mode(self, mode: str) -> None:
        self._check(lib.dogm_set_ego_mode(self.h, EGO_MODES[mode]), "setEgoMode")

    def set_resampling_mode(self, mode: str) -> None:
        self._check(lib.dogm_set_resampling_mode(self.h, RESAMPLING_MODES[mode]), "setResamplingMode")

    def set_graph(self, enable: bool = True) -> None:
        self._check(lib.dogm_set_graph(self.h, int(enable)), "setGraph")

    def set_cdf_mode(self, mode: str) -> None:
        """Resampling CDF arithmetic: 'exact' (default), 'fp64' or 'fp32'."""
        self._check(lib.dogm_set_cdf_mode(self.h, CDF_MODES[mode]), "setCdfMode")

    def position(self) -> tuple[float, float, float]:
        return (lib.dogm_get_position_x(self.h), lib.dogm_get_position_y(self.h), lib.dogm_get_yaw(self.h))

    # ---- kernel timing -------------------------------------------------------------------------
    @staticmethod
    def kernel_names() -> list[str]:
        return [lib.dogm_kernel_name(i).decode() for i in range(lib.dogm_kernel_count())]

    def set_kernel_timing(self, name: str, enable: bool = True) -> None:
        self._check(lib.dogm_set_kernel_timing(self.h, self.kernel_names().index(name), int(enable)), "timing")

    def kernel_timing(self, name: str) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_longlong()
        self._check(lib.dogm_get_kernel_timing(self.h, self.kernel_names().index(name), C.byref(ms), C.byref(n)),
                    "timing")
        return ms.value, n.value

    def reset_kernel_timing(self) -> None:
        self._check(lib.dogm_reset_kernel_timing(self.h), "timing")


def vehicle_footprint(params: Params) -> int:
    """Device bytes one vehicle with these params holds (dogm_vehicle_footprint)."""
    cp = params.to_c()
    b = C.c_uint64()
    st = lib.dogm_vehicle_footprint(C.byref(cp), C.byref(b))
    if st != DOGM_OK:
        raise DogmError(st, "vehicle_footprint")
    return int(b.value)


class Batch:
    """Many independent grids ("vehicles") updated in the same kernel launches (dogm_batch_*)."""

    def __init__(self):
        h = C.c_void_p()
        st = lib.dogm_batch_create(C.byref(h))
        if st != DOGM_OK:
            raise DogmError(st, "dogm_batch_create")
        self.h = h
        self.vehicles: dict[int, Params] = {}

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.dogm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != DOGM_OK:
            raise DogmError(st, f"{what}: {lib.dogm_batch_last_error(self.h).decode()}")

    def add_vehicle(self, params: Params, seed: int = DOGM_DEFAULT_SEED, instance: int = 0) -> int:
        cp = params.to_c()
        vid = C.c_int()
        self._check(lib.dogm_batch_add_vehicle(self.h, C.byref(cp), seed, instance, C.byref(vid)), "addVehicle")
        self.vehicles[vid.value] = params
        return vid.value

    def remove_vehicle(self, vid: int) -> None:
        self._check(lib.dogm_batch_remove_vehicle(self.h, vid), "removeVehicle")
        self.vehicles.pop(vid, None)

    def device_bytes(self) -> int:
        return int(lib.dogm_batch_device_bytes(self.h))

    def measurement_buffer(self, vid: int) -> int:
        p = C.c_void_p()
        self._check(lib.dogm_batch_measurement_buffer(self.h, vid, C.byref(p)), "measurementBuffer")
        return p.value

    def grid_cell_count(self, vid: int) -> int:
        p = self.vehicles[vid]
        gs = int(np.float32(p.size) / np.float32(p.resolution))
        return gs * gs

    def update(self, vids, meas, poses, dts) -> None:
        """One frame for the vehicles `vids`.  meas[i]: host structured array, DeviceBuffer, or None
        (the grid already sits in the vehicle's measurement buffer); poses[i] = (x, y, yaw)."""
        n = len(vids)
        ids = np.ascontiguousarray(vids, dtype=np.int32)
        pose = np.asarray(poses, dtype=np.float32).reshape(n, 3)
        xs, ys, yaws = (np.ascontiguousarray(pose[:, k]) for k in range(3))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float32), (n,)))
        host = [m for m in meas if m is not None and not isinstance(m, DeviceBuffer)]
        dev = [m for m in meas if isinstance(m, DeviceBuffer)]
        if host and dev:
            raise ValueError("one step takes host or device measurement grids, not both")
        keep = []
        ptrs = (C.c_void_p * n)()
        for i, m in enumerate(meas):
            if m is None:
                ptrs[i] = None
            elif isinstance(m, DeviceBuffer):
                ptrs[i] = m.ptr
            else:
                a = np.ascontiguousarray(m, dtype=MEAS_CELL_DTYPE)
                keep.append(a)
                ptrs[i] = a.ctypes.data
        self._check(lib.dogm_batch_update(self.h, n, _ptr(ids), C.cast(ptrs, C.c_void_p), _ptr(xs), _ptr(ys),
                                          _ptr(yaws), _ptr(dt), 1 if dev else 0), "batchUpdate")

    def get_grid_cells(self, vid: int) -> np.ndarray:
        out = np.empty(self.grid_cell_count(vid), dtype=GRID_CELL_DTYPE)
        self._check(lib.dogm_batch_get_grid_cells(self.h, vid, _ptr(out)), "getGridCells")
        return out

    def get_particles(self, vid: int) -> dict:
        n = self.vehicles[vid].particle_count
        state = np.empty((n, 4), np.float32)
        idx = np.empty(n, np.int32)
        w = np.empty(n, np.float32)
        a = np.empty(n, np.uint8)
        self._check(lib.dogm_batch_get_particles(self.h, vid, _ptr(state), _ptr(idx), _ptr(w), _ptr(a)),
                    "getParticles")
        return {"state": state, "grid_cell_idx": idx, "weight": w, "associated": a}

    def get_resampled_indices(self, vid: int) -> np.ndarray:
        out = np.empty(self.vehicles[vid].particle_count, np.int32)
        self._check(lib.dogm_batch_get_resampled_indices(self.h, vid, _ptr(out)), "getResampledIndices")
        return out

    def set_cdf_mode(self, vid: int, mode: str) -> None:
        self._check(lib.dogm_batch_set_cdf_mode(self.h, vid, CDF_MODES[mode]), "setCdfMode")

    def set_ego_mode(self, vid: int, mode: str) -> None:
        self._check(lib.dogm_batch_set_ego_mode(self.h, vid, EGO_MODES[mode]), "setEgoMode")

    def set_resampling_mode(self, vid: int, mode: str) -> None:
        self._check(lib.dogm_batch_set_resampling_mode(self.h, vid, RESAMPLING_MODES[mode]), "setResamplingMode")

    def set_extraction(self, what: int, min_occupancy: float = 0.7, min_velocity: float = 4.0) -> None:
        self._check(lib.dogm_batch_set_extraction(self.h, what, min_occupancy, min_velocity), "setExtraction")

    def set_lanes(self, lanes: int) -> None:
        self._check(lib.dogm_batch_set_lanes(self.h, lanes), "setLanes")

    def set_graph(self, enable: bool = True) -> None:
        self._check(lib.dogm_batch_set_graph(self.h, int(enable)), "setGraph")

    def get_occupancy(self, vid: int) -> np.ndarray:
        out = np.empty(self.grid_cell_count(vid), np.float32)
        self._check(lib.dogm_batch_get_occupancy(self.h, vid, _ptr(out)), "getOccupancy")
        return out

    def get_occupancy_image(self, vid: int) -> np.ndarray:
        out = np.empty(self.grid_cell_count(vid), np.uint8)
        self._check(lib.dogm_batch_get_occupancy_image(self.h, vid, _ptr(out)), "getOccupancyImage")
        return out

    def get_cells_with_velocity(self, vid: int, capacity: int | None = None) -> tuple[np.ndarray, int]:
        return _velocity_cells(lambda out, cap, cnt: lib.dogm_batch_get_cells_with_velocity(
            self.h, vid, out, cap, cnt), self.grid_cell_count(vid), capacity, self._check)

    def set_kernel_timing(self, name: str, enable: bool = True) -> None:
        self._check(lib.dogm_batch_set_kernel_timing(self.h, DOGM.kernel_names().index(name), int(enable)),
                    "timing")

    def kernel_timing(self, name: str) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_longlong()
        self._check(lib.dogm_batch_get_kernel_timing(self.h, DOGM.kernel_names().index(name), C.byref(ms),
                                                     C.byref(n)), "timing")
        return ms.value, n.value

    def reset_kernel_timing(self) -> None:
        self._check(lib.dogm_batch_reset_kernel_timing(self.h), "timing")


class Server:
    """Continuous-batching server over resident vehicles (dogm_server_*)."""

    def __init__(self, hbm_budget_bytes: int, max_batch: int = 0):
        cfg = dogm_server_config(hbm_budget_bytes, max_batch)
        h = C.c_void_p()
        st = lib.dogm_server_create(C.byref(cfg), C.byref(h))
        if st != DOGM_OK:
            raise DogmError(st, "dogm_server_create")
        self.h = h
        self.vehicles: dict[int, Params] = {}
        self._keep: dict[int, object] = {}

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.dogm_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != DOGM_OK:
            raise DogmError(st, f"{what}: {lib.dogm_server_last_error(self.h).decode()}")

    def clock(self) -> float:
        return lib.dogm_server_clock(self.h)

    def admit(self, params: Params, seed: int = DOGM_DEFAULT_SEED, instance: int = 0) -> int | None:
        """Vehicle id, or None when the footprint does not fit the remaining budget."""
        cp = params.to_c()
        vid = C.c_int()
        st = lib.dogm_server_admit(self.h, C.byref(cp), seed, instance, C.byref(vid))
        if st == 2:  # DOGM_ERR_OUT_OF_MEMORY: over budget
            return None
        self._check(st, "admit")
        self.vehicles[vid.value] = params
        return vid.value

    def release(self, vid: int) -> None:
        self._check(lib.dogm_server_release(self.h, vid), "release")
        self.vehicles.pop(vid, None)

    def resident_bytes(self) -> int:
        return int(lib.dogm_server_resident_bytes(self.h))

    def submit(self, vid: int, meas, pose, dt: float, t_arrival: float) -> int:
        """meas: DeviceBuffer or host structured array (kept alive until the frame completes)."""
        x, y, yaw = pose
        t = C.c_uint64()
        if isinstance(meas, DeviceBuffer):
            st = lib.dogm_server_submit(self.h, vid, C.c_void_p(meas.ptr), 1, x, y, yaw, dt, t_arrival, C.byref(t))
            keep = meas
        else:
            keep = np.ascontiguousarray(meas, dtype=MEAS_CELL_DTYPE)
            st = lib.dogm_server_submit(self.h, vid, _ptr(keep), 0, x, y, yaw, dt, t_arrival, C.byref(t))
        self._check(st, "submit")
        self._keep[t.value] = keep
        return t.value

    def submit_scan(self, vid: int, lidar: "Lidar", ranges, pose, dt: float, t_arrival: float) -> int:
        """A frame given as lidar ranges; its grid is generated on the device inside the step."""
        x, y, yaw = pose
        r = np.ascontiguousarray(ranges, np.float64)
        assert r.size == lidar.beam_count
        t = C.c_uint64()
        self._check(lib.dogm_server_submit_scan(self.h, vid, lidar.h, _ptr(r), x, y, yaw, dt, t_arrival, C.byref(t)),
                    "submitScan")
        return t.value

    def queued(self) -> int:
        return lib.dogm_server_queued(self.h)

    def step(self, now: float) -> tuple[int, float]:
        n = C.c_int()
        secs = C.c_double()
        self._check(lib.dogm_server_step(self.h, now, C.byref(n), C.byref(secs)), "step")
        return n.value, secs.value

    def start(self) -> None:
        self._check(lib.dogm_server_start(self.h), "start")

    def stop(self) -> None:
        self._check(lib.dogm_server_stop(self.h), "stop")

    def wait_idle(self) -> None:
        self._check(lib.dogm_server_wait_idle(self.h), "wait_idle")

    def poll(self, max_results: int = 4096) -> list[dict]:
        buf = (dogm_frame_result * max_results)()
        n = C.c_int()
        self._check(lib.dogm_server_poll(self.h, buf, max_results, C.byref(n)), "poll")
        out = []
        for i in range(n.value):
            r = buf[i]
            self._keep.pop(r.ticket, None)
            out.append({"ticket": r.ticket, "vehicle_id": r.vehicle_id, "status": r.status,
                        "batch_size": r.batch_size, "t_arrival": r.t_arrival, "t_start": r.t_start,
                        "t_done": r.t_done})
        return out

    def get_grid_cells(self, vid: int) -> np.ndarray:
        p = self.vehicles[vid]
        gs = int(np.float32(p.size) / np.float32(p.resolution))
        out = np.empty(gs * gs, dtype=GRID_CELL_DTYPE)
        self._check(lib.dogm_server_get_grid_cells(self.h, vid, _ptr(out)), "getGridCells")
        return out

    def get_particles(self, vid: int) -> dict:
        n = self.vehicles[vid].particle_count
        state = np.empty((n, 4), np.float32)
        idx = np.empty(n, np.int32)
        w = np.empty(n, np.float32)
        a = np.empty(n, np.uint8)
        self._check(lib.dogm_server_get_particles(self.h, vid, _ptr(state), _ptr(idx), _ptr(w), _ptr(a)),
                    "getParticles")
        return {"state": state, "grid_cell_idx": idx, "weight": w, "associated": a}

    def stats(self) -> tuple[int, int]:
        a, b = C.c_longlong(), C.c_longlong()
        self._check(lib.dogm_server_stats(self.h, C.byref(a), C.byref(b)), "stats")
        return a.value, b.value

    def set_extraction(self, what: int, min_occupancy: float = 0.7, min_velocity: float = 4.0) -> None:
        self._check(lib.dogm_server_set_extraction(self.h, what, min_occupancy, min_velocity), "setExtraction")

    def _cells(self, vid: int) -> int:
        p = self.vehicles[vid]
        gs = int(np.float32(p.size) / np.float32(p.resolution))
        return gs * gs

    def get_occupancy(self, vid: int) -> np.ndarray:
        out = np.empty(self._cells(vid), np.float32)
        self._check(lib.dogm_server_get_occupancy(self.h, vid, _ptr(out)), "getOccupancy")
        return out

    def get_occupancy_image(self, vid: int) -> np.ndarray:
        out = np.empty(self._cells(vid), np.uint8)
        self._check(lib.dogm_server_get_occupancy_image(self.h, vid, _ptr(out)), "getOccupancyImage")
        return out

    def get_cells_with_velocity(self, vid: int, capacity: int | None = None) -> tuple[np.ndarray, int]:
        return _velocity_cells(lambda out, cap, cnt: lib.dogm_server_get_cells_with_velocity(
            self.h, vid, out, cap, cnt), self._cells(vid), capacity, self._check)


def _velocity_cells(call, cells: int, capacity: int | None, check) -> tuple[np.ndarray, int]:
    cap = cells if capacity is None else int(capacity)
    out = np.empty(max(cap, 1), dtype=VELOCITY_CELL_DTYPE)
    cnt = C.c_int()
    check(call(_ptr(out), cap, C.byref(cnt)), "getCellsWithVelocity")
    return out[:min(cap, cnt.value)].copy(), cnt.value


def _as_particle_arrays(n, state, idx, weight, assoc):
    s = None if state is None else np.ascontiguousarray(state, dtype=np.float32).reshape(n, 4)
    i = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32).reshape(n)
    w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32).reshape(n)
    a = None if assoc is None else np.ascontiguousarray(assoc, dtype=np.uint8).reshape(n)
    return s, i, w, a


class Lidar:
    """Measurement-grid generator on the device (dogm_lidar_*): ranges in, measurement grids out."""

    def __init__(self, grid_size: int, beam_count: int, resolution: float, max_range: float, occ_mass: float = 0.7,
                 free_mass: float = 0.4, beam_cos=None, beam_sin=None):
        self.params = dogm_lidar_params(grid_size, beam_count, resolution, max_range, occ_mass, free_mass)
        self.grid_size = grid_size
        self.beam_count = beam_count
        cs = None if beam_cos is None else np.ascontiguousarray(beam_cos, np.float64)
        sn = None if beam_sin is None else np.ascontiguousarray(beam_sin, np.float64)
        h = C.c_void_p()
        st = lib.dogm_lidar_create(C.byref(self.params), _ptr(cs), _ptr(sn), C.byref(h))
        if st != DOGM_OK:
            raise DogmError(st, "dogm_lidar_create failed")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.dogm_lidar_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != DOGM_OK:
            raise DogmError(st, f"{what}: {lib.dogm_lidar_last_error(self.h).decode()}")

    def generate(self, ranges, device_out: int) -> None:
        """One scan into a device buffer of grid_size^2 measurement cells (its address)."""
        r = np.ascontiguousarray(ranges, np.float64)
        assert r.size == self.beam_count
        self._check(lib.dogm_lidar_generate(self.h, _ptr(r), C.c_void_p(device_out)), "generate")

    def generate_many(self, ranges, device_outs) -> None:
        r = np.ascontiguousarray(ranges, np.float64).reshape(-1)
        n = len(device_outs)
        assert r.size == n * self.beam_count
        outs = (C.c_void_p * n)(*device_outs)
        self._check(lib.dogm_lidar_generate_many(self.h, n, _ptr(r), C.cast(outs, C.c_void_p)), "generateMany")


def dbscan(cells: np.ndarray, eps: float = 3.0, min_cells: int = 5) -> tuple[np.ndarray, int]:
    """(cluster id per cell, -1 = noise; number of clusters) of velocity-cell records."""
    c = np.ascontiguousarray(cells, dtype=VELOCITY_CELL_DTYPE)
    ids = np.empty(max(c.size, 1), np.int32)
    k = C.c_int()
    st = lib.dogm_dbscan(_ptr(c), c.size, eps, min_cells, _ptr(ids), C.byref(k))
    if st != DOGM_OK:
        raise DogmError(st, "dogm_dbscan")
    return ids[:c.size].copy(), k.value


def cluster_means(cells: np.ndarray, ids: np.ndarray, clusters: int) -> np.ndarray:
    c = np.ascontiguousarray(cells, dtype=VELOCITY_CELL_DTYPE)
    i = np.ascontiguousarray(ids, dtype=np.int32)
    out = np.zeros(max(clusters, 1), dtype=CLUSTER_DTYPE)
    st = lib.dogm_cluster_means(_ptr(c), c.size, _ptr(i), clusters, _ptr(out))
    if st != DOGM_OK:
        raise DogmError(st, "dogm_cluster_means")
    return out[:clusters].copy()


class Evaluator:
    """DBSCAN + assignment + MAE / RMSE over frames (dogm_evaluator_*)."""

    def __init__(self, resolution: float, max_assignment_distance: float = 5.0, eps: float = 3.0,
                 min_cells: int = 5):
        h = C.c_void_p()
        st = lib.dogm_evaluator_create(resolution, max_assignment_distance, eps, min_cells, C.byref(h))
        if st != DOGM_OK:
            raise DogmError(st, "dogm_evaluator_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.dogm_evaluator_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_step(self, cells: np.ndarray, truth: np.ndarray) -> np.ndarray:
        """Clusters of the frame (metres) after accumulating their errors against `truth`."""
        c = np.ascontiguousarray(cells, dtype=VELOCITY_CELL_DTYPE)
        t = np.ascontiguousarray(truth, dtype=OBJECT_DTYPE)
        cap = max(c.size, 1)
        out = np.zeros(cap, dtype=CLUSTER_DTYPE)
        k = C.c_int()
        st = lib.dogm_evaluator_add_step(self.h, _ptr(c), c.size, _ptr(t), t.size, _ptr(out), cap, C.byref(k))
        if st != DOGM_OK:
            raise DogmError(st, "dogm_evaluator_add_step")
        return out[:k.value].copy()

    def summary(self) -> dict:
        mae = np.zeros(1, OBJECT_DTYPE)
        rmse = np.zeros(1, OBJECT_DTYPE)
        d, u = C.c_longlong(), C.c_longlong()
        lib.dogm_evaluator_summary(self.h, _ptr(mae), _ptr(rmse), C.byref(d), C.byref(u))
        as_dict = lambda a: {k: float(a[k][0]) for k in OBJECT_DTYPE.names}  # noqa: E731
        return {"mae": as_dict(mae), "rmse": as_dict(rmse), "detections": d.value, "unassigned": u.value}


# ---- tuning (dogm_tuning_*) -------------------------------------------------------------------------
SORT_FORMS = {"reduce_scan": 0, "onesweep": 1}
TUNING_KEYS = ("sort", "merge_slice", "pred_block", "seg_items", "lanes")


def tuning_dict(t: dogm_tuning) -> dict:
    d = {k: getattr(t, k) for k in TUNING_KEYS}
    d["sort"] = {v: k for k, v in SORT_FORMS.items()}[d["sort"]]
    return d


TUNING_BUILTIN = {"sort": "reduce_scan", "merge_slice": 0, "pred_block": 512, "seg_items": 8, "lanes": 2}


def _tuning_c(d: dict) -> dogm_tuning:
    vals = dict(TUNING_BUILTIN)
    vals.update(d)
    return dogm_tuning(SORT_FORMS[vals["sort"]], int(vals["merge_slice"]), int(vals["pred_block"]),
                       int(vals["seg_items"]), int(vals["lanes"]))


def tuning_override(d: dict | None) -> None:
    """Force the launch choices `d` (keys of TUNING_KEYS; missing keys take the built-in choice) for
    every following step of the process; None returns to the table."""
    if d is None:
        st = lib.dogm_tuning_override(None)
    else:
        t = _tuning_c(d)
        st = lib.dogm_tuning_override(C.byref(t))
    if st != DOGM_OK:
        raise DogmError(st, f"dogm_tuning_override({d})")


def tuning_lookup(vehicles: int, particles: int) -> dict:
    t = dogm_tuning()
    st = lib.dogm_tuning_lookup(vehicles, particles, C.byref(t))
    if st != DOGM_OK:
        raise DogmError(st, "dogm_tuning_lookup")
    return tuning_dict(t)


def tuning_load(path: str | None) -> None:
    st = lib.dogm_tuning_load(None if path is None else str(path).encode())
    if st != DOGM_OK:
        raise DogmError(st, f"dogm_tuning_load({path})")
