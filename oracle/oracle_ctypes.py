"""ctypes binding of the CPU checker (oracle/liboracle_dogm.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's host-core
baseline leg.  The product (dynamic-occupancy-grid-map_amd/) never imports it.  Method names match
dogm_ctypes.DOGM so parity tests drive both the same way.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle_dogm.so"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        vp, i, f = C.c_void_p, C.c_int, C.c_float
        sigs = {
            "oracle_create": (vp, [vp, C.c_uint64, C.c_uint32, i]),
            "oracle_destroy": (None, [vp]),
            "oracle_update_measurement_grid": (i, [vp, vp]),
            "oracle_update_pose": (None, [vp, f, f, f]),
            "oracle_initialize_particles": (None, [vp]),
            "oracle_particle_prediction": (None, [vp, f]),
            "oracle_particle_assignment": (None, [vp]),
            "oracle_grid_cell_occupancy_update": (None, [vp, f]),
            "oracle_update_persistent_particles": (None, [vp]),
            "oracle_initialize_new_particles": (None, [vp]),
            "oracle_statistical_moments": (None, [vp]),
            "oracle_resampling": (None, [vp]),
            "oracle_update_grid": (i, [vp, vp, f, f, f, f]),
            "oracle_get": (i, [vp, i, vp]),
            "oracle_set": (i, [vp, i, vp]),
            "oracle_get_i32": (i, [vp, i, vp]),
            "oracle_set_i32": (i, [vp, i, vp]),
            "oracle_get_epoch": (C.c_uint32, [vp]),
            "oracle_threads": (i, []),
            "oracle_set_epoch": (None, [vp, C.c_uint32]),
            "oracle_position_x": (f, [vp]),
            "oracle_position_y": (f, [vp]),
            "oracle_yaw": (f, [vp]),
            "oracle_set_flags": (None, [vp, i, i]),
            "oracle_set_cdf_mode": (None, [vp, i]),
            "oracle_spacing_fractions": (None, [C.c_uint64, C.c_uint32, C.c_uint32, i, vp]),
            "oracle_philox": (None, [C.c_uint32] * 6 + [vp]),
            "oracle_log_f32": (f, [f]),
            "oracle_sincos_2pi": (None, [f, C.POINTER(f), C.POINTER(f)]),
            "oracle_box_muller": (None, [C.c_uint32, C.c_uint32, C.POINTER(f), C.POINTER(f)]),
            "oracle_uniform01": (f, [C.c_uint32]),
            "oracle_uniform01_oc": (f, [C.c_uint32]),
            "oracle_fx_from_f32": (None, [f, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
            "oracle_fx_from_f32_nonneg_mismatches": (C.c_int, [C.c_void_p, C.c_int]),
            "oracle_fx_from_f64_nonneg_mismatches": (C.c_int, [C.c_void_p, C.c_int]),
            "oracle_exponential_sign_violations": (C.c_int, []),
            "oracle_fx_from_f64": (None, [C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
            "oracle_fx_to_f32": (f, [C.c_int64, C.c_uint64]),
            "oracle_fx_to_f64": (C.c_double, [C.c_int64, C.c_uint64]),
            "oracle_fx_scale_u64": (None, [C.c_int64, C.c_uint64, C.c_uint64, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_uint64)]),
            "oracle_order_end": (i, [C.c_int64, C.c_uint64, C.c_int64, C.c_uint64, i]),
            "oracle_occupancy": (None, [vp, vp, i, vp, vp]),
            "oracle_cells_with_velocity": (i, [i] + [vp] * 7 + [f, f, vp]),
            "oracle_velocity_normalized": (f, [f, f, f, f, f]),
            "oracle_lidar_grid": (None, [i, i, C.c_double, C.c_double, f, f, vp, vp, vp, vp]),
            "oracle_moment_sums": (None, [vp, i, vp, vp]),
            "oracle_set_ego_mode": (None, [vp, i]),
            "oracle_set_draw_mode": (None, [vp, i]),
            "oracle_two_sum_f32": (None, [f, f, C.POINTER(f), C.POINTER(f)]),
            "oracle_f32c_add": (None, [f, f, f, f, C.POINTER(f), C.POINTER(f)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# Field ids (oracle_dogm.cpp, enum Field).
F_P_X, F_P_Y, F_P_VX, F_P_VY, F_P_W = 0, 1, 2, 3, 4
F_B_X, F_B_Y, F_B_VX, F_B_VY, F_B_W = 5, 6, 7, 8, 9
F_WEIGHT_ARRAY, F_BORN = 10, 11
F_GRID0 = 16
GRID_FIELDS = ["new_born_occ_mass", "pers_occ_mass", "free_mass", "occ_mass", "pred_occ_mass", "mu_A", "mu_UA",
               "w_A", "w_UA", "mean_x_vel", "mean_y_vel", "var_x_vel", "var_y_vel", "covar_xy_vel"]
I_P_IDX, I_P_ASSOC, I_B_IDX, I_B_ASSOC, I_G_START, I_G_END, I_ANCESTORS = 100, 101, 102, 103, 104, 105, 106

GRID_CELL_DTYPE = np.dtype([
    ("start_idx", "<i4"), ("end_idx", "<i4"), ("new_born_occ_mass", "<f4"), ("pers_occ_mass", "<f4"),
    ("free_mass", "<f4"), ("occ_mass", "<f4"), ("pred_occ_mass", "<f4"), ("mu_A", "<f4"), ("mu_UA", "<f4"),
    ("w_A", "<f4"), ("w_UA", "<f4"), ("mean_x_vel", "<f4"), ("mean_y_vel", "<f4"), ("var_x_vel", "<f4"),
    ("var_y_vel", "<f4"), ("covar_xy_vel", "<f4"),
])
MEAS_CELL_DTYPE = np.dtype([("free_mass", "<f4"), ("occ_mass", "<f4"), ("likelihood", "<f4"), ("p_A", "<f4")])


class OracleDOGM:
    """The checker: fp64=False is the bit-exact f32 twin of the GPU path, fp64=True evaluates the
    same algorithm in double precision."""

    def __init__(self, params, seed: int = 123456, instance: int = 0, fp64: bool = False):
        self.params = params
        self._cp = params.to_c()
        self.fp64 = fp64
        self.h = lib().oracle_create(C.addressof(self._cp), seed, instance, int(fp64))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.grid_size = int(params.size / params.resolution)
        self.grid_cell_count = self.grid_size * self.grid_size
        self.particle_count = params.particle_count
        self.new_born_particle_count = params.new_born_particle_count

    def close(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- stages --------------------------------------------------------------------------------
    def update_grid(self, meas, x, y, yaw, dt):
        m = np.ascontiguousarray(meas, dtype=MEAS_CELL_DTYPE)
        lib().oracle_update_grid(self.h, m.ctypes.data, x, y, yaw, dt)

    def update_measurement_grid(self, meas):
        m = np.ascontiguousarray(meas, dtype=MEAS_CELL_DTYPE)
        lib().oracle_update_measurement_grid(self.h, m.ctypes.data)

    def update_pose(self, x, y, yaw):
        lib().oracle_update_pose(self.h, x, y, yaw)

    def initialize_particles(self):
        lib().oracle_initialize_particles(self.h)

    def particle_prediction(self, dt):
        lib().oracle_particle_prediction(self.h, dt)

    def particle_assignment(self):
        lib().oracle_particle_assignment(self.h)

    def grid_cell_occupancy_update(self, dt):
        lib().oracle_grid_cell_occupancy_update(self.h, dt)

    def update_persistent_particles(self):
        lib().oracle_update_persistent_particles(self.h)

    def initialize_new_particles(self):
        lib().oracle_initialize_new_particles(self.h)

    def statistical_moments(self):
        lib().oracle_statistical_moments(self.h)

    def resampling(self):
        lib().oracle_resampling(self.h)

    # ---- state transport -----------------------------------------------------------------------
    def _getf(self, field, n):
        out = np.empty(n, np.float64)
        assert lib().oracle_get(self.h, field, out.ctypes.data) == n
        return out

    def _setf(self, field, arr, n):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(n)
        assert lib().oracle_set(self.h, field, a.ctypes.data) == n

    def _geti(self, field, n):
        out = np.empty(n, np.int32)
        assert lib().oracle_get_i32(self.h, field, out.ctypes.data) == n
        return out

    def _seti(self, field, arr, n):
        a = np.ascontiguousarray(arr, dtype=np.int32).reshape(n)
        assert lib().oracle_set_i32(self.h, field, a.ctypes.data) == n

    def _dt(self):
        return np.float64 if self.fp64 else np.float32

    def _get_set(self, base, iidx, iassoc, n):
        st = np.stack([self._getf(base + k, n) for k in range(4)], axis=1).astype(self._dt())
        return {"state": st, "grid_cell_idx": self._geti(iidx, n),
                "weight": self._getf(base + 4, n).astype(self._dt()),
                "associated": self._geti(iassoc, n).astype(np.uint8)}

    def get_particles(self):
        return self._get_set(F_P_X, I_P_IDX, I_P_ASSOC, self.particle_count)

    def get_birth_particles(self):
        return self._get_set(F_B_X, I_B_IDX, I_B_ASSOC, self.new_born_particle_count)

    def _put_set(self, base, iidx, iassoc, n, state, idx, weight, assoc):
        if state is not None:
            s = np.asarray(state).reshape(n, 4)
            for k in range(4):
                self._setf(base + k, s[:, k], n)
        if idx is not None:
            self._seti(iidx, idx, n)
        if weight is not None:
            self._setf(base + 4, weight, n)
        if assoc is not None:
            self._seti(iassoc, np.asarray(assoc).astype(np.int32), n)

    def set_particles(self, state=None, grid_cell_idx=None, weight=None, associated=None):
        self._put_set(F_P_X, I_P_IDX, I_P_ASSOC, self.particle_count, state, grid_cell_idx, weight, associated)

    def set_birth_particles(self, state=None, grid_cell_idx=None, weight=None, associated=None):
        self._put_set(F_B_X, I_B_IDX, I_B_ASSOC, self.new_born_particle_count, state, grid_cell_idx, weight,
                      associated)

    def get_grid_cells(self):
        """Grid fields as a dict of arrays (float64 values exact for the f32 mode)."""
        C_ = self.grid_cell_count
        out = {"start_idx": self._geti(I_G_START, C_), "end_idx": self._geti(I_G_END, C_)}
        for k, name in enumerate(GRID_FIELDS):
            out[name] = self._getf(F_GRID0 + k, C_).astype(self._dt())
        return out

    def set_grid_cells(self, cells):
        C_ = self.grid_cell_count
        self._seti(I_G_START, cells["start_idx"], C_)
        self._seti(I_G_END, cells["end_idx"], C_)
        for k, name in enumerate(GRID_FIELDS):
            self._setf(F_GRID0 + k, cells[name], C_)

    def get_weight_array(self):
        return self._getf(F_WEIGHT_ARRAY, self.particle_count).astype(self._dt())

    def set_weight_array(self, w):
        self._setf(F_WEIGHT_ARRAY, w, self.particle_count)

    def get_born_masses(self):
        return self._getf(F_BORN, self.grid_cell_count).astype(self._dt())

    def set_born_masses(self, m):
        self._setf(F_BORN, m, self.grid_cell_count)

    def get_resampled_indices(self):
        return self._geti(I_ANCESTORS, self.particle_count)

    @property
    def epoch(self) -> int:
        return lib().oracle_get_epoch(self.h)

    @epoch.setter
    def epoch(self, e: int) -> None:
        lib().oracle_set_epoch(self.h, e)

    def set_flags(self, first_pose: bool, first_meas: bool) -> None:
        lib().oracle_set_flags(self.h, int(first_pose), int(first_meas))

    def set_draw_mode(self, mode: str) -> None:
        lib().oracle_set_draw_mode(self.h, {"multinomial": 0, "systematic": 1}[mode])

    def set_ego_mode(self, mode: str) -> None:
        lib().oracle_set_ego_mode(self.h, {"reference": 0, "metric": 1}[mode])

    def set_cdf_mode(self, mode: str) -> None:
        """Resampling CDF arithmetic of the f32 checker: 'exact', 'fp64' or 'fp32'."""
        lib().oracle_set_cdf_mode(self.h, {"exact": 0, "fp64": 1, "fp32": 2, "fp32c": 3}[mode])

    def position(self):
        return (lib().oracle_position_x(self.h), lib().oracle_position_y(self.h), lib().oracle_yaw(self.h))


# ---- numerics helpers for known-answer tests -------------------------------------------------------
def threads() -> int:
    """OpenMP threads the checker's parallel loops use (OMP_NUM_THREADS, else the host's cores)."""
    return int(lib().oracle_threads())


def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().oracle_philox(*ctr, *key, out)
    return tuple(out)


def log_f32(x: float) -> float:
    return lib().oracle_log_f32(x)


def sincos_2pi(t: float):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincos_2pi(t, C.byref(s), C.byref(c))
    return s.value, c.value


def box_muller(a: int, b: int):
    z0, z1 = C.c_float(), C.c_float()
    lib().oracle_box_muller(a, b, C.byref(z0), C.byref(z1))
    return z0.value, z1.value


def uniform01(x: int) -> float:
    return lib().oracle_uniform01(x)


def uniform01_oc(x: int) -> float:
    return lib().oracle_uniform01_oc(x)


def _to_int(hi, lo) -> int:
    return (hi.value << 64) | lo.value


def _split(v: int):
    u = v & ((1 << 128) - 1)
    hi = u >> 64
    if hi >= 1 << 63:
        hi -= 1 << 64
    return C.c_int64(hi), C.c_uint64(u & ((1 << 64) - 1))


def fx_from_f32(v: float) -> int:
    hi, lo = C.c_int64(), C.c_uint64()
    lib().oracle_fx_from_f32(v, C.byref(hi), C.byref(lo))
    return _to_int(hi, lo)


def fx_from_f32_nonneg_mismatches(bits) -> int:
    """Count of the float bit patterns in `bits` (uint32 array) where the kernels' sign-free
    conversion differs from fx_from_f32."""
    import numpy as np
    b = np.ascontiguousarray(bits, dtype=np.uint32)
    return lib().oracle_fx_from_f32_nonneg_mismatches(b.ctypes.data, len(b))


def fx_from_f64_nonneg_mismatches(bits) -> int:
    """Count of the double bit patterns in `bits` (uint64 array) where the sign-free conversion
    differs from fx_from_f64."""
    import numpy as np
    b = np.ascontiguousarray(bits, dtype=np.uint64)
    return lib().oracle_fx_from_f64_nonneg_mismatches(b.ctypes.data, len(b))


def exponential_sign_violations() -> int:
    return lib().oracle_exponential_sign_violations()


def fx_from_f64(v: float) -> int:
    hi, lo = C.c_int64(), C.c_uint64()
    lib().oracle_fx_from_f64(v, C.byref(hi), C.byref(lo))
    return _to_int(hi, lo)


def fx_to_f32(v: int) -> float:
    return lib().oracle_fx_to_f32(*_split(v))


def fx_to_f64(v: int) -> float:
    return lib().oracle_fx_to_f64(*_split(v))


def fx_scale_u64(total: int, u: int) -> int:
    hi, lo = C.c_int64(), C.c_uint64()
    lib().oracle_fx_scale_u64(*_split(total), u, C.byref(hi), C.byref(lo))
    return _to_int(hi, lo)


def order_end(prefix: int, total: int, v: int) -> int:
    return lib().oracle_order_end(*_split(prefix), *_split(total), v)


def spacing_fractions(seed: int, instance: int, epoch: int, n: int) -> np.ndarray:
    """The n draw fractions u_k in [0, 1) of one resampling (shared by every CDF mode)."""
    out = np.empty(n, np.float64)
    lib().oracle_spacing_fractions(seed, instance, epoch, n, out.ctypes.data)
    return out


def occupancy(occ, free_mass) -> tuple[np.ndarray, np.ndarray]:
    """Pignistic occupancy and 8-bit image of the cells (image_creation.cpp:19-22, 123-125)."""
    o = np.ascontiguousarray(occ, np.float32)
    fr = np.ascontiguousarray(free_mass, np.float32)
    p = np.empty(o.size, np.float32)
    img = np.empty(o.size, np.uint8)
    lib().oracle_occupancy(o.ctypes.data, fr.ctypes.data, o.size, p.ctypes.data, img.ctypes.data)
    return p, img


def cells_with_velocity(grid_size: int, cells, min_occupancy: float, min_velocity: float) -> np.ndarray:
    """Row-major indices of the cells computeCellsWithVelocity keeps (image_creation.cpp:24-66);
    `cells` is a structured array with the dogm_grid_cell field names."""
    fields = ["occ_mass", "free_mass", "mean_x_vel", "mean_y_vel", "var_x_vel", "var_y_vel", "covar_xy_vel"]
    arrs = [np.ascontiguousarray(cells[k], np.float32) for k in fields]
    out = np.empty(max(arrs[0].size, 1), np.int32)
    n = lib().oracle_cells_with_velocity(grid_size, *[a.ctypes.data for a in arrs], min_occupancy, min_velocity,
                                         out.ctypes.data)
    return out[:n].copy()


def velocity_normalized(mean_x: float, mean_y: float, var_x: float, var_y: float, covar: float) -> float:
    """v^T S^-1 v of one cell as the demo evaluates it (image_creation.cpp:35-50)."""
    return lib().oracle_velocity_normalized(mean_x, mean_y, var_x, var_y, covar)


def lidar_grid(grid_size: int, resolution: float, max_range: float, occ_mass: float, free_mass: float,
               beam_cos, beam_sin, ranges) -> np.ndarray:
    """Measurement grid of one scan (structured array with the dogm_measurement_cell fields)."""
    cs = np.ascontiguousarray(beam_cos, np.float64)
    sn = np.ascontiguousarray(beam_sin, np.float64)
    r = np.ascontiguousarray(ranges, np.float64)
    out = np.zeros(grid_size * grid_size, dtype=[("free_mass", "<f4"), ("occ_mass", "<f4"), ("likelihood", "<f4"),
                                                 ("p_A", "<f4")])
    lib().oracle_lidar_grid(grid_size, cs.size, resolution, max_range, occ_mass, free_mass, cs.ctypes.data,
                            sn.ctypes.data, r.ctypes.data, out.ctypes.data)
    return out


def moment_sums(keys, x) -> np.ndarray:
    """Per-cell sums of the f32 path's velocity-moment channels in the device's f64 association.
    keys: sorted cell index per particle; x: (5, n) float32 channel values.  Returns (5, n) float64
    with the cell's sum at the cell's last particle and NaN elsewhere."""
    k = np.ascontiguousarray(keys, np.int32)
    v = np.ascontiguousarray(x, np.float32).reshape(5, k.size)
    out = np.full((5, k.size), np.nan)
    lib().oracle_moment_sums(k.ctypes.data, k.size, v.ctypes.data, out.ctypes.data)
    return out


def two_sum_f32(a: float, b: float):
    s, e = C.c_float(), C.c_float()
    lib().oracle_two_sum_f32(a, b, C.byref(s), C.byref(e))
    return np.float32(s.value), np.float32(e.value)


def f32c_from_double(x: float):
    hi = np.float32(x)
    return hi, np.float32(x - float(hi))


def f32c_add(a, b):
    hi, lo = C.c_float(), C.c_float()
    lib().oracle_f32c_add(a[0], a[1], b[0], b[1], C.byref(hi), C.byref(lo))
    return np.float32(hi.value), np.float32(lo.value)
