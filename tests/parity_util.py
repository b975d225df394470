"""Helpers shared by the parity tests: scenarios and bitwise comparisons of GPU vs checker state."""
from __future__ import annotations

import numpy as np

import scene as scene_mod
from dogm_ctypes import Params

GRID_FIELDS = ["start_idx", "end_idx", "new_born_occ_mass", "pers_occ_mass", "free_mass", "occ_mass",
               "pred_occ_mass", "mu_A", "mu_UA", "w_A", "w_UA", "mean_x_vel", "mean_y_vel", "var_x_vel",
               "var_y_vel", "covar_xy_vel"]


def small_params(gs=48, res=0.5, n=3000, b=300, **kw) -> Params:
    base = dict(size=gs * res, resolution=res, particle_count=n, new_born_particle_count=b, persistence_prob=0.99,
                stddev_process_noise_position=0.1, stddev_process_noise_velocity=2.0, birth_prob=0.02,
                stddev_velocity=30.0, init_max_velocity=15.0, freespace_discount=0.01)
    base.update(kw)
    p = Params(**base)
    assert int(np.float32(p.size) / np.float32(p.resolution)) == gs
    return p


def small_scene(gs=48, res=0.5, n_boxes=5, n_beams=180, ego_speed=6.0, yaw_rate=0.1, seed=3):
    cfg = scene_mod.SceneConfig(grid_size=gs, resolution=res, n_boxes=n_boxes, n_beams=n_beams,
                                max_range=0.5 * gs * res, ego_speed=ego_speed, ego_yaw_rate=yaw_rate, seed=seed)
    return scene_mod.Scene(cfg)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32)


def _diff_msg(name, a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    bad = np.nonzero(a != b)[0]
    i = int(bad[0])
    return f"{name}: {bad.size} of {a.size} differ; first at {i}: gpu={a[i]!r} checker={b[i]!r}"


def assert_bits_equal(name, gpu, ref):
    g, r = _u32(gpu), _u32(ref)
    assert g.shape == r.shape, f"{name}: shape {g.shape} vs {r.shape}"
    if not np.array_equal(g, r):
        raise AssertionError(_diff_msg(name, np.asarray(gpu, np.float32), np.asarray(ref, np.float32)))


def assert_int_equal(name, gpu, ref):
    g, r = np.asarray(gpu), np.asarray(ref)
    assert g.shape == r.shape, f"{name}: shape {g.shape} vs {r.shape}"
    if not np.array_equal(g, r):
        raise AssertionError(_diff_msg(name, g, r))


def assert_particles_equal(tag, g: dict, o: dict):
    assert_int_equal(f"{tag} grid_cell_idx", g["grid_cell_idx"], o["grid_cell_idx"])
    assert_bits_equal(f"{tag} state", g["state"], o["state"])
    assert_bits_equal(f"{tag} weight", g["weight"], o["weight"])
    assert_int_equal(f"{tag} associated", g["associated"], o["associated"])


def assert_grid_equal(tag, gcells: np.ndarray, ocells: dict, fields=GRID_FIELDS):
    for f in fields:
        if f in ("start_idx", "end_idx"):
            assert_int_equal(f"{tag} {f}", gcells[f], ocells[f])
        else:
            assert_bits_equal(f"{tag} {f}", gcells[f], ocells[f])


def assert_state_equal(tag, gpu, orc, *, births=True, ancestors=False, particles=True):
    """Every observable array of the filter, bit for bit.  particles=False skips the particle and
    birth read-backs (they write out a resampled set that the next update would otherwise read through
    its ancestors)."""
    if particles:
        assert_particles_equal(f"{tag} particles", gpu.get_particles(), orc.get_particles())
    assert_grid_equal(f"{tag} grid", gpu.get_grid_cells(), orc.get_grid_cells())
    assert_bits_equal(f"{tag} weight_array", gpu.get_weight_array(), orc.get_weight_array())
    assert_bits_equal(f"{tag} born masses", gpu.get_born_masses(), orc.get_born_masses())
    if particles and births and gpu.new_born_particle_count > 0:
        assert_particles_equal(f"{tag} births", gpu.get_birth_particles(), orc.get_birth_particles())
    if ancestors:
        assert_int_equal(f"{tag} ancestors", gpu.get_resampled_indices(), orc.get_resampled_indices())


def state_differences(g_parts: dict, o_parts: dict, g_grid, o_grid) -> list[str]:
    """Every particle array and grid field that differs, with its count of differing entries (an
    empty list when the states are bit-identical)."""
    out = []
    pairs = [(f"particles {k}", g_parts[k], o_parts[k]) for k in ("grid_cell_idx", "state", "weight", "associated")]
    pairs += [(f"grid {f}", g_grid[f], o_grid[f]) for f in GRID_FIELDS]
    for name, a, b in pairs:
        a, b = np.asarray(a), np.asarray(b)
        if a.dtype.kind == "f":
            a, b = _u32(a), _u32(b)
        if a.shape != b.shape:
            out.append(f"{name}: shape {a.shape} vs {b.shape}")
            continue
        bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
        if bad.size:
            out.append(f"{name}: {bad.size} of {a.shape[0]} differ (first at {int(bad[0])})")
    return out


def assert_vehicle_equal(tag, g_parts: dict, o_parts: dict, g_grid, o_grid):
    """Particles and grid bit for bit; a failure lists every array that differs, so the earliest
    stage whose output differs can be read off."""
    d = state_differences(g_parts, o_parts, g_grid, o_grid)
    if d:
        raise AssertionError(f"{tag}: " + "; ".join(d))
