"""GPU parity: the gfx950 path against the f32 checker, bit for bit, through the C-ABI.

Bit-exact on everything: particle cell indices, sorted order, cell start/end bounds, resampled
ancestor indices, and (because every reduction is an exact fixed-point sum and every element
formula is the same f32 operation sequence on both sides) every mass, weight and moment as well.
"""
import sys

import numpy as np
import pytest

from dogm_ctypes import DOGM, DeviceBuffer, Params
from oracle_ctypes import OracleDOGM
from parity_util import (assert_grid_equal, assert_int_equal, assert_particles_equal, assert_state_equal,
                         small_params, small_scene)

pytestmark = pytest.mark.gpu

STAGES = [
    ("particle_prediction", (0.1,)),
    ("particle_assignment", ()),
    ("grid_cell_occupancy_update", (0.1,)),
    ("update_persistent_particles", ()),
    ("initialize_new_particles", ()),
    ("statistical_moments", ()),
    ("resampling", ()),
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_stagewise_bit_exact():
    """Every public stage of the reference (dogm.h:148-156), frame by frame, with ego motion."""
    p = small_params()
    sc = small_scene()
    gpu = DOGM(p, seed=11)
    orc = OracleDOGM(p, seed=11)
    for f in range(4):
        _say(f"stagewise frame {f}")
        m = sc.measurement(f)
        gpu.update_measurement_grid(m)
        orc.update_measurement_grid(m)
        assert_state_equal(f"frame {f} measurement", gpu, orc)
        x, y, yaw = sc.ego_pose(f)
        gpu.update_pose(x, y, yaw)
        orc.update_pose(x, y, yaw)
        assert gpu.position() == orc.position()
        assert_state_equal(f"frame {f} pose", gpu, orc)
        for name, args in STAGES:
            getattr(gpu, name)(*args)
            getattr(orc, name)(*args)
            assert_state_equal(f"frame {f} {name}", gpu, orc, ancestors=(name == "resampling"))


def test_update_grid_bit_exact_with_ego_shift():
    """The fused per-frame path (shift applied inside predict / occupancy) against the checker's
    eager stage sequence."""
    p = small_params(gs=64, res=0.25, n=20000, b=2000)
    sc = small_scene(gs=64, res=0.25, n_boxes=6, n_beams=240, ego_speed=5.0, yaw_rate=0.2, seed=5)
    gpu = DOGM(p, seed=3)
    orc = OracleDOGM(p, seed=3)
    dm = DeviceBuffer(gpu.grid_cell_count * 16)
    for f in range(8):
        _say(f"update_grid frame {f}")
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        if f % 2:
            dm.upload(m)
            gpu.update_grid(dm, x, y, yaw, 0.1)
        else:
            gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert gpu.position() == orc.position()
        assert_state_equal(f"frame {f}", gpu, orc, ancestors=True)
    dm.free()


def test_plain_small_workload_bit_exact():
    """The 250x250 / 300k / 30k workload with its scene, several frames."""
    import scene
    w = scene.WORKLOADS["plain_small"]
    p = Params(**w["params"])
    sc = scene.make_scene("plain_small")
    gpu = DOGM(p)
    orc = OracleDOGM(p)
    for f in range(4):
        _say(f"plain_small frame {f}")
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        gpu.update_grid(m, x, y, yaw, w["dt"])
        orc.update_grid(m, x, y, yaw, w["dt"])
        assert_state_equal(f"frame {f}", gpu, orc, ancestors=True)


def test_single_cell_spanning_many_tiles():
    """All evidence in one cell: after init every particle sits in that cell, so the per-cell
    sums cross every reduction tile (exercises the cross-tile fix-up)."""
    p = small_params(gs=16, res=1.0, n=50000, b=5000, stddev_process_noise_position=0.0,
                     stddev_process_noise_velocity=0.0, init_max_velocity=0.0)
    m = np.zeros(256, dtype=[("free_mass", "<f4"), ("occ_mass", "<f4"), ("likelihood", "<f4"), ("p_A", "<f4")])
    m["likelihood"] = 1.0
    m["p_A"] = 1.0
    m["occ_mass"][8 + 16 * 8] = 0.9
    gpu = DOGM(p, seed=9)
    orc = OracleDOGM(p, seed=9)
    for f in range(3):
        _say(f"single cell frame {f}")
        gpu.update_grid(m, 0.0, 0.0, 0.0, 0.1)
        orc.update_grid(m, 0.0, 0.0, 0.0, 0.1)
        assert_state_equal(f"frame {f}", gpu, orc, ancestors=True)
    cells = gpu.get_grid_cells()
    c = 8 + 16 * 8
    assert cells["end_idx"][c] - cells["start_idx"][c] + 1 > p.particle_count // 2


def test_zero_measurement_and_no_births():
    """All-zero evidence (uniform initial spread) and B = 0."""
    p = small_params(gs=32, res=0.5, n=5000, b=0)
    m = np.zeros(32 * 32, dtype=[("free_mass", "<f4"), ("occ_mass", "<f4"), ("likelihood", "<f4"), ("p_A", "<f4")])
    m["likelihood"] = 1.0
    m["p_A"] = 1.0
    gpu = DOGM(p)
    orc = OracleDOGM(p)
    for f in range(3):
        _say(f"zero measurement frame {f}")
        gpu.update_grid(m, 0.1 * f, 0.0, 0.0, 0.1)
        orc.update_grid(m, 0.1 * f, 0.0, 0.0, 0.1)
        assert_state_equal(f"frame {f}", gpu, orc, ancestors=True, births=False)


@pytest.mark.parametrize("gs,n,b", [(1, 1, 1), (10, 2, 1), (3, 7, 0), (5, 513, 3)])
def test_tiny_sizes(gs, n, b):
    p = small_params(gs=gs, res=1.0, n=n, b=b)
    sc = small_scene(gs=max(gs, 2), res=1.0, n_boxes=1, n_beams=16, ego_speed=1.5, seed=gs)
    gpu = DOGM(p)
    orc = OracleDOGM(p)
    for f in range(3):
        m = sc.measurement(f)[: gs * gs] if gs > 1 else np.array([(0.1, 0.6, 1.0, 1.0)], dtype=sc.measurement(0).dtype)
        x, y, yaw = sc.ego_pose(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"gs={gs} n={n} b={b} frame {f}", gpu, orc, ancestors=True, births=b > 0)


def test_sort_stability_and_bounds_against_numpy():
    """Sorted order equals numpy's stable argsort of the predicted cell indices, and start/end are
    the first/last occurrences (an independent formulation of the assignment stage)."""
    p = small_params(gs=200, res=0.5, n=300000, b=10)
    gpu = DOGM(p, seed=21)
    rng = np.random.default_rng(0)
    st = np.zeros((p.particle_count, 4), np.float32)
    st[:, 0] = rng.uniform(0, 200, p.particle_count)
    st[:, 1] = rng.uniform(0, 200, p.particle_count)
    st[:, 2] = np.arange(p.particle_count, dtype=np.float32)  # tags the original slot
    idx = (np.clip(st[:, 0].astype(np.int64), 0, 199) + 200 * np.clip(st[:, 1].astype(np.int64), 0, 199))
    gpu.set_particles(st, idx.astype(np.int32), np.full(p.particle_count, 1e-3, np.float32),
                      np.zeros(p.particle_count, np.uint8))
    gpu.particle_assignment()
    g = gpu.get_particles()
    order = np.argsort(idx, kind="stable")
    assert_int_equal("sorted cell idx", g["grid_cell_idx"], idx[order])
    assert_int_equal("sorted slots", g["state"][:, 2].astype(np.int64), order)
    cells = gpu.get_grid_cells()
    uniq, first = np.unique(idx[order], return_index=True)
    last = np.r_[first[1:], p.particle_count] - 1
    start = np.full(40000, -1)
    end = np.full(40000, -1)
    start[uniq] = first
    end[uniq] = last
    assert_int_equal("start", cells["start_idx"], start)
    assert_int_equal("end", cells["end_idx"], end)


def test_run_twice_identical():
    p = small_params()
    sc = small_scene()
    outs = []
    for _ in range(2):
        gpu = DOGM(p, seed=5)
        for f in range(3):
            x, y, yaw = sc.ego_pose(f)
            gpu.update_grid(sc.measurement(f), x, y, yaw, 0.1)
        outs.append((gpu.get_particles(), gpu.get_grid_cells()))
        gpu.close()
    assert_particles_equal("rerun", outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1].view(np.uint8), outs[1][1].view(np.uint8))


def test_errors_are_reported():
    p = small_params()
    from dogm_ctypes import DogmError, lib
    gpu = DOGM(p)
    assert lib.dogm_update_grid(gpu.h, None, 0.0, 0.0, 0.0, 0.1, 0) == 1  # DOGM_ERR_INVALID_ARGUMENT
    assert b"null" in lib.dogm_last_error(gpu.h)
    with pytest.raises(DogmError):
        DOGM(small_params(n=0))
    bad = np.zeros(p.particle_count, np.int32)
    bad[7] = gpu.grid_cell_count
    with pytest.raises(DogmError):
        gpu.set_particles(grid_cell_idx=bad)
    # The handle stays usable after a rejected call.
    gpu.particle_prediction(0.1)


def test_full_updates_mixed_with_grid_writers():
    """Full updates skip re-zeroing cells that were empty in the previous full update; every other
    writer of the grid must make the next full update rewrite all cells.  Mixes full updates with
    grid-cell injection (stale non-zero moments, mu and bounds in empty cells), a stage-by-stage
    frame, an eager pose update and the metric ego mode, comparing every field after each call."""
    p = small_params(gs=64, res=0.25, n=20000, b=2000)
    sc = small_scene(gs=64, res=0.25, n_boxes=6, n_beams=240, ego_speed=5.0, yaw_rate=0.2, seed=9)
    gpu = DOGM(p, seed=21)
    orc = OracleDOGM(p, seed=21)
    rng = np.random.default_rng(4)

    def full(f):
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"full frame {f}", gpu, orc, ancestors=True)

    f = 0
    for _ in range(3):
        full(f)
        f += 1
    _say("inject stale fields")
    cells = gpu.get_grid_cells()
    for name in ("mu_A", "mu_UA", "mean_x_vel", "mean_y_vel", "var_x_vel", "var_y_vel", "covar_xy_vel"):
        cells[name] = rng.uniform(0.5, 2.0, cells.shape).astype(np.float32)
    cells["start_idx"] = rng.integers(0, p.particle_count, cells.shape)
    cells["end_idx"] = cells["start_idx"]
    gpu.set_grid_cells(cells)
    orc.set_grid_cells({k: cells[k] for k in cells.dtype.names})
    for _ in range(2):
        full(f)
        f += 1
    _say("stage-by-stage frame")
    m = sc.measurement(f)
    x, y, yaw = sc.ego_pose(f)
    for o in (gpu, orc):
        o.update_measurement_grid(m)
        o.update_pose(x, y, yaw)
        for name, args in STAGES[:2]:
            getattr(o, name)(*args)
    assert_state_equal(f"stages frame {f}", gpu, orc)
    for o in (gpu, orc):
        for name, args in STAGES[2:]:
            getattr(o, name)(*args)
    assert_state_equal(f"stages frame {f} end", gpu, orc, ancestors=True)
    f += 1
    full(f)
    f += 1
    _say("eager pose update")
    x, y, yaw = sc.ego_pose(f)
    gpu.update_pose(x + 1.0, y - 0.75, yaw)
    orc.update_pose(x + 1.0, y - 0.75, yaw)
    assert_state_equal("eager pose", gpu, orc)
    for _ in range(2):
        full(f)
        f += 1
    _say("metric ego mode")
    gpu.set_ego_mode("metric")
    orc.set_ego_mode("metric")
    for _ in range(3):
        full(f)
        f += 1


def test_full_updates_after_setting_unequal_weights():
    """Full updates normally start from equal weights (the last resampling's, or the initial 1 / N),
    which the prediction then leaves in place for the sorting pass to form the predicted weights;
    weights written through the C-ABI between updates are unequal and must be read per particle.
    Both paths, and the switch back after the next resampling, match the checker."""
    p = small_params(gs=60, res=0.5, n=20000, b=2000)
    sc = small_scene(gs=60, res=0.5)
    gpu = DOGM(p, seed=5)
    orc = OracleDOGM(p, seed=5)
    rng = np.random.default_rng(31)
    for f in range(6):
        _say(f"unequal weights frame {f}")
        if f in (2, 4):
            w = (rng.uniform(0.0, 2.0, p.particle_count) / p.particle_count).astype(np.float32)
            gpu.set_particles(weight=w)
            orc.set_particles(weight=w)
        x, y, yaw = sc.ego_pose(f)
        m = sc.measurement(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"frame {f}", gpu, orc, ancestors=True, births=True)


def test_deferred_resampled_sets():
    """A resampling leaves the new particle set as ancestor indices, which the next full update's
    prediction reads the states through; every other reader or writer of the particles first writes
    the set out.  Full updates run back to back without particle read-backs (the prediction path),
    with a birth read-back, an eager pose update, a stage-by-stage frame, a particle write and a
    particle read-back between them; every array matches the checker."""
    p = small_params(gs=64, res=0.25, n=20000, b=2000)
    sc = small_scene(gs=64, res=0.25, n_boxes=6, n_beams=240, ego_speed=5.0, yaw_rate=0.2, seed=13)
    gpu = DOGM(p, seed=17)
    orc = OracleDOGM(p, seed=17)
    rng = np.random.default_rng(8)
    f = 0

    def full(check_particles=False):
        nonlocal f
        _say(f"deferred sets frame {f}")
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"frame {f}", gpu, orc, ancestors=True, particles=check_particles)
        f += 1

    for _ in range(4):
        full()
    full(check_particles=True)
    full()
    _say("birth read-back")
    assert_particles_equal("births", gpu.get_birth_particles(), orc.get_birth_particles())
    full()
    full()
    _say("eager pose update")
    x, y, yaw = sc.ego_pose(f)
    gpu.update_pose(x + 1.0, y - 0.75, yaw)
    orc.update_pose(x + 1.0, y - 0.75, yaw)
    full()
    _say("stage-by-stage frame")
    m = sc.measurement(f)
    x, y, yaw = sc.ego_pose(f)
    for o in (gpu, orc):
        o.update_measurement_grid(m)
        o.update_pose(x, y, yaw)
        for name, args in STAGES:
            getattr(o, name)(*args)
    assert_state_equal(f"stages frame {f}", gpu, orc, ancestors=True, particles=False)
    f += 1
    full()
    _say("particle write")
    w = (rng.uniform(0.0, 2.0, p.particle_count) / p.particle_count).astype(np.float32)
    gpu.set_particles(weight=w)
    orc.set_particles(weight=w)
    full()
    full()
    _say("metric ego mode")
    gpu.set_ego_mode("metric")
    orc.set_ego_mode("metric")
    full()
    full(check_particles=True)


def test_hostile_weights_in_resampling():
    """Resampling from injected weights with negative values, NaN, zeros, subnormals and a wide
    dynamic range (persistent and birth weights): negative and NaN weights count as zero in the CDF on
    both sides, so ancestors and particles stay bit-exact."""
    p = small_params(gs=40, res=0.5, n=12000, b=1500)
    sc = small_scene(gs=40, res=0.5)
    gpu = DOGM(p, seed=17)
    orc = OracleDOGM(p, seed=17)
    rng = np.random.default_rng(8)
    for f in range(2):
        x, y, yaw = sc.ego_pose(f)
        m = sc.measurement(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)

    def hostile(n):
        w = (10.0 ** rng.uniform(-12, 2, n)).astype(np.float32)
        k = rng.integers(0, 6, n)
        w[k == 0] = -w[k == 0]
        w[k == 1] = np.nan
        w[k == 2] = 0.0
        w[k == 3] = np.float32(1e-40)  # subnormal
        return w

    for rep in range(2):
        _say(f"hostile weights round {rep}")
        for o in (gpu, orc):
            o.particle_prediction(0.1)
            o.particle_assignment()
        w, bw = hostile(p.particle_count), hostile(p.new_born_particle_count)
        gpu.set_weight_array(w)
        orc.set_weight_array(w)
        gpu.set_birth_particles(weight=bw)
        orc.set_birth_particles(weight=bw)
        gpu.resampling()
        orc.resampling()
        assert_int_equal("ancestors", gpu.get_resampled_indices(), orc.get_resampled_indices())
        assert_particles_equal(f"round {rep}", gpu.get_particles(), orc.get_particles())
