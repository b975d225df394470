"""Full-size GPU parity: the gfx950 path against the f32 checker, bit for bit, at the benchmark sizes.

* 1200 x 1200 cells, 2e6 + 2e5 particles (the headline workload): first frame (particle
  initialisation) and two more with a grid shift every frame, every observable array compared.
* 2048 x 2048 cells, 8e6 + 8e5 particles (the largest workload, precision study size): two frames in
  the exact and the FP64 CDF modes.
* A batch of two full-size vehicles (different Philox instances) against two plain handles:
  batching never changes a result.

The checker runs on the host's OpenMP threads (a few hundred ms per full-size frame).
"""
import sys

import numpy as np
import pytest

import scene
from dogm_ctypes import DOGM, Batch, DeviceBuffer, Params
from oracle_ctypes import OracleDOGM
from parity_util import assert_grid_equal, assert_int_equal, assert_particles_equal, assert_state_equal

pytestmark = pytest.mark.gpu

LARGE = dict(size=204.8, resolution=0.1, particle_count=8_000_000, new_born_particle_count=800_000,
             persistence_prob=0.99, stddev_process_noise_position=0.1, stddev_process_noise_velocity=2.0,
             birth_prob=0.02, stddev_velocity=30.0, init_max_velocity=15.0, freespace_discount=0.01)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _full_scene():
    w = scene.WORKLOADS["plain_full"]
    return Params(**w["params"]), scene.make_scene("plain_full", seed=1), float(w["dt"])


def test_headline_size_bit_exact():
    p, sc, dt = _full_scene()
    gpu = DOGM(p, seed=123456)
    orc = OracleDOGM(p, seed=123456)
    dm = DeviceBuffer(gpu.grid_cell_count * 16)
    for f in range(3):
        _say(f"full size frame {f}")
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        dm.upload(m)
        gpu.update_grid(dm, x, y, yaw, dt)
        orc.update_grid(m, x, y, yaw, dt)
        assert gpu.position() == orc.position()
        assert_state_equal(f"1200^2 frame {f}", gpu, orc, ancestors=True)
    dm.free()
    gpu.close()
    orc.close()


@pytest.mark.parametrize("mode", ["exact", "fp64"])
def test_largest_size_bit_exact(mode):
    p = Params(**LARGE)
    cfg = scene.SceneConfig(grid_size=2048, resolution=0.1, n_boxes=60, n_beams=1440, max_range=100.0,
                            ego_speed=10.0, ego_yaw_rate=0.1, seed=1)
    sc = scene.Scene(cfg)
    gpu = DOGM(p, seed=99)
    orc = OracleDOGM(p, seed=99)
    gpu.set_cdf_mode(mode)
    orc.set_cdf_mode(mode)
    for f in range(2):
        _say(f"2048^2 {mode} frame {f}")
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"2048^2 {mode} frame {f}", gpu, orc, ancestors=True)
    gpu.close()
    orc.close()


def test_full_size_batch_matches_plain():
    p, sc, dt = _full_scene()
    b = Batch()
    ids = [b.add_vehicle(p, 123456, inst) for inst in (0, 5)]
    plain = [DOGM(p, seed=123456, instance=inst) for inst in (0, 5)]
    for f in range(3):
        _say(f"full size batch frame {f}")
        m = sc.measurement(f)
        pose = sc.ego_pose(f)
        b.update(ids, [m, m], [pose, pose], dt)
        for g in plain:
            g.update_grid(m, *pose, dt)
    for vid, g in zip(ids, plain):
        assert_particles_equal(f"vehicle {vid}", b.get_particles(vid), g.get_particles())
        gc, bc = g.get_grid_cells(), b.get_grid_cells(vid)
        for name in gc.dtype.names:
            assert np.array_equal(gc[name].view(np.uint32), bc[name].view(np.uint32)), name
        g.close()
    b.close()
