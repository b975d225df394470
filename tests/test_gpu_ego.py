"""Metric ego compensation on the GPU against the checker, bit for bit: full frames with a moving,
turning ego, the pose update alone (every grid field), and a batch mixing both modes."""
import sys

import numpy as np
import pytest

from dogm_ctypes import DOGM, Batch
from oracle_ctypes import OracleDOGM
from parity_util import assert_grid_equal, assert_particles_equal, assert_state_equal, small_params, small_scene

pytestmark = pytest.mark.gpu
SEED = 17


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene(gs=48, seed=5):
    sc = small_scene(gs=gs, res=0.5, ego_speed=4.0, yaw_rate=0.3, seed=seed)
    sc.cfg.sensor_frame = "vehicle"
    return sc


def test_metric_frames_match_checker():
    p = small_params(gs=48, res=0.5, n=20000, b=2000)
    sc = _scene()
    gpu, orc = DOGM(p, seed=SEED, instance=3), OracleDOGM(p, seed=SEED, instance=3)
    for o in (gpu, orc):
        o.set_ego_mode("metric")
    for f in range(5):
        _say(f"metric frame {f}")
        x, y, yaw = sc.ego_pose(f)
        m = sc.measurement(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"metric frame {f}", gpu, orc)
        assert gpu.position() == pytest.approx(orc.position())


def test_metric_pose_update_alone():
    p = small_params(gs=40, res=0.5, n=6000, b=600)
    sc = _scene(gs=40, seed=9)
    gpu, orc = DOGM(p, seed=SEED, instance=4), OracleDOGM(p, seed=SEED, instance=4)
    for f in range(2):
        x, y, yaw = sc.ego_pose(f)
        gpu.update_grid(sc.measurement(f), x, y, yaw, 0.1)
        orc.update_grid(sc.measurement(f), x, y, yaw, 0.1)
    for o in (gpu, orc):
        o.set_ego_mode("metric")
    for k, (x, y, yaw) in enumerate([(0.3, 0.2, 0.5), (1.4, -0.9, 0.5), (1.4, -0.9, -0.25), (2.9, 0.6, 0.1)]):
        _say(f"metric pose update {k}")
        gpu.update_pose(x, y, yaw)
        orc.update_pose(x, y, yaw)
        assert_particles_equal(f"pose {k}", gpu.get_particles(), orc.get_particles())
        assert_grid_equal(f"pose {k}", gpu.get_grid_cells(), orc.get_grid_cells())


def test_batch_mixing_modes_matches_plain():
    vs = [(small_params(gs=40, res=0.5, n=8000, b=800), _scene(gs=40, seed=20 + k)) for k in range(2)]
    modes = ["metric", "reference"]
    with Batch() as b:
        ids = [b.add_vehicle(p, SEED, 50 + k) for k, (p, _) in enumerate(vs)]
        plains = []
        for k, (p, _) in enumerate(vs):
            b.set_ego_mode(ids[k], modes[k])
            g = DOGM(p, seed=SEED, instance=50 + k)
            g.set_ego_mode(modes[k])
            plains.append(g)
        for f in range(4):
            _say(f"mixed ego modes frame {f}")
            meas = [sc.measurement(f) for _, sc in vs]
            poses = [sc.ego_pose(f) for _, sc in vs]
            b.update(ids, meas, poses, 0.1)
            for k, g in enumerate(plains):
                g.update_grid(meas[k], *poses[k], 0.1)
                assert_particles_equal(f"frame {f} vehicle {k}", b.get_particles(ids[k]), g.get_particles())
                assert_grid_equal(f"frame {f} vehicle {k}", b.get_grid_cells(ids[k]), g.get_grid_cells())
