"""Resampling CDF modes on the GPU against the checker, bit for bit (exact / FP64 / FP32 /
compensated FP32 scans)."""
import sys

import numpy as np
import pytest

from dogm_ctypes import DOGM, Batch
from oracle_ctypes import OracleDOGM
from parity_util import (assert_grid_equal, assert_int_equal, assert_particles_equal, assert_state_equal,
                         small_params, small_scene)

pytestmark = pytest.mark.gpu

SEED = 44


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32c"])
def test_float_modes_full_frames(mode):
    p = small_params(gs=48, res=0.5, n=30000, b=3000)
    sc = small_scene(gs=48, res=0.5)
    gpu = DOGM(p, seed=SEED, instance=2)
    orc = OracleDOGM(p, seed=SEED, instance=2)
    gpu.set_cdf_mode(mode)
    orc.set_cdf_mode(mode)
    for f in range(3):
        _say(f"{mode} frame {f}")
        x, y, yaw = sc.ego_pose(f)
        m = sc.measurement(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"{mode} frame {f}", gpu, orc, ancestors=True)


@pytest.mark.parametrize("mode", ["exact", "fp64", "fp32", "fp32c"])
def test_log_uniform_weights_many_tiles(mode):
    """Weights 10^(-9 U) injected before resampling (the precision workload's weights)."""
    n, b = 150_000, 15_000
    p = small_params(gs=64, res=0.5, n=n, b=b)
    sc = small_scene(gs=64, res=0.5)
    gpu = DOGM(p, seed=SEED, instance=7)
    orc = OracleDOGM(p, seed=SEED, instance=7)
    gpu.set_cdf_mode(mode)
    orc.set_cdf_mode(mode)
    rng = np.random.default_rng(9)
    for f in range(2):
        _say(f"{mode} log-uniform frame {f}")
        m = sc.measurement(f)
        x, y, yaw = sc.ego_pose(f)
        for o in (gpu, orc):
            o.update_measurement_grid(m)
            o.update_pose(x, y, yaw)
            o.particle_prediction(0.1)
            o.particle_assignment()
            o.grid_cell_occupancy_update(0.1)
            o.update_persistent_particles()
            o.initialize_new_particles()
            o.statistical_moments()
        w = (10.0 ** (-9.0 * rng.random(n))).astype(np.float32)
        bw = (10.0 ** (-9.0 * rng.random(b))).astype(np.float32)
        for o in (gpu, orc):
            o.set_weight_array(w)
            o.set_birth_particles(weight=bw)
            o.resampling()
        assert_int_equal(f"{mode} ancestors", gpu.get_resampled_indices(), orc.get_resampled_indices())
        assert_particles_equal(f"{mode} particles", gpu.get_particles(), orc.get_particles())


def test_batch_mixes_modes():
    vs = [(small_params(gs=40, res=0.5, n=8000, b=800), small_scene(gs=40, res=0.5, seed=50 + k)) for k in range(4)]
    modes = ["exact", "fp64", "fp32", "fp32c"]
    with Batch() as batch:
        ids = [batch.add_vehicle(p, SEED, 60 + k) for k, (p, _) in enumerate(vs)]
        for k, vid in enumerate(ids):
            batch.set_cdf_mode(vid, modes[k])
        plains = []
        for k, (p, _) in enumerate(vs):
            g = DOGM(p, seed=SEED, instance=60 + k)
            g.set_cdf_mode(modes[k])
            plains.append(g)
        for f in range(3):
            _say(f"mixed-mode batch frame {f}")
            meas = [sc.measurement(f) for _, sc in vs]
            poses = [sc.ego_pose(f) for _, sc in vs]
            batch.update(ids, meas, poses, 0.1)
            for k, g in enumerate(plains):
                x, y, yaw = poses[k]
                g.update_grid(meas[k], x, y, yaw, 0.1)
                assert_particles_equal(f"frame {f} vehicle {k}", batch.get_particles(ids[k]), g.get_particles())
                assert_grid_equal(f"frame {f} vehicle {k}", batch.get_grid_cells(ids[k]), g.get_grid_cells())
                assert_int_equal(f"frame {f} vehicle {k} ancestors", batch.get_resampled_indices(ids[k]),
                                 g.get_resampled_indices())


@pytest.mark.parametrize("mode", ["exact", "fp64", "fp32"])
def test_systematic_full_frames(mode):
    p = small_params(gs=48, res=0.5, n=30000, b=3000)
    sc = small_scene(gs=48, res=0.5)
    gpu = DOGM(p, seed=SEED, instance=5)
    orc = OracleDOGM(p, seed=SEED, instance=5)
    for o in (gpu, orc):
        o.set_cdf_mode(mode)
    gpu.set_resampling_mode("systematic")
    orc.set_draw_mode("systematic")
    for f in range(3):
        _say(f"systematic {mode} frame {f}")
        x, y, yaw = sc.ego_pose(f)
        m = sc.measurement(f)
        gpu.update_grid(m, x, y, yaw, 0.1)
        orc.update_grid(m, x, y, yaw, 0.1)
        assert_state_equal(f"systematic {mode} frame {f}", gpu, orc, ancestors=True)


def test_batch_mixes_draw_modes():
    vs = [(small_params(gs=40, res=0.5, n=8000, b=800), small_scene(gs=40, res=0.5, seed=80 + k)) for k in range(2)]
    draws = ["systematic", "multinomial"]
    with Batch() as batch:
        ids = [batch.add_vehicle(p, SEED, 90 + k) for k, (p, _) in enumerate(vs)]
        plains = []
        for k, (p, _) in enumerate(vs):
            batch.set_resampling_mode(ids[k], draws[k])
            g = DOGM(p, seed=SEED, instance=90 + k)
            g.set_resampling_mode(draws[k])
            plains.append(g)
        for f in range(3):
            _say(f"mixed draw modes frame {f}")
            meas = [sc.measurement(f) for _, sc in vs]
            poses = [sc.ego_pose(f) for _, sc in vs]
            batch.update(ids, meas, poses, 0.1)
            for k, g in enumerate(plains):
                g.update_grid(meas[k], *poses[k], 0.1)
                assert_particles_equal(f"frame {f} vehicle {k}", batch.get_particles(ids[k]), g.get_particles())
                assert_int_equal(f"frame {f} vehicle {k} ancestors", batch.get_resampled_indices(ids[k]),
                                 g.get_resampled_indices())
