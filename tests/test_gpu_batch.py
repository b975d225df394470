"""Batched form: many grids of uneven size in the same kernel launches (dogm_batch_*).

Every vehicle's result must be bit-identical to a plain handle created with the same params, seed
and instance and fed the same frames, whichever other vehicles share its steps and whenever it
joins.  One vehicle is also checked against the f32 checker directly.
"""
import sys

import numpy as np
import pytest

from dogm_ctypes import DOGM, Batch, DeviceBuffer, vehicle_footprint
from oracle_ctypes import OracleDOGM
from parity_util import (assert_grid_equal, assert_int_equal, assert_particles_equal, small_params,
                         small_scene)

pytestmark = pytest.mark.gpu

SEED = 77

# (grid side, resolution, N, B): uneven sizes, N not a multiple of the tile sizes, one B = 0.
FLEET = [
    (40, 0.5, 3000, 300),
    (64, 0.4, 9001, 901),
    (24, 0.5, 777, 0),
    (96, 0.25, 20000, 2000),
]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vehicle(k):
    gs, res, n, b = FLEET[k]
    return small_params(gs=gs, res=res, n=n, b=b), small_scene(gs=gs, res=res, seed=10 + k)


def _assert_vehicle_equal(tag, batch, vid, plain):
    assert_particles_equal(f"{tag} particles", batch.get_particles(vid), plain.get_particles())
    assert_grid_equal(f"{tag} grid", batch.get_grid_cells(vid), plain.get_grid_cells())
    assert_int_equal(f"{tag} ancestors", batch.get_resampled_indices(vid), plain.get_resampled_indices())


def test_footprint_matches_allocation():
    p, _ = _vehicle(1)
    with Batch() as b:
        b.add_vehicle(p, SEED, 1)
        assert b.device_bytes() == vehicle_footprint(p)


def test_batch_equals_alone_every_step():
    vs = [_vehicle(k) for k in range(len(FLEET))]
    with Batch() as batch:
        ids = [batch.add_vehicle(p, SEED, k) for k, (p, _) in enumerate(vs)]
        plains = [DOGM(p, seed=SEED, instance=k) for k, (p, _) in enumerate(vs)]
        for f in range(4):
            _say(f"batch frame {f}")
            meas = [sc.measurement(f) for _, sc in vs]
            poses = [sc.ego_pose(f) for _, sc in vs]
            batch.update(ids, meas, poses, 0.1)
            for k, pl in enumerate(plains):
                x, y, yaw = poses[k]
                pl.update_grid(meas[k], x, y, yaw, 0.1)
                _assert_vehicle_equal(f"frame {f} vehicle {k}", batch, ids[k], pl)


def test_mixed_schedule_and_late_join():
    """Vehicles step on different subsets (continuous batching); one joins late and one leaves."""
    vs = [_vehicle(k) for k in range(len(FLEET))]
    schedule = [[0, 1], [0, 2], [1, 2, 0], [3, 0], [3, 1, 2], [2, 3, 0]]
    with Batch() as batch:
        ids = {}
        plains = {}
        frames = {k: 0 for k in range(len(FLEET))}
        for step, members in enumerate(schedule):
            _say(f"mixed step {step}: {members}")
            for k in members:
                if k not in ids:
                    p, _ = vs[k]
                    ids[k] = batch.add_vehicle(p, SEED, 100 + k)
                    plains[k] = DOGM(p, seed=SEED, instance=100 + k)
            meas, poses = [], []
            for k in members:
                _, sc = vs[k]
                meas.append(sc.measurement(frames[k]))
                poses.append(sc.ego_pose(frames[k]))
            batch.update([ids[k] for k in members], meas, poses, 0.1)
            for i, k in enumerate(members):
                x, y, yaw = poses[i]
                plains[k].update_grid(meas[i], x, y, yaw, 0.1)
                frames[k] += 1
                _assert_vehicle_equal(f"step {step} vehicle {k}", batch, ids[k], plains[k])
            if step == 4:
                batch.remove_vehicle(ids.pop(1))
                plains.pop(1).close()
                vs[1] = None


def test_batch_vehicle_against_checker_and_device_measurements():
    """A batched vehicle fed through its device measurement buffer equals the f32 checker."""
    p, sc = _vehicle(0)
    q, sq = _vehicle(3)
    orc = OracleDOGM(p, seed=SEED, instance=5)
    with Batch() as batch:
        a = batch.add_vehicle(p, SEED, 5)
        b = batch.add_vehicle(q, SEED, 6)
        buf = DeviceBuffer(sq.measurement(0).nbytes)
        for f in range(3):
            _say(f"checker frame {f}")
            m = sc.measurement(f)
            # vehicle b: measurement written straight into its buffer on the device
            mb = np.ascontiguousarray(sq.measurement(f))
            buf.upload(mb)
            from dogm_ctypes import lib
            assert lib.dogm_memcpy(batch.measurement_buffer(b), buf.ptr, mb.nbytes, 2) == 0
            batch.update([a, b], [m, None], [sc.ego_pose(f), sq.ego_pose(f)], [0.1, 0.1])
            x, y, yaw = sc.ego_pose(f)
            orc.update_grid(m, x, y, yaw, 0.1)
            assert_particles_equal(f"frame {f} particles", batch.get_particles(a), orc.get_particles())
            assert_grid_equal(f"frame {f} grid", batch.get_grid_cells(a), orc.get_grid_cells())
            assert_int_equal(f"frame {f} ancestors", batch.get_resampled_indices(a), orc.get_resampled_indices())
        buf.free()


def test_batch_rejects_bad_steps():
    p, sc = _vehicle(2)
    with Batch() as batch:
        a = batch.add_vehicle(p, SEED, 1)
        m = sc.measurement(0)
        with pytest.raises(RuntimeError):
            batch.update([a, a], [m, m], [(0, 0, 0), (0, 0, 0)], 0.1)
        with pytest.raises(RuntimeError):
            batch.update([a + 99], [m], [(0, 0, 0)], 0.1)
        batch.update([a], [m], [(0, 0, 0)], 0.1)
