"""Step graphs: an update step captured once and replayed gives the same state as eager launches,
frame after frame, with eager work (read-backs, extraction, stage calls) between the replays."""
import sys

import numpy as np
import pytest

from dogm_ctypes import DOGM, EXTRACT_IMAGE, EXTRACT_VELOCITY, Batch
from parity_util import assert_bits_equal, assert_grid_equal, assert_int_equal, assert_particles_equal, \
    small_params, small_scene

pytestmark = pytest.mark.gpu
SEED = 23


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(tag, a, b):
    assert_particles_equal(f"{tag} particles", a.get_particles(), b.get_particles())
    assert_grid_equal(f"{tag} grid", a.get_grid_cells(), b.get_grid_cells())
    assert_int_equal(f"{tag} ancestors", a.get_resampled_indices(), b.get_resampled_indices())


def test_plain_graph_replays_match_eager():
    p = small_params(gs=48, res=0.5, n=20000, b=2000)
    sc = small_scene(gs=48, res=0.5)
    g, e = DOGM(p, seed=SEED, instance=1), DOGM(p, seed=SEED, instance=1)
    g.set_graph(True)
    for f in range(7):
        _say(f"graph plain frame {f}")
        x, y, yaw = sc.ego_pose(f)
        m = sc.measurement(f)
        g.update_grid(m, x, y, yaw, 0.1)
        e.update_grid(m, x, y, yaw, 0.1)
        _same(f"frame {f}", g, e)
        # eager work between replays
        assert_bits_equal(f"frame {f} occupancy", g.get_occupancy(), e.get_occupancy())
        if f == 3:
            for h in (g, e):
                h.set_cdf_mode("fp64")  # a new signature: recaptured
        if f == 5:
            for h in (g, e):
                h.particle_prediction(0.1)  # stage calls between graph steps
                h.particle_assignment()
            _same(f"frame {f} stages", g, e)


def test_batch_graph_replays_match_eager():
    vs = [(small_params(gs=gs, res=0.5, n=n, b=n // 10), small_scene(gs=gs, res=0.5, seed=30 + k))
          for k, (gs, n) in enumerate([(40, 9000), (56, 16000), (32, 5000)])]
    with Batch() as gb, Batch() as eb:
        gids = [gb.add_vehicle(p, SEED, 10 + k) for k, (p, _) in enumerate(vs)]
        eids = [eb.add_vehicle(p, SEED, 10 + k) for k, (p, _) in enumerate(vs)]
        gb.set_graph(True)
        for b in (gb, eb):
            b.set_extraction(EXTRACT_IMAGE | EXTRACT_VELOCITY, 0.55, 0.5)
        for f in range(6):
            _say(f"graph batch frame {f}")
            meas = [sc.measurement(f) for _, sc in vs]
            poses = [sc.ego_pose(f) for _, sc in vs]
            gb.update(gids, meas, poses, 0.1)
            eb.update(eids, meas, poses, 0.1)
            for k in range(3):
                tag = f"frame {f} vehicle {k}"
                assert_particles_equal(tag, gb.get_particles(gids[k]), eb.get_particles(eids[k]))
                assert_grid_equal(tag, gb.get_grid_cells(gids[k]), eb.get_grid_cells(eids[k]))
                assert_int_equal(f"{tag} image", gb.get_occupancy_image(gids[k]), eb.get_occupancy_image(eids[k]))
                rg, cg = gb.get_cells_with_velocity(gids[k])
                re_, ce = eb.get_cells_with_velocity(eids[k])
                assert cg == ce and rg.tobytes() == re_.tobytes(), tag
