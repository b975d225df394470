"""Result extraction on the GPU (pignistic occupancy, 8-bit image, cells with velocity) against the
checker applied to the same cells, bit for bit; batched and served vehicles against plain handles."""
import sys

import numpy as np
import pytest

import oracle_ctypes as oc
from dogm_ctypes import (DOGM, EXTRACT_IMAGE, EXTRACT_OCCUPANCY, EXTRACT_VELOCITY, GRID_CELL_DTYPE, Batch, Server,
                         vehicle_footprint)
from parity_util import assert_bits_equal, assert_int_equal, small_params, small_scene

pytestmark = pytest.mark.gpu

SEED = 31
THRESHOLDS = [(0.7, 4.0), (0.55, 0.5), (0.0, 0.0)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_velocity(tag, recs, count, cells, gs, min_occ, min_vel):
    want = oc.cells_with_velocity(gs, cells, min_occ, min_vel)
    assert count == want.size, f"{tag}: {count} cells vs {want.size}"
    idx = (recs["y"].astype(np.int64) * gs + recs["x"].astype(np.int64))
    assert_int_equal(f"{tag} cells", idx, want[:recs.size])
    assert np.array_equal(recs["x"], (want[:recs.size] % gs).astype(np.float32))
    assert recs["data"].tobytes() == cells[want[:recs.size]].tobytes(), f"{tag}: record data"


def _check_handle(tag, g, gs):
    cells = g.get_grid_cells()
    p, img = oc.occupancy(cells["occ_mass"], cells["free_mass"])
    assert_bits_equal(f"{tag} occupancy", g.get_occupancy(), p)
    assert_int_equal(f"{tag} image", g.get_occupancy_image(), img)
    meas = g.get_measurement_cells()
    mp, mimg = oc.occupancy(meas["occ_mass"], meas["free_mass"])
    assert_bits_equal(f"{tag} measurement occupancy", g.get_occupancy("measurement"), mp)
    assert_int_equal(f"{tag} measurement image", g.get_occupancy_image("measurement"), mimg)
    total = 0
    for mo, mv in THRESHOLDS:
        recs, count = g.get_cells_with_velocity(mo, mv)
        _check_velocity(f"{tag} ({mo}, {mv})", recs, count, cells, gs, mo, mv)
        total += count
    return total


def test_plain_handle_frames():
    gs = 64
    p = small_params(gs=gs, res=0.5, n=40000, b=4000)
    sc = small_scene(gs=gs, res=0.5, n_boxes=8)
    g = DOGM(p, seed=SEED, instance=1)
    moving = 0
    for f in range(4):
        _say(f"plain extraction frame {f}")
        x, y, yaw = sc.ego_pose(f)
        g.update_grid(sc.measurement(f), x, y, yaw, 0.1)
        moving = _check_handle(f"frame {f}", g, gs)
    assert moving > 0


def test_injected_grid_all_cells_and_capacity():
    """90,000 cells all passing (thresholds -inf): two record chunks; truncated capacities."""
    gs = 300
    p = small_params(gs=gs, res=0.2, n=2000, b=200)
    g = DOGM(p, seed=SEED, instance=2)
    rng = np.random.default_rng(4)
    cells = np.zeros(gs * gs, dtype=GRID_CELL_DTYPE)
    for k in GRID_CELL_DTYPE.names[2:]:
        cells[k] = rng.normal(size=gs * gs).astype(np.float32)
    cells["start_idx"] = rng.integers(-1, 2000, gs * gs)
    cells["end_idx"] = rng.integers(-1, 2000, gs * gs)
    cells["var_x_vel"] = np.abs(cells["var_x_vel"]) + 0.1
    cells["var_y_vel"] = np.abs(cells["var_y_vel"]) + 0.1
    cells["covar_xy_vel"][::7] = 0.0
    cells["var_x_vel"][::13] = 0.0   # some singular covariances
    cells["covar_xy_vel"][::13] = 0.0
    g.set_grid_cells(cells)
    _say("injected grid, thresholds -inf")
    recs, count = g.get_cells_with_velocity(-np.inf, -np.inf)
    assert count == gs * gs
    _check_velocity("all", recs, count, cells, gs, -np.inf, -np.inf)
    for cap in (0, 1, 777, 65536, 65537):
        _say(f"capacity {cap}")
        recs, count = g.get_cells_with_velocity(-np.inf, -np.inf, capacity=cap)
        assert count == gs * gs and recs.size == cap
        _check_velocity(f"capacity {cap}", recs, count, cells, gs, -np.inf, -np.inf)
    for mo, mv in [(0.3, 1.0), (0.6, 3.0)]:
        _say(f"injected grid ({mo}, {mv})")
        recs, count = g.get_cells_with_velocity(mo, mv)
        _check_velocity(f"({mo}, {mv})", recs, count, cells, gs, mo, mv)
    p_occ, img = oc.occupancy(cells["occ_mass"], cells["free_mass"])
    assert_bits_equal("injected occupancy", g.get_occupancy(), p_occ)
    assert_int_equal("injected image", g.get_occupancy_image(), img)


def _vehicles():
    sizes = [(40, 9000, 900), (64, 30000, 3000), (24, 4000, 400)]
    return [(small_params(gs=gs, res=0.5, n=n, b=b), small_scene(gs=gs, res=0.5, seed=70 + k), gs)
            for k, (gs, n, b) in enumerate(sizes)]


@pytest.mark.parametrize("in_step", [True, False])
def test_batch_matches_plain(in_step):
    vs = _vehicles()
    what = EXTRACT_OCCUPANCY | EXTRACT_IMAGE | EXTRACT_VELOCITY
    with Batch() as batch:
        ids = [batch.add_vehicle(p, SEED, 80 + k) for k, (p, _, _) in enumerate(vs)]
        # what = 0: nothing extracted inside the steps, the getters extract on demand with these thresholds
        batch.set_extraction(what if in_step else 0, 0.55, 0.5)
        plains = [DOGM(p, seed=SEED, instance=80 + k) for k, (p, _, _) in enumerate(vs)]
        for f in range(3):
            _say(f"batch extraction (in step: {in_step}) frame {f}")
            meas = [sc.measurement(f) for _, sc, _ in vs]
            poses = [sc.ego_pose(f) for _, sc, _ in vs]
            batch.update(ids, meas, poses, 0.1)
            for k, g in enumerate(plains):
                x, y, yaw = poses[k]
                g.update_grid(meas[k], x, y, yaw, 0.1)
                gs = vs[k][2]
                tag = f"frame {f} vehicle {k}"
                assert_bits_equal(f"{tag} occupancy", batch.get_occupancy(ids[k]), g.get_occupancy())
                assert_int_equal(f"{tag} image", batch.get_occupancy_image(ids[k]), g.get_occupancy_image())
                recs, count = batch.get_cells_with_velocity(ids[k])
                prec, pcount = g.get_cells_with_velocity(0.55, 0.5)
                assert count == pcount and recs.tobytes() == prec.tobytes(), tag
                _check_velocity(tag, recs, count, batch.get_grid_cells(ids[k]), gs, 0.55, 0.5)


def test_server_extraction_matches_plain():
    vs = _vehicles()[:2]
    what = EXTRACT_IMAGE | EXTRACT_VELOCITY
    budget = sum(vehicle_footprint(p) for p, _, _ in vs) + (64 << 20)
    with Server(budget) as srv:
        srv.set_extraction(what, 0.55, 0.5)
        ids = [srv.admit(p, SEED, 90 + k) for k, (p, _, _) in enumerate(vs)]
        assert all(i is not None for i in ids)
        plains = [DOGM(p, seed=SEED, instance=90 + k) for k, (p, _, _) in enumerate(vs)]
        for f in range(3):
            _say(f"server extraction frame {f}")
            for k, (_, sc, _) in enumerate(vs):
                srv.submit(ids[k], sc.measurement(f), sc.ego_pose(f), 0.1, srv.clock())
            while srv.queued():
                srv.step(srv.clock())
            assert all(r["status"] == 0 for r in srv.poll())
            for k, g in enumerate(plains):
                x, y, yaw = vs[k][1].ego_pose(f)
                g.update_grid(vs[k][1].measurement(f), x, y, yaw, 0.1)
                tag = f"server frame {f} vehicle {k}"
                assert_int_equal(f"{tag} image", srv.get_occupancy_image(ids[k]), g.get_occupancy_image())
                recs, count = srv.get_cells_with_velocity(ids[k])
                prec, pcount = g.get_cells_with_velocity(0.55, 0.5)
                assert count == pcount and recs.tobytes() == prec.tobytes(), tag
                # not extracted in the step: computed on demand
                assert_bits_equal(f"{tag} occupancy", srv.get_occupancy(ids[k]), g.get_occupancy())
