"""The batched and serving workloads at their real shapes on the GPU.

A fleet of 64 vehicles with grid sides drawn from {128, 256, 512, 768, 1024} at 0.1 m, N = 1.5 x cells,
B = N / 10, each with its own scene (5-30 boxes, 720-beam lidar) and ego trajectory at 2-15 m/s, so the
grid shifts on every frame.  Every frame's measurement grids are generated on the device from the
vehicles' lidar scans.

* Batched: one dogm_batch_update per frame over all 64 vehicles, eager and with step graphs (the
  graph is captured on frame 1 and replayed on frames 2 and 3).  Every vehicle is compared on every
  frame with a plain handle (same params, seed and instance) fed the same grid, and the graph batch
  with the eager batch, bit for bit.  A 128-side and a 1024-side vehicle are compared with the f32
  checker, which is fed the checker's own grid generator (byte-equal to the device grid, checked).
* Serving: the same fleet admitted to a Server under a 200 GB budget, frames submitted as lidar scans
  (grids generated inside the step) with extraction on.  Every served vehicle equals its plain
  handle (state, occupancy image and velocity cells) and a 1024-side vehicle equals the checker.

Each vehicle must reproduce the stage sequence of dogm/src/dogm.cu:115-131.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import fleet as fl
import oracle_ctypes as oc
from dogm_ctypes import (EXTRACT_IMAGE, EXTRACT_VELOCITY, MEAS_CELL_DTYPE, DOGM, Batch, DeviceBuffer, Lidar, Server,
                         lib, vehicle_footprint)
from oracle_ctypes import OracleDOGM
from parity_util import assert_int_equal, assert_vehicle_equal

pytestmark = pytest.mark.gpu

SEED = 123456
FLEET_SIZE = 64
FRAMES = 4
DT = 0.1


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Fleet:
    """Vehicles, their scenes, per-frame scans and one device generator per grid side."""

    def __init__(self):
        self.vehicles = fl.make_fleet(FLEET_SIZE, seed=2024, min_speed=2.0)
        self.sides = sorted({v.side for v in self.vehicles})
        assert self.sides == list(fl.SIDES), f"the fleet must cover every side: {self.sides}"
        self.scenes = [fl.vehicle_scene(v, i) for i, v in enumerate(self.vehicles)]
        self.ranges = [[sc.ranges(f) for f in range(FRAMES)] for sc in self.scenes]
        self.by_side = {s: [i for i, v in enumerate(self.vehicles) if v.side == s] for s in self.sides}
        self.lidar = {}
        for s in self.sides:
            sc = self.scenes[self.by_side[s][0]]
            cs, sn = sc.beam_directions()
            self.lidar[s] = Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn)
        # the vehicles checked against the checker: the first of the smallest and of the largest side
        self.checked = [self.by_side[self.sides[0]][0], self.by_side[self.sides[-1]][0]]

    def pose(self, i, f):
        return self.scenes[i].ego_pose(f)

    def generate(self, f, outs):
        """Frame f's grid of every vehicle i into the device address outs[i]."""
        for s, js in self.by_side.items():
            self.lidar[s].generate_many(np.stack([self.ranges[j][f] for j in js]), [outs[j] for j in js])

    def checker_grid(self, i, f):
        sc = self.scenes[i]
        c = sc.cfg
        cs, sn = sc.beam_directions()
        return oc.lidar_grid(c.grid_size, c.resolution, c.max_range, c.occ_mass, c.free_mass, cs, sn,
                             self.ranges[i][f])

    def close(self):
        for lid in self.lidar.values():
            lid.close()


@pytest.fixture(scope="module")
def fleet():
    f = Fleet()
    yield f
    f.close()


def _download_meas(ptr, cells):
    out = np.empty(cells, dtype=MEAS_CELL_DTYPE)
    assert lib.dogm_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 1) == 0
    return out


def test_fleet_shape(fleet):
    """The fleet is the batched workload: every side, N = 1.5 C, B = N / 10, a shift every frame."""
    for v in fleet.vehicles:
        c = v.side * v.side
        assert v.params.particle_count == int(1.5 * c)
        assert v.params.new_born_particle_count == v.params.particle_count // 10
        assert 5 <= v.boxes <= 30 and v.speed >= 2.0
    total = sum(vehicle_footprint(v.params) for v in fleet.vehicles)
    _say(f"fleet: {FLEET_SIZE} vehicles, sides {[len(fleet.by_side[s]) for s in fleet.sides]}, "
         f"{total / 2**30:.1f} GiB resident")


def test_fleet_batched_eager_and_graph_equal_plain_and_checker(fleet):
    n = FLEET_SIZE
    eager, graph = Batch(), Batch()
    graph.set_graph(True)
    ids_e = [eager.add_vehicle(v.params, SEED, i) for i, v in enumerate(fleet.vehicles)]
    ids_g = [graph.add_vehicle(v.params, SEED, i) for i, v in enumerate(fleet.vehicles)]
    plains = [DOGM(v.params, seed=SEED, instance=i) for i, v in enumerate(fleet.vehicles)]
    bufs = [DeviceBuffer(v.side * v.side * 16) for v in fleet.vehicles]
    checkers = {i: OracleDOGM(fleet.vehicles[i].params, seed=SEED, instance=i) for i in fleet.checked}
    last_pos = [None] * n
    try:
        for f in range(FRAMES):
            _say(f"fleet batched frame {f}: generate grids")
            fleet.generate(f, [eager.measurement_buffer(i) for i in ids_e])
            fleet.generate(f, [graph.measurement_buffer(i) for i in ids_g])
            fleet.generate(f, [b.ptr for b in bufs])
            poses = [fleet.pose(i, f) for i in range(n)]
            _say(f"fleet batched frame {f}: eager step")
            eager.update(ids_e, [None] * n, poses, DT)
            _say(f"fleet batched frame {f}: graph step")
            graph.update(ids_g, [None] * n, poses, DT)
            _say(f"fleet batched frame {f}: plain handles")
            for i, g in enumerate(plains):
                g.update_grid(bufs[i], *poses[i], DT)
            lib.dogm_device_synchronize()
            for i, g in enumerate(plains):
                tag = f"frame {f} vehicle {i} (side {fleet.vehicles[i].side})"
                pos = g.position()
                if f > 0:
                    assert pos != last_pos[i], f"{tag}: the grid must shift on every frame"
                last_pos[i] = pos
                pe, pp = eager.get_particles(ids_e[i]), g.get_particles()
                ge, gp = eager.get_grid_cells(ids_e[i]), g.get_grid_cells()
                assert_vehicle_equal(f"{tag} batch vs plain", pe, pp, ge, gp)
                del pp, gp
                assert_vehicle_equal(f"{tag} graph vs eager", graph.get_particles(ids_g[i]), pe,
                                     graph.get_grid_cells(ids_g[i]), ge)
                assert_int_equal(f"{tag} ancestors", eager.get_resampled_indices(ids_e[i]), g.get_resampled_indices())
            for i, orc in checkers.items():
                tag = f"frame {f} vehicle {i} (side {fleet.vehicles[i].side}) vs checker"
                _say(tag)
                m = fleet.checker_grid(i, f)
                got = _download_meas(eager.measurement_buffer(ids_e[i]), m.size)
                assert got.tobytes() == m.tobytes(), f"{tag}: device grid"
                orc.update_grid(m, *poses[i], DT)
                assert_vehicle_equal(tag, eager.get_particles(ids_e[i]), orc.get_particles(),
                                     eager.get_grid_cells(ids_e[i]), orc.get_grid_cells())
                assert_int_equal(f"{tag} ancestors", eager.get_resampled_indices(ids_e[i]),
                                 orc.get_resampled_indices())
    finally:
        for b in bufs:
            b.free()
        for g in plains:
            g.close()
        for orc in checkers.values():
            orc.close()
        eager.close()
        graph.close()


def test_fleet_served_from_scans_equal_plain_and_checker(fleet):
    n = FLEET_SIZE
    min_occ, min_vel = 0.7, 4.0
    plains = [DOGM(v.params, seed=SEED, instance=i) for i, v in enumerate(fleet.vehicles)]
    bufs = [DeviceBuffer(v.side * v.side * 16) for v in fleet.vehicles]
    ci = fleet.checked[-1]
    orc = OracleDOGM(fleet.vehicles[ci].params, seed=SEED, instance=ci)
    with Server(hbm_budget_bytes=200 * 10**9) as s:
        s.set_extraction(EXTRACT_IMAGE | EXTRACT_VELOCITY, min_occ, min_vel)
        vids = [s.admit(v.params, SEED, i) for i, v in enumerate(fleet.vehicles)]
        assert all(v is not None for v in vids), "the fleet fits the 200 GB budget"
        now = 0.0
        try:
            for f in range(FRAMES):
                _say(f"fleet served frame {f}: submit scans and step")
                for i, v in enumerate(fleet.vehicles):
                    s.submit_scan(vids[i], fleet.lidar[v.side], fleet.ranges[i][f], fleet.pose(i, f), DT,
                                  now + 1e-4 * i)
                results = []
                while s.queued():
                    _, secs = s.step(now + 0.01)
                    now += max(secs, 1e-3)
                    results += s.poll(1 << 12)
                results += s.poll(1 << 12)
                assert len(results) == n and all(r["status"] == 0 for r in results)
                assert max(r["batch_size"] for r in results) > 1, "mixed-size batches per step"
                _say(f"fleet served frame {f}: plain handles")
                fleet.generate(f, [b.ptr for b in bufs])
                for i, g in enumerate(plains):
                    g.update_grid(bufs[i], *fleet.pose(i, f), DT)
                lib.dogm_device_synchronize()
                for i, g in enumerate(plains):
                    tag = f"served frame {f} vehicle {i} (side {fleet.vehicles[i].side})"
                    assert_vehicle_equal(tag, s.get_particles(vids[i]), g.get_particles(), s.get_grid_cells(vids[i]),
                                         g.get_grid_cells())
                    assert_int_equal(f"{tag} image", s.get_occupancy_image(vids[i]), g.get_occupancy_image())
                    cs, ncs = s.get_cells_with_velocity(vids[i])
                    cp, ncp = g.get_cells_with_velocity(min_occ, min_vel)
                    assert ncs == ncp and cs.tobytes() == cp.tobytes(), f"{tag} velocity cells"
                tag = f"served frame {f} vehicle {ci} vs checker"
                _say(tag)
                m = fleet.checker_grid(ci, f)
                orc.update_grid(m, *fleet.pose(ci, f), DT)
                assert_vehicle_equal(tag, s.get_particles(vids[ci]), orc.get_particles(), s.get_grid_cells(vids[ci]),
                                     orc.get_grid_cells())
        finally:
            for b in bufs:
                b.free()
            for g in plains:
                g.close()
            orc.close()
