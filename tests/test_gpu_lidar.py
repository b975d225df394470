"""Measurement-grid generation on the GPU against the checker, bit for bit, and a batched step fed
from grids generated into the vehicles' measurement buffers."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_ctypes as oc
import scene as scene_mod
from dogm_ctypes import MEAS_CELL_DTYPE, Batch, DeviceBuffer, Lidar, lib
from parity_util import assert_grid_equal, assert_particles_equal, small_params, small_scene

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _download(buf: DeviceBuffer, cells: int) -> np.ndarray:
    out = np.empty(cells, dtype=MEAS_CELL_DTYPE)
    assert lib.dogm_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.ptr), out.nbytes, 1) == 0
    return out


def _checker(sc, frame):
    c = sc.cfg
    cs, sn = sc.beam_directions()
    return oc.lidar_grid(c.grid_size, c.resolution, c.max_range, c.occ_mass, c.free_mass, cs, sn, sc.ranges(frame))


@pytest.mark.parametrize("name", ["small", "plain_small", "plain_full"])
def test_generate_matches_checker(name):
    sc = small_scene(gs=64, res=0.5) if name == "small" else scene_mod.make_scene(name)
    cs, sn = sc.beam_directions()
    cells = sc.cfg.grid_size ** 2
    with Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn) as lidar:
        buf = DeviceBuffer(cells * 16)
        for f in (0, 2, 5):
            _say(f"lidar {name} frame {f}")
            lidar.generate(sc.ranges(f), buf.ptr)
            got = _download(buf, cells)
            assert got.tobytes() == _checker(sc, f).tobytes(), f"{name} frame {f}"
        buf.free()


def test_generate_many_and_edge_ranges():
    sc = small_scene(gs=80, res=0.25, n_beams=240, seed=12)
    cs, sn = sc.beam_directions()
    cells = sc.cfg.grid_size ** 2
    ranges = [sc.ranges(f) for f in range(5)]
    ranges[3] = ranges[3].copy()
    ranges[3][::7] = np.nan
    ranges[3][1::7] = 1e9
    ranges[3][2::7] = 0.0
    with Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn) as lidar:
        bufs = [DeviceBuffer(cells * 16) for _ in ranges]
        _say("lidar generate_many x5")
        lidar.generate_many(np.stack(ranges), [b.ptr for b in bufs])
        for i, b in enumerate(bufs):
            want = oc.lidar_grid(sc.cfg.grid_size, sc.cfg.resolution, sc.cfg.max_range, sc.cfg.occ_mass,
                                 sc.cfg.free_mass, cs, sn, ranges[i])
            assert _download(b, cells).tobytes() == want.tobytes(), f"grid {i}"
        for b in bufs:
            b.free()
    # default directions: 2 pi b / n with the host C library's cos / sin; the checker gets the same
    libm = C.CDLL("libm.so.6")
    libm.cos.restype = libm.sin.restype = C.c_double
    libm.cos.argtypes = libm.sin.argtypes = [C.c_double]
    n = sc.cfg.n_beams
    ang = [2.0 * np.pi * b / n for b in range(n)]
    with Lidar(**sc.lidar_params()) as lidar:
        buf = DeviceBuffer(cells * 16)
        lidar.generate(ranges[0], buf.ptr)
        want = oc.lidar_grid(sc.cfg.grid_size, sc.cfg.resolution, sc.cfg.max_range, sc.cfg.occ_mass,
                             sc.cfg.free_mass, [libm.cos(a) for a in ang], [libm.sin(a) for a in ang], ranges[0])
        assert _download(buf, cells).tobytes() == want.tobytes()
        buf.free()


def test_batch_fed_from_device_generated_grids():
    vs = [(small_params(gs=48, res=0.5, n=8000, b=800), small_scene(gs=48, res=0.5, seed=40 + k)) for k in range(3)]
    with Batch() as gen_batch, Batch() as host_batch:
        gids = [gen_batch.add_vehicle(p, 7, 30 + k) for k, (p, _) in enumerate(vs)]
        hids = [host_batch.add_vehicle(p, 7, 30 + k) for k, (p, _) in enumerate(vs)]
        sc0 = vs[0][1]
        cs, sn = sc0.beam_directions()
        with Lidar(**sc0.lidar_params(), beam_cos=cs, beam_sin=sn) as lidar:
            for f in range(3):
                _say(f"device-generated grids, frame {f}")
                lidar.generate_many(np.stack([sc.ranges(f) for _, sc in vs]),
                                    [gen_batch.measurement_buffer(v) for v in gids])
                poses = [sc.ego_pose(f) for _, sc in vs]
                gen_batch.update(gids, [None] * 3, poses, 0.1)
                host_batch.update(hids, [sc.measurement(f) for _, sc in vs], poses, 0.1)
                for k in range(3):
                    assert_particles_equal(f"frame {f} vehicle {k}", gen_batch.get_particles(gids[k]),
                                           host_batch.get_particles(hids[k]))
                    assert_grid_equal(f"frame {f} vehicle {k}", gen_batch.get_grid_cells(gids[k]),
                                      host_batch.get_grid_cells(hids[k]))


def test_server_scan_frames_match_plain_handles():
    """Frames submitted as scans (grids generated inside the serving step) give the same filter state
    as plain handles fed the numpy generator's grids."""
    from dogm_ctypes import DOGM, Server, vehicle_footprint
    vs = [(small_params(gs=gs, res=0.5, n=6000, b=600), small_scene(gs=gs, res=0.5, seed=60 + k))
          for k, gs in enumerate((40, 56, 40))]
    lidars = {}
    for _, sc in vs:
        if sc.cfg.grid_size not in lidars:
            cs, sn = sc.beam_directions()
            lidars[sc.cfg.grid_size] = Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn)
    budget = sum(vehicle_footprint(p) for p, _ in vs) + (16 << 20)
    with Server(budget) as srv:
        ids = [srv.admit(p, 5, 70 + k) for k, (p, _) in enumerate(vs)]
        plains = [DOGM(p, seed=5, instance=70 + k) for k, (p, _) in enumerate(vs)]
        for f in range(3):
            _say(f"server scan frames, frame {f}")
            for k, (_, sc) in enumerate(vs):
                srv.submit_scan(ids[k], lidars[sc.cfg.grid_size], sc.ranges(f), sc.ego_pose(f), 0.1, srv.clock())
            while srv.queued():
                srv.step(srv.clock())
            assert all(r["status"] == 0 for r in srv.poll())
            for k, (_, sc) in enumerate(vs):
                x, y, yaw = sc.ego_pose(f)
                plains[k].update_grid(sc.measurement(f), x, y, yaw, 0.1)
                assert_particles_equal(f"frame {f} vehicle {k}", srv.get_particles(ids[k]), plains[k].get_particles())
                assert_grid_equal(f"frame {f} vehicle {k}", srv.get_grid_cells(ids[k]), plains[k].get_grid_cells())
    for lid in lidars.values():
        lid.close()
