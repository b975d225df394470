"""Serving form: admission by device-memory budget, continuous batching of mixed-size vehicles, and
every result equal to the same vehicle run alone (bit for bit)."""
import random
import sys
import time

import numpy as np
import pytest

from dogm_ctypes import DOGM, DeviceBuffer, Server, vehicle_footprint
from parity_util import assert_grid_equal, assert_particles_equal, small_params, small_scene

pytestmark = pytest.mark.gpu

SEED = 91
FLEET = [(40, 0.5, 3000, 300), (64, 0.4, 9001, 901), (24, 0.5, 777, 0), (80, 0.25, 12000, 1200)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vehicle(k):
    gs, res, n, b = FLEET[k]
    return small_params(gs=gs, res=res, n=n, b=b), small_scene(gs=gs, res=res, seed=30 + k)


def test_admission_by_budget():
    pa, _ = _vehicle(0)
    pb, _ = _vehicle(1)
    fa, fb = vehicle_footprint(pa), vehicle_footprint(pb)
    with Server(hbm_budget_bytes=fa + fb - 1) as s:
        a = s.admit(pa, SEED, 1)
        assert a is not None and s.resident_bytes() == fa
        assert s.admit(pb, SEED, 2) is None, "over budget must be refused"
        s.release(a)
        assert s.resident_bytes() == 0
        b = s.admit(pb, SEED, 2)
        assert b is not None and s.resident_bytes() == fb


def _frames_alone(k, nframes, instance):
    p, sc = _vehicle(k)
    g = DOGM(p, seed=SEED, instance=instance)
    for f in range(nframes):
        x, y, yaw = sc.ego_pose(f)
        g.update_grid(sc.measurement(f), x, y, yaw, 0.1)
    return g


def test_mixed_arrivals_step_mode_equal_alone():
    """Random interleaved arrivals (several frames of one vehicle may be queued at once); each step
    takes at most one frame per vehicle; per-vehicle order and results match alone runs."""
    rng = random.Random(5)
    nframes = {0: 4, 1: 3, 2: 5, 3: 3}
    with Server(hbm_budget_bytes=1 << 40) as s:
        vids = {}
        bufs = []
        arrivals = []
        for k, nf in nframes.items():
            p, sc = _vehicle(k)
            vids[k] = s.admit(p, SEED, 200 + k)
            t = rng.uniform(0, 0.05)
            for f in range(nf):
                t += rng.uniform(0.0, 0.12)
                arrivals.append((t, k, f))
        arrivals.sort()
        results = []
        now = 0.0
        i = 0
        while i < len(arrivals) or s.queued():
            # admit everything that has arrived, in bursts
            burst = rng.randint(1, 4)
            while i < len(arrivals) and (arrivals[i][0] <= now or burst > 0):
                t, k, f = arrivals[i]
                _, sc = _vehicle(k)
                d = DeviceBuffer(sc.measurement(f).nbytes)
                d.upload(np.ascontiguousarray(sc.measurement(f)))
                bufs.append(d)
                s.submit(vids[k], d, sc.ego_pose(f), 0.1, t)
                now = max(now, t)
                burst -= 1
                i += 1
            n, secs = s.step(now)
            assert n >= 1
            now += secs
            results += s.poll()
        _say(f"{len(results)} frames in {s.stats()[0]} steps")
        assert len(results) == sum(nframes.values())
        assert all(r["status"] == 0 for r in results)
        assert any(r["batch_size"] > 1 for r in results)
        by_vehicle = {}
        for r in results:
            by_vehicle.setdefault(r["vehicle_id"], []).append(r["ticket"])
        for tickets in by_vehicle.values():
            assert tickets == sorted(tickets), "a vehicle's frames must complete in submission order"
        for k, nf in nframes.items():
            g = _frames_alone(k, nf, 200 + k)
            assert_particles_equal(f"vehicle {k} particles", s.get_particles(vids[k]), g.get_particles())
            assert_grid_equal(f"vehicle {k} grid", s.get_grid_cells(vids[k]), g.get_grid_cells())
            g.close()
        for d in bufs:
            d.free()


def test_worker_thread_continuous_batching():
    """Frames submitted while the worker is stepping join later steps; results equal alone runs."""
    nframes = 4
    with Server(hbm_budget_bytes=1 << 40) as s:
        vids = {k: s.admit(_vehicle(k)[0], SEED, 300 + k) for k in range(len(FLEET))}
        meas = {(k, f): _vehicle(k)[1].measurement(f) for k in range(len(FLEET)) for f in range(nframes)}
        s.start()
        for f in range(nframes):
            for k in range(len(FLEET)):
                _, sc = _vehicle(k)
                s.submit(vids[k], meas[(k, f)], sc.ego_pose(f), 0.1, s.clock())
            time.sleep(0.002)
        s.wait_idle()
        s.stop()
        res = s.poll()
        assert len(res) == nframes * len(FLEET) and all(r["status"] == 0 for r in res)
        assert all(r["t_done"] >= r["t_start"] >= r["t_arrival"] - 1e-3 for r in res)
        for k in range(len(FLEET)):
            g = _frames_alone(k, nframes, 300 + k)
            assert_particles_equal(f"vehicle {k} particles", s.get_particles(vids[k]), g.get_particles())
            assert_grid_equal(f"vehicle {k} grid", s.get_grid_cells(vids[k]), g.get_grid_cells())
            g.close()
