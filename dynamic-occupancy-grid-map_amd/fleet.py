"""Fleets of simulated vehicles for the batched and serving forms, and a virtual-time driver.

A fleet vehicle has its own grid size (sides drawn from {128, 256, 512, 768, 1024} cells at 0.1 m),
N = 1.5 x cells persistent particles, B = N / 10 births, its own seeded scene (5-30 boxes, 720-beam
lidar) and ego motion.  Frames arrive at 10 Hz per vehicle with +-20 ms jitter and a random phase, so
the merged stream of many vehicles is close to Poisson.

`simulate` runs such a stream through a `dogm_ctypes.Server` in virtual time: arrivals are submitted
when the virtual clock passes them, each step takes the real GPU duration measured by the server,
and the latency of a frame is (step end - arrival).  This gives the latency distribution of the
real service times without waiting out the simulated seconds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

import scene as scene_mod
from dogm_ctypes import DeviceBuffer, Lidar, Params

SIDES = (128, 256, 512, 768, 1024)
RESOLUTION = 0.1


def fleet_params(side: int, particles_per_cell: float = 1.5) -> Params:
    n = int(particles_per_cell * side * side)
    return Params(size=side * RESOLUTION, resolution=RESOLUTION, particle_count=n, new_born_particle_count=n // 10,
                  persistence_prob=0.99, stddev_process_noise_position=0.1, stddev_process_noise_velocity=2.0,
                  birth_prob=0.02, stddev_velocity=30.0, init_max_velocity=15.0, freespace_discount=0.01)


@dataclass
class FleetVehicle:
    side: int
    params: Params
    speed: float      # m/s
    yaw_rate: float   # rad/s
    phase: float      # s, first arrival
    boxes: int

    def pose(self, frame: int, dt: float = 0.1) -> tuple[float, float, float]:
        t = frame * dt
        if abs(self.yaw_rate) < 1e-9:
            return (self.speed * t, 0.0, 0.0)
        r = self.speed / self.yaw_rate
        th = self.yaw_rate * t
        return (r * math.sin(th), r * (1.0 - math.cos(th)), th)


def make_fleet(count: int, seed: int = 7, sides=SIDES, min_speed: float = 0.0) -> list[FleetVehicle]:
    """`count` vehicles with sides drawn from `sides`; speeds uniform in [min_speed, 15) m/s (above
    1 m/s every 0.1 s frame moves the grid by at least one 0.1 m cell)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        side = int(rng.choice(sides))
        out.append(FleetVehicle(side=side, params=fleet_params(side), speed=float(rng.uniform(min_speed, 15.0)),
                                yaw_rate=float(rng.uniform(-0.2, 0.2)), phase=float(rng.uniform(0.0, 0.1)),
                                boxes=int(rng.integers(5, 31))))
    return out


def vehicle_scene(fv: FleetVehicle, index: int, seed: int = 3) -> scene_mod.Scene:
    """The vehicle's own scene: its box count, 720-beam lidar reaching the grid edge, and an ego
    trajectory with its speed and yaw rate (Scene.ego_pose equals FleetVehicle.pose)."""
    cfg = scene_mod.SceneConfig(grid_size=fv.side, resolution=RESOLUTION, n_boxes=fv.boxes, n_beams=720,
                                max_range=0.5 * fv.side * RESOLUTION, ego_speed=fv.speed, ego_yaw_rate=fv.yaw_rate,
                                seed=seed * 1000 + index)
    return scene_mod.Scene(cfg)


def side_scene(side: int, seed: int = 3) -> scene_mod.Scene:
    """The scene (15 moving boxes, 720-beam lidar reaching the grid edge) of a grid side."""
    cfg = scene_mod.SceneConfig(grid_size=side, resolution=RESOLUTION, n_boxes=15, n_beams=720,
                                max_range=0.5 * side * RESOLUTION, ego_speed=10.0, ego_yaw_rate=0.1, seed=seed + side)
    return scene_mod.Scene(cfg)


class ScanBank:
    """A few lidar scans per grid side and one device generator per side: frames are submitted as
    scans and their grids generated on the device inside the serving step."""

    def __init__(self, sides, frames: int = 4, seed: int = 3, scenes=None):
        self.lidar = {}
        self.ranges = {}
        for side in sorted(set(sides)):
            sc = scenes[side] if scenes else side_scene(side, seed)
            cs, sn = sc.beam_directions()
            self.lidar[side] = Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn)
            self.ranges[side] = [sc.ranges(f) for f in range(frames)]

    def submit(self, server, vid, side, k, pose, dt, t_arrival):
        r = self.ranges[side]
        server.submit_scan(vid, self.lidar[side], r[k % len(r)], pose, dt, t_arrival)

    def close(self):
        for lid in self.lidar.values():
            lid.close()


def measurement_bank(sides, frames: int = 4, seed: int = 3) -> dict[int, list[DeviceBuffer]]:
    """A few device-resident measurement grids per grid side (720-beam lidar, moving boxes); the
    vehicles of one side cycle through them."""
    bank = {}
    for side in sorted(set(sides)):
        sc = side_scene(side, seed)
        bufs = []
        for f in range(frames):
            m = np.ascontiguousarray(sc.measurement(f))
            d = DeviceBuffer(m.nbytes)
            d.upload(m)
            bufs.append(d)
        bank[side] = bufs
    return bank


def arrivals(fleet: list[FleetVehicle], seconds: float, rate_multiplier: float = 1.0, period: float = 0.1,
             jitter: float = 0.02, seed: int = 11) -> list[tuple[float, int, int]]:
    """(t_arrival, vehicle index, frame number), sorted by time.  Each vehicle sends one frame per
    period / rate_multiplier with uniform jitter of +-jitter (scaled likewise)."""
    rng = np.random.default_rng(seed)
    per = period / rate_multiplier
    jit = jitter / rate_multiplier
    out = []
    for v, fv in enumerate(fleet):
        n = int(seconds / per)
        base = fv.phase / rate_multiplier + per * np.arange(n)
        t = base + rng.uniform(-jit, jit, size=n)
        t = np.maximum(t, 0.0)
        t = np.maximum.accumulate(t)  # a vehicle's frames keep their order
        out += [(float(t[k]), v, k) for k in range(n)]
    out.sort()
    return out


def run_wallclock(server, fleet, vids, bank, events, dt: float = 0.1) -> dict:
    """Runs `events` through `server` in real time: the server's worker thread steps while this thread
    submits every frame when the wall clock reaches its arrival time (the scheduled arrival, on the
    server's clock, is the frame's t_arrival, so a late submission counts against the latency);
    latency = result time - arrival time, both on the server's clock."""
    import time
    steps0, frames0 = server.stats()
    lat, batch = [], []

    def collect():
        for r in server.poll(1 << 16):
            if r["status"] != 0:
                raise RuntimeError(f"frame {r['ticket']} failed with status {r['status']}")
            lat.append(r["t_done"] - r["t_arrival"])
            batch.append(r["batch_size"])

    n_ev = len(events)
    server.start()
    t0 = server.clock()
    i = 0
    try:
        while i < n_ev:
            now = server.clock() - t0
            while i < n_ev and events[i][0] <= now:
                ta, v, k = events[i]
                fv = fleet[v]
                if isinstance(bank, ScanBank):
                    bank.submit(server, vids[v], fv.side, k, fv.pose(k), dt, t0 + ta)
                else:
                    server.submit(vids[v], bank[fv.side][k % len(bank[fv.side])], fv.pose(k), dt, t0 + ta)
                i += 1
            collect()
            if i < n_ev:
                wait = events[i][0] - (server.clock() - t0)
                if wait > 2e-4:
                    time.sleep(min(wait, 2e-3))
    finally:
        server.stop()  # runs the queue empty, then joins the worker
    collect()
    span = server.clock() - t0
    steps1, frames1 = server.stats()
    lat = np.asarray(lat)
    return {
        "frames": int(lat.size),
        "wall_seconds": round(span, 3),
        "frames_per_s": round(lat.size / span, 2) if span > 0 else None,
        "offered_frames_per_s": round(n_ev / (events[-1][0] - events[0][0]), 2) if n_ev > 1 else None,
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "max_ms": round(1e3 * float(lat.max()), 3),
        "steps": steps1 - steps0,
        "mean_batch": round(float(np.mean(batch)), 2),
        "clock": "wall (server worker thread; arrivals submitted in real time)",
    }


def simulate(server, fleet, vids, bank, events, dt: float = 0.1, progress=None) -> dict:
    """Runs `events` through `server` in virtual time; returns latency and throughput figures."""
    t = 0.0
    i = 0
    busy = 0.0
    steps = 0
    lat = []
    batch = []
    n_ev = len(events)
    while i < n_ev or server.queued():
        while i < n_ev and events[i][0] <= t:
            ta, v, k = events[i]
            fv = fleet[v]
            if isinstance(bank, ScanBank):
                bank.submit(server, vids[v], fv.side, k, fv.pose(k), dt, ta)
            else:
                server.submit(vids[v], bank[fv.side][k % len(bank[fv.side])], fv.pose(k), dt, ta)
            i += 1
        if not server.queued():
            t = events[i][0]
            continue
        n, secs = server.step(t)
        t += secs
        busy += secs
        steps += 1
        for r in server.poll(1 << 16):
            if r["status"] != 0:
                raise RuntimeError(f"frame {r['ticket']} failed with status {r['status']}")
            lat.append(r["t_done"] - r["t_arrival"])
            batch.append(r["batch_size"])
        if progress and steps % 500 == 0:
            progress(f"{i}/{n_ev} frames submitted, virtual t = {t:.2f} s")
    span = max(t, events[-1][0]) - events[0][0] if events else 0.0
    lat = np.asarray(lat)
    return {
        "frames": int(lat.size),
        "virtual_seconds": round(span, 4),
        "frames_per_s": round(lat.size / span, 2) if span > 0 else None,
        "offered_frames_per_s": round(n_ev / (events[-1][0] - events[0][0]), 2) if n_ev > 1 else None,
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "max_ms": round(1e3 * float(lat.max()), 3),
        "steps": steps,
        "mean_batch": round(float(np.mean(batch)), 2),
        "gpu_busy_fraction": round(busy / span, 4) if span > 0 else None,
    }
