"""Synthetic measurement grids for the filter workloads (the step before updateGrid).

The reference renders a polar inverse-sensor-model texture into a Cartesian grid with OpenGL
(dogm/demo/simulator/mapping/laser_to_meas_grid.cu:25-70).  This generator ray-casts a 2-D lidar
against moving oriented boxes on the host instead and applies a two-level inverse sensor model:
occupied mass ``occ`` at the cell of the first hit, free mass ``free`` on the cells a beam crosses
before it, unknown (0, 0) elsewhere.  Likelihood and p_A are 1 as in
measurement_grid.cu:129-130.  Grids are ego-centred (the lidar sits at the grid centre) and
row-major with index = x + grid_size * y.  Every draw comes from a seeded numpy Generator, so a
scene is reproducible from its seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MEAS_CELL_DTYPE = np.dtype([("free_mass", "<f4"), ("occ_mass", "<f4"), ("likelihood", "<f4"), ("p_A", "<f4")])


@dataclass
class SceneConfig:
    grid_size: int = 250
    resolution: float = 0.2
    n_boxes: int = 8
    box_length: float = 4.5
    box_width: float = 1.8
    max_box_speed: float = 10.0
    n_beams: int = 360
    max_range: float = 50.0
    occ_mass: float = 0.7
    free_mass: float = 0.4
    ego_speed: float = 0.0
    ego_yaw_rate: float = 0.0
    dt: float = 0.1
    seed: int = 1
    # "world": grid axes along the world axes (beams at fixed world angles); "vehicle": the sensor and
    # its grid turn with the ego heading (beam b at 2 pi b / n_beams from the heading)
    sensor_frame: str = "world"


class Scene:
    def __init__(self, cfg: SceneConfig):
        self.cfg = cfg
        rng = np.random.default_rng(cfg.seed)
        half = 0.5 * cfg.grid_size * cfg.resolution
        # Box centres inside the initial field of view, away from the sensor.
        n = cfg.n_boxes
        pos = rng.uniform(-0.85 * half, 0.85 * half, size=(n, 2))
        near = np.linalg.norm(pos, axis=1) < 3.0
        pos[near] += 3.0 * np.sign(pos[near] + 1e-9)
        self.box_pos0 = pos
        self.box_vel = rng.uniform(-cfg.max_box_speed, cfg.max_box_speed, size=(n, 2))
        heading = np.arctan2(self.box_vel[:, 1], self.box_vel[:, 0])
        self.box_heading = heading
        self.ego0 = np.zeros(2)

    def ego_pose(self, frame: int) -> tuple[float, float, float]:
        """Unicycle ego trajectory: speed v along the heading, heading turning at the yaw rate."""
        c = self.cfg
        t = frame * c.dt
        w, v = c.ego_yaw_rate, c.ego_speed
        if abs(w) < 1e-12:
            return float(v * t), 0.0, 0.0
        yaw = w * t
        x = v / w * np.sin(yaw)
        y = v / w * (1.0 - np.cos(yaw))
        return float(x), float(y), float(yaw)

    def _box_edges(self, frame: int, ego_xy: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Edge segments (start, direction) of every box in sensor-centred coordinates."""
        c = self.cfg
        t = frame * c.dt
        centre = self.box_pos0 + self.box_vel * t - ego_xy[None, :]
        ch, sh = np.cos(self.box_heading), np.sin(self.box_heading)
        hl, hw = 0.5 * c.box_length, 0.5 * c.box_width
        ax = np.stack([ch, sh], 1)
        ay = np.stack([-sh, ch], 1)
        corners = np.stack([
            centre + hl * ax + hw * ay, centre - hl * ax + hw * ay,
            centre - hl * ax - hw * ay, centre + hl * ax - hw * ay], 1)  # (n, 4, 2)
        starts = corners.reshape(-1, 2)
        ends = np.roll(corners, -1, axis=1).reshape(-1, 2)
        return starts, ends - starts

    def objects(self, frame: int) -> np.ndarray:
        """Ground truth of the boxes inside the grid at `frame`: centre in grid metres (the sensor
        at 0.5 * grid_size * resolution) and world velocity [m/s], as (x, y, vx, vy) float32 rows."""
        c = self.cfg
        ex, ey, yaw = self.ego_pose(frame)
        half = 0.5 * c.grid_size * c.resolution
        rel = self.box_pos0 + self.box_vel * (frame * c.dt) - np.array([ex, ey])[None, :]
        vel = self.box_vel
        if c.sensor_frame == "vehicle":
            cy, sy = np.cos(yaw), np.sin(yaw)
            rot = np.array([[cy, sy], [-sy, cy]])  # world -> vehicle axes
            rel = rel @ rot.T
            vel = vel @ rot.T
        centre = rel + half
        inside = np.all((centre >= 0.0) & (centre < 2.0 * half), axis=1)
        out = np.zeros(int(inside.sum()), dtype=[("x", "<f4"), ("y", "<f4"), ("vx", "<f4"), ("vy", "<f4")])
        out["x"], out["y"] = centre[inside, 0], centre[inside, 1]
        out["vx"], out["vy"] = vel[inside, 0], vel[inside, 1]
        return out

    def beam_directions(self) -> tuple[np.ndarray, np.ndarray]:
        """(cos, sin) of every beam angle 2 pi b / n_beams, as the generators here use them."""
        ang = 2.0 * np.pi * np.arange(self.cfg.n_beams) / self.cfg.n_beams
        return np.cos(ang), np.sin(ang)

    def lidar_params(self) -> dict:
        """Arguments of the device generator (dogm_lidar_params) for this scene's sensor."""
        c = self.cfg
        return dict(grid_size=c.grid_size, beam_count=c.n_beams, resolution=c.resolution, max_range=c.max_range,
                    occ_mass=c.occ_mass, free_mass=c.free_mass)

    def ranges(self, frame: int) -> np.ndarray:
        """Distance to the first hit per beam (inf when nothing is hit within max_range)."""
        c = self.cfg
        ex, ey, yaw = self.ego_pose(frame)
        p, e = self._box_edges(frame, np.array([ex, ey]) + self.ego0)
        ang = 2.0 * np.pi * np.arange(c.n_beams) / c.n_beams
        if c.sensor_frame == "vehicle":
            ang = ang + yaw  # world directions of the beams
        d = np.stack([np.cos(ang), np.sin(ang)], 1)  # (b, 2)
        # Solve t d = p + s e for every beam/edge pair.
        den = d[:, None, 0] * e[None, :, 1] - d[:, None, 1] * e[None, :, 0]
        with np.errstate(divide="ignore", invalid="ignore"):
            tt = (p[None, :, 0] * e[None, :, 1] - p[None, :, 1] * e[None, :, 0]) / den
            ss = (p[None, :, 0] * d[:, None, 1] - p[None, :, 1] * d[:, None, 0]) / den
        ok = (np.abs(den) > 1e-12) & (tt > 0) & (ss >= 0) & (ss <= 1)
        tt = np.where(ok, tt, np.inf)
        r = tt.min(axis=1)
        r[r > c.max_range] = np.inf
        return r

    def measurement(self, frame: int) -> np.ndarray:
        c = self.cfg
        gs, res = c.grid_size, c.resolution
        r = self.ranges(frame)
        ang = 2.0 * np.pi * np.arange(c.n_beams) / c.n_beams
        centre = 0.5 * gs * res
        step = 0.5 * res
        n_steps = int(np.ceil(c.max_range / step)) + 1
        dist = np.arange(n_steps) * step
        lim = np.minimum(r, c.max_range)
        free_pts = dist[None, :] < lim[:, None]
        xs = centre + np.cos(ang)[:, None] * dist[None, :]
        ys = centre + np.sin(ang)[:, None] * dist[None, :]
        ix = np.floor(xs / res).astype(np.int64)
        iy = np.floor(ys / res).astype(np.int64)
        inside = (ix >= 0) & (ix < gs) & (iy >= 0) & (iy < gs)
        free_mask = np.zeros(gs * gs, bool)
        sel = free_pts & inside
        free_mask[ix[sel] + gs * iy[sel]] = True
        hit = np.isfinite(r)
        hx = np.floor((centre + np.cos(ang[hit]) * r[hit]) / res).astype(np.int64)
        hy = np.floor((centre + np.sin(ang[hit]) * r[hit]) / res).astype(np.int64)
        hin = (hx >= 0) & (hx < gs) & (hy >= 0) & (hy < gs)
        occ_mask = np.zeros(gs * gs, bool)
        occ_mask[hx[hin] + gs * hy[hin]] = True
        m = np.zeros(gs * gs, dtype=MEAS_CELL_DTYPE)
        m["free_mass"][free_mask & ~occ_mask] = np.float32(c.free_mass)
        m["occ_mass"][occ_mask] = np.float32(c.occ_mass)
        m["likelihood"] = 1.0
        m["p_A"] = 1.0
        return m


# Workloads of the benchmark configurations (grid sizes, particle counts and scenes).
WORKLOADS = {
    "plain_small": dict(
        params=dict(size=50.0, resolution=0.2, particle_count=300_000, new_born_particle_count=30_000,
                    persistence_prob=0.99, stddev_process_noise_position=0.1, stddev_process_noise_velocity=2.0,
                    birth_prob=0.02, stddev_velocity=30.0, init_max_velocity=15.0, freespace_discount=0.01),
        scene=dict(grid_size=250, resolution=0.2, n_boxes=8, n_beams=360, max_range=50.0),
        frames=20, dt=0.1),
    "plain_full": dict(
        params=dict(size=120.0, resolution=0.1, particle_count=2_000_000, new_born_particle_count=200_000,
                    persistence_prob=0.99, stddev_process_noise_position=0.1, stddev_process_noise_velocity=2.0,
                    birth_prob=0.02, stddev_velocity=30.0, init_max_velocity=15.0, freespace_discount=0.01),
        scene=dict(grid_size=1200, resolution=0.1, n_boxes=40, n_beams=1080, max_range=60.0, ego_speed=10.0,
                   ego_yaw_rate=0.1),
        frames=50, dt=0.1),
}


def make_scene(workload: str, seed: int = 1, sensor_frame: str = "world") -> Scene:
    w = WORKLOADS[workload]
    return Scene(SceneConfig(seed=seed, dt=w["dt"], sensor_frame=sensor_frame, **w["scene"]))
