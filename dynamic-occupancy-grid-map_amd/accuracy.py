"""Tracking accuracy of the filter against the scenes' ground truth (the reference demo's
evaluation loop, dogm/demo/main.cpp:82-101 with utils/precision_evaluator.cpp:37-122).

Per frame: updateGrid, the cells passing the velocity test (occupancy >= 0.7, v^T S^-1 v >= 4,
main.cpp:51-52, extracted on the device), DBSCAN (eps 3 cells, 5 cells) and the cluster means in
metres, each assigned to the closest ground-truth box within 5 m; MAE and RMSE of position and
velocity over all assigned clusters.  Frames before `skip` (the filter's warm-up) are not scored.

    python dynamic-occupancy-grid-map_amd/accuracy.py [--workload plain_small] [--frames 20]
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

import scene as scene_mod
from dogm_ctypes import DOGM, EXTRACT_VELOCITY, Batch, Evaluator, Lidar, Params, Server

MIN_OCC, MIN_VEL = 0.7, 4.0


def run(workload: str = "plain_small", frames: int | None = None, skip: int = 5, seed: int = 1,
        form: str = "plain", cdf_mode: str = "exact", vehicles: int = 1, ego_mode: str = "reference",
        sensor_frame: str = "world") -> dict:
    """form "plain": one handle; "batched": `vehicles` vehicles of the same scene (instances 0..)
    stepped together, all scored; "serving": the same vehicles admitted to a server, every frame
    submitted as a lidar scan (the grid is generated on the device inside the step) and the cells
    with velocity extracted inside the step (exact CDF and the reference's ego compensation).  sensor_frame "vehicle" turns the lidar and its grid with the ego
    heading (the yaw is then passed to the filter; with a world-aligned sensor the filter gets yaw 0,
    as the reference demo passes, main.cpp:91)."""
    w = scene_mod.WORKLOADS[workload]
    p = Params(**w["params"])
    sc = scene_mod.make_scene(workload, seed=seed, sensor_frame=sensor_frame)

    def pose(f):
        x, y, yaw = sc.ego_pose(f)
        return x, y, (yaw if sensor_frame == "vehicle" else 0.0)

    frames = frames or w["frames"]
    dt = float(w["dt"])
    ev = Evaluator(p.resolution)
    per_frame = []
    if form == "plain":
        g = DOGM(p, seed=123456)
        g.set_cdf_mode(cdf_mode)
        g.set_ego_mode(ego_mode)
        for f in range(frames):
            x, y, yaw = pose(f)
            g.update_grid(sc.measurement(f), x, y, yaw, dt)
            if f >= skip:
                cells, _ = g.get_cells_with_velocity(MIN_OCC, MIN_VEL)
                cl = ev.add_step(cells, sc.objects(f))
                per_frame.append({"frame": f, "cells": int(cells.size), "clusters": int(cl.size)})
        g.close()
    elif form == "serving":
        if cdf_mode != "exact" or ego_mode != "reference":
            raise ValueError("the serving form runs the exact CDF and the reference ego compensation")
        srv = Server(hbm_budget_bytes=256 << 30)
        srv.set_extraction(EXTRACT_VELOCITY, MIN_OCC, MIN_VEL)  # before any admission
        ids = [srv.admit(p, 123456, i) for i in range(vehicles)]
        assert all(v is not None for v in ids)
        cs, sn = sc.beam_directions()
        with Lidar(**sc.lidar_params(), beam_cos=cs, beam_sin=sn) as lidar:
            for f in range(frames):
                t = f * dt
                r = sc.ranges(f)
                for vid in ids:
                    srv.submit_scan(vid, lidar, r, pose(f), dt, t_arrival=t)
                while srv.queued():
                    srv.step(t)
                if f >= skip:
                    for vid in ids:
                        cells, _ = srv.get_cells_with_velocity(vid)
                        cl = ev.add_step(cells, sc.objects(f))
                        per_frame.append({"frame": f, "vehicle": vid, "cells": int(cells.size),
                                          "clusters": int(cl.size)})
        srv.close()
    else:
        b = Batch()
        ids = [b.add_vehicle(p, 123456, i) for i in range(vehicles)]
        b.set_extraction(EXTRACT_VELOCITY, MIN_OCC, MIN_VEL)  # inside every step, for all vehicles
        for vid in ids:
            b.set_cdf_mode(vid, cdf_mode)
            b.set_ego_mode(vid, ego_mode)
        for f in range(frames):
            m = sc.measurement(f)
            b.update(ids, [m] * vehicles, [pose(f)] * vehicles, dt)
            if f >= skip:
                for vid in ids:
                    cells, _ = b.get_cells_with_velocity(vid)
                    cl = ev.add_step(cells, sc.objects(f))
                    per_frame.append({"frame": f, "vehicle": vid, "cells": int(cells.size), "clusters": int(cl.size)})
        b.close()
    s = ev.summary()
    ev.close()
    s.update({"workload": workload, "form": form, "cdf_mode": cdf_mode, "ego_mode": ego_mode,
              "sensor_frame": sensor_frame, "frames": frames, "scored_from": skip, "per_frame": per_frame})
    return s


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="plain_small", choices=sorted(scene_mod.WORKLOADS))
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--skip", type=int, default=5)
    ap.add_argument("--form", default="plain", choices=["plain", "batched", "serving"])
    ap.add_argument("--cdf-mode", default="exact", choices=["exact", "fp64", "fp32c", "fp32"])
    ap.add_argument("--vehicles", type=int, default=1)
    ap.add_argument("--ego-mode", default="reference", choices=["reference", "metric"])
    ap.add_argument("--sensor-frame", default="world", choices=["world", "vehicle"])
    a = ap.parse_args()
    r = run(a.workload, a.frames, a.skip, form=a.form, cdf_mode=a.cdf_mode, vehicles=a.vehicles,
            ego_mode=a.ego_mode, sensor_frame=a.sensor_frame)
    print(json.dumps(r))
    sys.exit(0)
