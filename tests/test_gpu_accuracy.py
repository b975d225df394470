"""Accuracy harness on the GPU path: extraction + DBSCAN + metrics over a scene with ground truth.
A batched vehicle scores exactly like the plain handle with the same seed and instance."""
import sys

import pytest

import accuracy

pytestmark = pytest.mark.gpu


def test_batched_scores_like_plain():
    print("accuracy plain vs batched", flush=True)
    sys.stdout.flush()
    a = accuracy.run("plain_small", frames=12, skip=4, form="plain")
    b = accuracy.run("plain_small", frames=12, skip=4, form="batched", vehicles=1)
    for k in ("mae", "rmse", "detections", "unassigned"):
        assert a[k] == b[k], k
    assert [(r["frame"], r["cells"], r["clusters"]) for r in a["per_frame"]] == \
        [(r["frame"], r["cells"], r["clusters"]) for r in b["per_frame"]]
    assert a["detections"] > 0


def test_tracking_quality_bounds():
    """The moving boxes are found and their velocity estimated: a regression guard on the whole
    path (bounds several times the values measured on this scene)."""
    print("accuracy plain_small 20 frames", flush=True)
    r = accuracy.run("plain_small", frames=20, skip=5)
    assert r["detections"] >= 20
    assert r["mae"]["vx"] < 4.0 and r["mae"]["vy"] < 4.0
    assert r["mae"]["x"] < 3.0 and r["mae"]["y"] < 3.0


def test_serving_scores_like_batched():
    """Frames submitted to the server as scans (grids generated on the device) and the velocity
    cells extracted inside the steps score exactly like the batched form fed host-generated grids."""
    print("accuracy batched vs serving", flush=True)
    sys.stdout.flush()
    a = accuracy.run("plain_small", frames=12, skip=4, form="batched", vehicles=2)
    b = accuracy.run("plain_small", frames=12, skip=4, form="serving", vehicles=2)
    for k in ("mae", "rmse", "detections", "unassigned"):
        assert a[k] == b[k], k
    assert [(r["frame"], r["cells"], r["clusters"]) for r in a["per_frame"]] == \
        [(r["frame"], r["cells"], r["clusters"]) for r in b["per_frame"]]
    assert a["detections"] > 0


@pytest.mark.parametrize("workload,frames,ego", [("plain_small", 16, "reference"), ("plain_full", 10, "metric")])
def test_harness_on_gpu_cells_matches_checker(workload, frames, ego):
    """The cells the GPU extracts each frame (plain handle) are clustered and scored by the library's
    harness (dogm_dbscan, dogm_cluster_means, dogm_evaluator_*) and by the checker's restatement
    (oracle/oracle_cluster.py, after dbscan.cpp:24-114 and precision_evaluator.cpp:37-122): the same
    partition and cluster ids every frame, and the same detections, MAE and RMSE."""
    import numpy as np

    import oracle_cluster as oc
    import scene as scene_mod
    from dogm_ctypes import DOGM, Evaluator, Params, dbscan

    print(f"accuracy harness vs checker on {workload}", flush=True)
    w = scene_mod.WORKLOADS[workload]
    p = Params(**w["params"])
    sc = scene_mod.make_scene(workload, seed=1)
    ev = Evaluator(p.resolution)
    g = DOGM(p, seed=123456)
    g.set_ego_mode(ego)
    chk_frames = []
    partitions = 0
    try:
        for f in range(frames):
            x, y, yaw = sc.ego_pose(f)
            g.update_grid(sc.measurement(f), x, y, yaw if ego == "metric" else 0.0, float(w["dt"]))
            if f < 4:
                continue
            cells, _ = g.get_cells_with_velocity(accuracy.MIN_OCC, accuracy.MIN_VEL)
            truth = sc.objects(f)
            ev.add_step(cells, truth)
            xs, ys = cells["x"], cells["y"]
            vxs, vys = cells["data"]["mean_x_vel"], cells["data"]["mean_y_vel"]
            if cells.size:
                ids, k = dbscan(cells, 3.0, 5)
                cid, kc = oc.dbscan(xs, ys, 3.0, 5)
                assert k == kc and np.array_equal(ids, cid), f"frame {f}: DBSCAN partition"
                partitions += int(k > 0)
            chk_frames.append((xs, ys, vxs, vys, [tuple(t) for t in truth]))
    finally:
        g.close()
    mae, rmse, det, un = oc.evaluate(chk_frames, p.resolution)
    s = ev.summary()
    assert partitions > 0 and det > 0
    assert (s["detections"], s["unassigned"]) == (det, un)
    assert [s["mae"][k] for k in ("x", "y", "vx", "vy")] == [float(np.float32(v)) for v in mae]
    assert [s["rmse"][k] for k in ("x", "y", "vx", "vy")] == [float(np.float32(v)) for v in rmse]
