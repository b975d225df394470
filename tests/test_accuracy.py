"""Accuracy harness (host code in the library): DBSCAN, cluster means and the MAE / RMSE evaluator
against the checker's restatement, closed forms, and scikit-learn's DBSCAN for the core points."""
import numpy as np
import pytest

import oracle_cluster as oc
from dogm_ctypes import VELOCITY_CELL_DTYPE, Evaluator, cluster_means, dbscan


def _cells(xy, v=None):
    c = np.zeros(len(xy), dtype=VELOCITY_CELL_DTYPE)
    c["x"] = [p[0] for p in xy]
    c["y"] = [p[1] for p in xy]
    if v is not None:
        c["data"]["mean_x_vel"] = [q[0] for q in v]
        c["data"]["mean_y_vel"] = [q[1] for q in v]
    return c


def _random_cells(rng, n, side, blobs=4):
    centres = rng.integers(0, side, size=(blobs, 2))
    pts = set()
    while len(pts) < n:
        if rng.random() < 0.8:
            c = centres[rng.integers(blobs)]
            p = (int(c[0] + rng.normal(0, 3)), int(c[1] + rng.normal(0, 2)))
        else:
            p = (int(rng.integers(0, side)), int(rng.integers(0, side)))
        pts.add(p)
    xy = sorted(pts, key=lambda p: (p[1], p[0]))  # row-major, as the extraction lists them
    v = rng.normal(0, 20, size=(len(xy), 2)).astype(np.float32)
    return _cells(xy, v)


def test_closed_forms():
    block = [(x, y) for y in range(5) for x in range(5)]
    far = [(x + 40, y + 10) for y in range(5) for x in range(5)]
    lone = [(20, 30), (70, 70)]
    c = _cells(sorted(block + far + lone, key=lambda p: (p[1], p[0])))
    ids, k = dbscan(c, 3.0, 5)
    assert k == 2
    xy = list(zip(c["x"].astype(int), c["y"].astype(int)))
    assert {xy[i] for i in range(len(xy)) if ids[i] == 0} == set(block)
    assert {xy[i] for i in range(len(xy)) if ids[i] == 1} == set(far)
    assert {xy[i] for i in range(len(xy)) if ids[i] == -1} == set(lone)
    # a chain of cells 2 apart: every cell has 3 neighbours within 2.0 (itself and two), so with
    # min_cells 3 the whole chain is one cluster and with 4 it is all noise
    chain = _cells([(2 * i, 0) for i in range(10)])
    assert dbscan(chain, 2.0, 3)[1] == 1
    ids4, k4 = dbscan(chain, 2.0, 4)
    assert k4 == 0 and np.all(ids4 == -1)
    m = cluster_means(c, ids, k)
    assert m["cells"].tolist() == [25, 25]
    assert (m["x"][0], m["y"][0]) == (2.0, 2.0) and (m["x"][1], m["y"][1]) == (42.0, 12.0)
    assert dbscan(_cells([]), 3.0, 5)[1] == 0


@pytest.mark.parametrize("seed,eps,min_cells", [(0, 3.0, 5), (1, 3.0, 5), (2, 1.5, 3), (3, 2.0, 4), (4, 4.0, 8)])
def test_dbscan_matches_checker(seed, eps, min_cells):
    rng = np.random.default_rng(seed)
    c = _random_cells(rng, 400, 120)
    ids, k = dbscan(c, eps, min_cells)
    oids, ok = oc.dbscan(c["x"], c["y"], eps, min_cells)
    assert k == ok and np.array_equal(ids, oids)
    got = cluster_means(c, ids, k)
    want = oc.cluster_means(c["x"], c["y"], c["data"]["mean_x_vel"], c["data"]["mean_y_vel"], oids, ok)
    for a, b in zip(got, want):
        assert (a["x"], a["y"], a["vx"], a["vy"], a["cells"]) == b


def test_core_points_match_scikit_learn():
    from sklearn.cluster import DBSCAN
    rng = np.random.default_rng(9)
    c = _random_cells(rng, 600, 150)
    ids, k = dbscan(c, 3.0, 5)
    X = np.stack([c["x"], c["y"]], 1).astype(np.float64)
    sk = DBSCAN(eps=3.0, min_samples=5).fit(X)
    core = np.zeros(len(c), bool)
    core[sk.core_sample_indices_] = True
    # noise agrees exactly; core points are partitioned identically up to the cluster numbering
    assert np.array_equal(ids == -1, sk.labels_ == -1)
    pairs = {(int(a), int(b)) for a, b in zip(ids[core], sk.labels_[core])}
    assert len(pairs) == len({a for a, _ in pairs}) == len({b for _, b in pairs}) == k


def test_evaluator_closed_form_and_checker():
    block = [(x, y) for y in range(10, 15) for x in range(20, 25)]       # mean cell (22, 12)
    c = _cells(block, [(30.0, -10.0)] * len(block))                      # 30, -10 cells/s
    ev = Evaluator(0.1)
    truth = np.array([(2.5, 1.6, 2.0, -1.0), (20.0, 20.0, 0.0, 0.0)],
                     dtype=[("x", "<f4"), ("y", "<f4"), ("vx", "<f4"), ("vy", "<f4")])
    cl = ev.add_step(c, truth)
    assert len(cl) == 1 and cl["cells"][0] == 25
    s = ev.summary()
    # mean (2.2, 1.2) m, (3.0, -1.0) m/s against (2.5, 1.6), (2.0, -1.0): errors 0.3, 0.4, 1.0, 0.0
    assert s["detections"] == 1 and s["unassigned"] == 0
    for k, v in zip(("x", "y", "vx", "vy"), (0.3, 0.4, 1.0, 0.0)):
        assert s["mae"][k] == pytest.approx(v, abs=1e-6) and s["rmse"][k] == pytest.approx(v, abs=1e-6)
    ev.add_step(c, truth[1:])                                            # nothing within 5 m
    assert ev.summary()["unassigned"] == 1
    # checker on random frames
    rng = np.random.default_rng(3)
    frames = []
    ev2 = Evaluator(0.2, 5.0, 3.0, 5)
    for _ in range(6):
        c = _random_cells(rng, 300, 100)
        truth = np.zeros(5, dtype=[("x", "<f4"), ("y", "<f4"), ("vx", "<f4"), ("vy", "<f4")])
        truth["x"] = rng.uniform(0, 20, 5)
        truth["y"] = rng.uniform(0, 20, 5)
        truth["vx"] = rng.normal(0, 3, 5)
        truth["vy"] = rng.normal(0, 3, 5)
        ev2.add_step(c, truth)
        frames.append((c["x"], c["y"], c["data"]["mean_x_vel"], c["data"]["mean_y_vel"],
                       [tuple(t) for t in truth]))
    mae, rmse, det, un = oc.evaluate(frames, 0.2)
    s = ev2.summary()
    assert (s["detections"], s["unassigned"]) == (det, un) and det > 0
    assert [s["mae"][k] for k in ("x", "y", "vx", "vy")] == [float(np.float32(v)) for v in mae]
    assert [s["rmse"][k] for k in ("x", "y", "vx", "vy")] == [float(np.float32(v)) for v in rmse]
