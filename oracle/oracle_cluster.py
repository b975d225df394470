"""TEST INFRASTRUCTURE ONLY: the checker's restatement of the accuracy harness (DBSCAN, cluster
means, assignment and MAE / RMSE) that the library's dogm_dbscan / dogm_cluster_means /
dogm_evaluator_* are compared against.  Loaded by tests/ only.

Follows the reference demo's evaluation: dogm/demo/utils/dbscan.cpp:24-114 (DBSCAN with a
neighbourhood that includes the point, min_cells core threshold, sqrtf distance :19-22),
precision_evaluator.cpp:37-122 (cluster mean in cell units times the resolution, closest object
within 5 m), metrics.cpp:7-53 (MAE, RMSE).  Neighbourhoods are brute force over all points, in
input order; expansion is breadth-first.  The reference's expandCluster labels the first |seeds|
entries of its point array instead of the seeds (dbscan.cpp:62-65); the textbook labelling is
restated instead, as in the library.
"""
from __future__ import annotations

from collections import deque

import numpy as np

UNCLASSIFIED, NOISE = -2, -1


def region(xs, ys, eps, i):
    f = np.float32
    out = []
    for j in range(len(xs)):
        dx = f(xs[i]) - f(xs[j])
        dy = f(ys[i]) - f(ys[j])
        if np.sqrt(f(dx * dx + dy * dy), dtype=np.float32) <= f(eps):
            out.append(j)
    return out


def dbscan(xs, ys, eps: float, min_cells: int):
    n = len(xs)
    cid = [UNCLASSIFIED] * n
    k = 0
    for p in range(n):
        if cid[p] != UNCLASSIFIED:
            continue
        nb = region(xs, ys, eps, p)
        if len(nb) < min_cells:
            cid[p] = NOISE
            continue
        seeds = deque()
        for q in nb:
            if cid[q] == UNCLASSIFIED:
                seeds.append(q)
            if cid[q] in (UNCLASSIFIED, NOISE):
                cid[q] = k
        while seeds:
            q = seeds.popleft()
            nq = region(xs, ys, eps, q)
            if len(nq) < min_cells:
                continue
            for r in nq:
                if cid[r] == UNCLASSIFIED:
                    seeds.append(r)
                if cid[r] in (UNCLASSIFIED, NOISE):
                    cid[r] = k
        k += 1
    return np.array(cid, np.int32), k


def cluster_means(xs, ys, vxs, vys, cid, k):
    """f32 sums in input order divided by the size (precision_evaluator.cpp:98-110)."""
    f = np.float32
    out = []
    for c in range(k):
        sx = sy = svx = svy = f(0)
        m = 0
        for i in range(len(xs)):
            if cid[i] == c:
                sx = f(sx + f(xs[i]))
                sy = f(sy + f(ys[i]))
                svx = f(svx + f(vxs[i]))
                svy = f(svy + f(vys[i]))
                m += 1
        out.append((f(sx / f(m)), f(sy / f(m)), f(svx / f(m)), f(svy / f(m)), m))
    return out


def evaluate(frames, resolution: float, max_distance: float = 5.0, eps: float = 3.0, min_cells: int = 5):
    """frames: [(xs, ys, vxs, vys, truth [(x, y, vx, vy)])]; returns MAE, RMSE, detections, unassigned."""
    f = np.float32
    abs_sum = np.zeros(4)
    sq_sum = np.zeros(4)
    det = un = 0
    for xs, ys, vxs, vys, truth in frames:
        if len(xs) == 0 or len(truth) == 0:
            continue
        cid, k = dbscan(xs, ys, eps, min_cells)
        for mx, my, mvx, mvy, _ in cluster_means(xs, ys, vxs, vys, cid, k):
            m = [f(mx * f(resolution)), f(my * f(resolution)), f(mvx * f(resolution)), f(mvy * f(resolution))]
            best, best_d = -1, None
            for o, t in enumerate(truth):
                d = np.sqrt(f(f(m[0] - f(t[0])) ** 2 + f(m[1] - f(t[1])) ** 2), dtype=np.float32)
                if d < f(max_distance) and (best < 0 or d < best_d):
                    best, best_d = o, d
            if best < 0:
                un += 1
                continue
            err = np.array([float(f(m[a] - f(truth[best][a]))) for a in range(4)])
            abs_sum += np.abs(err)
            sq_sum += err * err
            det += 1
    return abs_sum / det, np.sqrt(sq_sum / det), det, un
