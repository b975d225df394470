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
