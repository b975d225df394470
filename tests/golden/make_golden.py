"""Writes tests/golden/small_3frames.npz: inputs (parameters, measurement grids, poses) and the
checker's outputs after each of 3 updateGrid frames of a 16x16 scene with 64 particles and 8
births.  The reference ships no golden vectors; these pin the checker (and, through the GPU
tests, the HIP path) to fixed bits over time.  Regenerate only after an intended change of the
restated algorithm:

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path[:0] = [str(REPO / "dynamic-occupancy-grid-map_amd"), str(REPO / "oracle"), str(REPO / "tests")]

from oracle_ctypes import GRID_FIELDS, OracleDOGM  # noqa: E402
from parity_util import small_params, small_scene  # noqa: E402

PARAMS = dict(gs=16, res=0.5, n=64, b=8)
SEED = 2024
FRAMES = 3


def scenario():
    p = small_params(**PARAMS)
    sc = small_scene(gs=16, res=0.5, n_boxes=2, n_beams=64, ego_speed=5.0, yaw_rate=0.3, seed=11)
    meas = [sc.measurement(f) for f in range(FRAMES)]
    poses = [sc.ego_pose(f) for f in range(FRAMES)]
    return p, meas, poses


def run(dogm, meas, poses):
    out = {}
    for f in range(FRAMES):
        x, y, yaw = poses[f]
        dogm.update_grid(meas[f], x, y, yaw, 0.1)
        parts = dogm.get_particles()
        births = dogm.get_birth_particles()
        cells = dogm.get_grid_cells()
        out[f"f{f}_state"] = np.asarray(parts["state"], np.float32)
        out[f"f{f}_idx"] = np.asarray(parts["grid_cell_idx"], np.int32)
        out[f"f{f}_weight"] = np.asarray(parts["weight"], np.float32)
        out[f"f{f}_assoc"] = np.asarray(parts["associated"], np.uint8)
        out[f"f{f}_birth_state"] = np.asarray(births["state"], np.float32)
        out[f"f{f}_birth_weight"] = np.asarray(births["weight"], np.float32)
        out[f"f{f}_ancestors"] = np.asarray(dogm.get_resampled_indices(), np.int32)
        for k in ["start_idx", "end_idx"] + GRID_FIELDS:
            out[f"f{f}_grid_{k}"] = np.asarray(cells[k], np.int32 if k.endswith("idx") else np.float32)
    return out


def main():
    p, meas, poses = scenario()
    out = run(OracleDOGM(p, seed=SEED), meas, poses)
    np.savez_compressed(HERE / "small_3frames.npz", meas=np.stack(meas), poses=np.array(poses, np.float64), **out)
    print(f"wrote {HERE / 'small_3frames.npz'} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
