"""The checker's OpenMP loops: results do not depend on the thread count.

Each case runs the checker in a fresh process with OMP_NUM_THREADS fixed (OpenMP reads it once at
start-up) and digests every observable array after three frames with a grid shift every frame, in the
default reference ego mode and in the metric mode (its rotation path), with multinomial and
systematic draws.
"""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

from conftest import REPO

SCRIPT = r"""
import hashlib, sys
sys.path[:0] = [sys.argv[1] + "/dynamic-occupancy-grid-map_amd", sys.argv[1] + "/oracle", sys.argv[1] + "/tests"]
import numpy as np
import oracle_ctypes
from oracle_ctypes import OracleDOGM
from parity_util import small_params, small_scene
ego, draw = sys.argv[2], sys.argv[3]
p = small_params(gs=64, res=0.5, n=20000, b=2000)
sc = small_scene(gs=64, res=0.5)
o = OracleDOGM(p, seed=11, instance=2)
o.set_ego_mode(ego)
o.set_draw_mode(draw)
h = hashlib.sha256()
for f in range(3):
    x, y, yaw = sc.ego_pose(f)
    o.update_grid(sc.measurement(f), x, y, yaw, 0.1)
    for k, v in sorted(o.get_particles().items()):
        h.update(k.encode()); h.update(np.ascontiguousarray(v).tobytes())
    for k, v in sorted(o.get_grid_cells().items()):
        h.update(k.encode()); h.update(np.ascontiguousarray(v).tobytes())
    h.update(np.ascontiguousarray(o.get_resampled_indices()).tobytes())
print(oracle_ctypes.threads(), h.hexdigest())
"""


def _digest(threads: int, ego: str, draw: str) -> tuple[int, str]:
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", SCRIPT, str(REPO), ego, draw], env=env, check=True,
                         capture_output=True, text=True, timeout=300).stdout.split()
    return int(out[0]), out[1]


@pytest.mark.parametrize("ego,draw", [("reference", "multinomial"), ("metric", "systematic")])
def test_checker_results_do_not_depend_on_thread_count(ego, draw):
    n1, d1 = _digest(1, ego, draw)
    n5, d5 = _digest(5, ego, draw)
    assert (n1, n5) == (1, 5)
    assert d1 == d5
