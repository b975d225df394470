"""dogm_update_grid_noalloc (the update with its step buffers in a static device array) against
dogm_update_grid on twin handles, bit for bit, through the C-ABI.

Two handles of different size are stepped back to back through the static array (the second call
on a smaller grid than the first, so it runs over what the larger step left there), the large
one again afterwards, and calls of both entry points are mixed on one handle.  A grid whose step
needs more than the array holds is rejected with a clear error.
"""
import ctypes as C

import numpy as np
import pytest

from dogm_ctypes import DOGM, DOGM_ERR_INVALID_ARGUMENT, DOGM_OK, MEAS_CELL_DTYPE, DeviceBuffer, lib
from parity_util import assert_int_equal, assert_vehicle_equal, small_params, small_scene

pytestmark = pytest.mark.gpu

SEED = 99


def _noalloc(h: DOGM, meas, pose, dt=0.1):
    x, y, yaw = pose
    if isinstance(meas, DeviceBuffer):
        st = lib.dogm_update_grid_noalloc(h.h, C.c_void_p(meas.ptr), x, y, yaw, dt, 1)
    else:
        m = np.ascontiguousarray(meas, dtype=MEAS_CELL_DTYPE)
        st = lib.dogm_update_grid_noalloc(h.h, m.ctypes.data_as(C.c_void_p), x, y, yaw, dt, 0)
    assert st == DOGM_OK, lib.dogm_last_error(h.h).decode()


def _same(tag, a: DOGM, b: DOGM):
    assert_vehicle_equal(tag, a.get_particles(), b.get_particles(), a.get_grid_cells(), b.get_grid_cells())
    assert_int_equal(f"{tag} ancestors", a.get_resampled_indices(), b.get_resampled_indices())
    assert_int_equal(f"{tag} weight_array", a.get_weight_array().view(np.uint32),
                     b.get_weight_array().view(np.uint32))


def test_noalloc_matches_update_grid_large_then_small():
    big_p, big_sc = small_params(gs=128, res=0.5, n=60000, b=6000), small_scene(gs=128, res=0.5, seed=5)
    small_p, small_sc = small_params(gs=48, res=0.5, n=4000, b=400), small_scene(gs=48, res=0.5, seed=6)
    big, big_ref = DOGM(big_p, seed=SEED, instance=1), DOGM(big_p, seed=SEED, instance=1)
    small, small_ref = DOGM(small_p, seed=SEED, instance=2), DOGM(small_p, seed=SEED, instance=2)
    dbuf = DeviceBuffer(128 * 128 * 16)
    try:
        for f in range(4):
            print(f"no-alloc frame {f}", flush=True)
            mb, pb = big_sc.measurement(f), big_sc.ego_pose(f)
            if f % 2:  # device grids on odd frames (copied by the step's occupancy update)
                dbuf.upload(mb)
                _noalloc(big, dbuf, pb)
                big_ref.update_grid(dbuf, *pb, 0.1)
            else:
                _noalloc(big, mb, pb)
                big_ref.update_grid(mb, *pb, 0.1)
            _same(f"frame {f} large grid", big, big_ref)
            ms, ps = small_sc.measurement(f), small_sc.ego_pose(f)
            _noalloc(small, ms, ps)  # smaller step over the array the large one just used
            small_ref.update_grid(ms, *ps, 0.1)
            _same(f"frame {f} small grid", small, small_ref)
        # both entry points on one handle: the regular step rebuilds its own tables afterwards
        for f in range(4, 7):
            mb, pb = big_sc.measurement(f), big_sc.ego_pose(f)
            if f == 5:
                big.update_grid(mb, *pb, 0.1)
            else:
                _noalloc(big, mb, pb)
            big_ref.update_grid(mb, *pb, 0.1)
            _same(f"frame {f} mixed entry points", big, big_ref)
    finally:
        dbuf.free()
        for h in (big, big_ref, small, small_ref):
            h.close()


def test_noalloc_rejects_a_step_beyond_the_static_array():
    # 1.6e7 particles on a small grid: the step needs more than 16 MiB of step buffers.
    p = small_params(gs=64, res=0.5, n=16_000_000, b=1000)
    sc = small_scene(gs=64, res=0.5, seed=7)
    h = DOGM(p, seed=SEED)
    try:
        m = np.ascontiguousarray(sc.measurement(0), dtype=MEAS_CELL_DTYPE)
        x, y, yaw = sc.ego_pose(0)
        st = lib.dogm_update_grid_noalloc(h.h, m.ctypes.data_as(C.c_void_p), x, y, yaw, 0.1, 0)
        assert st == DOGM_ERR_INVALID_ARGUMENT
        msg = lib.dogm_last_error(h.h).decode()
        assert "static scratch" in msg and "bytes" in msg, msg
    finally:
        h.close()
