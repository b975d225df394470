"""Metric ego compensation of the checker (the library's opt-in DOGM_EGO_METRIC) against closed
forms: whole-cell translations with the remainder carried, the direction of the shift, rotation by
the yaw change about the grid centre for positions, velocities and grid fields, and translation along
a turned frame."""
import numpy as np

from oracle_ctypes import OracleDOGM
from parity_util import small_params


def _oracle(gs=20, res=0.5, n=4):
    p = small_params(gs=gs, res=res, n=n, b=1)
    o = OracleDOGM(p, seed=1, instance=0)
    o.set_ego_mode("metric")
    return o


def _put(o, xs, ys, vxs=None, vys=None):
    n = len(xs)
    st = np.zeros((n, 4), np.float32)
    st[:, 0], st[:, 1] = xs, ys
    if vxs is not None:
        st[:, 2], st[:, 3] = vxs, vys
    o.set_particles(state=st, grid_cell_idx=np.zeros(n, np.int32), weight=np.full(n, 0.25, np.float32),
                    associated=np.zeros(n, np.uint8))


def test_translation_carries_the_remainder():
    o = _oracle(res=0.5)
    _put(o, [10.0, 3.0, 7.5, 12.25], [10.0, 4.0, 1.0, 19.0])
    o.update_pose(0.0, 0.0, 0.0)
    xs = []
    for k in range(1, 11):
        o.update_pose(0.125 * k, 0.0, 0.0)        # 1/4 cell per frame along +x
        xs.append(float(o.get_particles()["state"][0, 0]))
    # content moves against the motion by whole cells: floor(k / 4) cells after frame k
    assert xs == [10.0 - (k // 4) for k in range(1, 11)]
    assert o.position()[0] == 1.0                 # 2 cells applied, 0.25 m of the motion pending


def test_rotation_quarter_turn():
    gs = 20
    o = _oracle(gs=gs)
    c = gs / 2
    _put(o, [c + 3.0, c, c + 0.5, c], [c, c + 2.0, c + 0.5, c - 4.0], [1.0, 0.0, 0.0, 2.0], [0.0, 1.0, 0.0, 0.0])
    cells = o.get_grid_cells()
    cells["occ_mass"][:] = 0
    cells["occ_mass"][int(c + 3) + gs * int(c)] = 0.5   # cell centre (c + 3.5, c + 0.5)
    o.set_grid_cells(cells)
    o.update_pose(0.0, 0.0, 0.0)
    o.update_pose(0.0, 0.0, np.float32(np.pi / 2))
    s = o.get_particles()["state"]
    # R(-90 deg): (dx, dy) -> (dy, -dx); velocities likewise
    np.testing.assert_allclose(s[:, 0], [c, c + 2.0, c + 0.5, c - 4.0], atol=1e-5)
    np.testing.assert_allclose(s[:, 1], [c - 3.0, c, c - 0.5, c], atol=1e-5)
    np.testing.assert_allclose(s[:, 2], [0.0, 1.0, 0.0, 0.0], atol=1e-6)
    np.testing.assert_allclose(s[:, 3], [-1.0, 0.0, 0.0, -2.0], atol=1e-6)
    occ = o.get_grid_cells()["occ_mass"].reshape(gs, gs)
    # the marked cell centre (3.5, 0.5) from c rotates to (0.5, -3.5): cell (c, c - 4)
    assert occ[int(c - 4), int(c)] == np.float32(0.5) and np.count_nonzero(occ) == 1


def test_rotation_round_trip_and_turned_translation():
    gs = 40
    o = _oracle(gs=gs, res=0.5)
    rng = np.random.default_rng(2)
    xs, ys = rng.uniform(5, 35, 4).astype(np.float32), rng.uniform(5, 35, 4).astype(np.float32)
    _put(o, xs, ys, [1.0, 0, -2, 3], [0.5, 1, 0, -1])
    o.update_pose(0.0, 0.0, 0.0)
    o.update_pose(0.0, 0.0, 0.3)
    o.update_pose(0.0, 0.0, 0.0)
    s = o.get_particles()["state"]
    np.testing.assert_allclose(s[:, 0], xs, atol=1e-4)
    np.testing.assert_allclose(s[:, 1], ys, atol=1e-4)
    # heading pi/2 (x axis of the grid along world +y): moving 1.6 m along world +y is 3.2 cells
    # forward, so the content moves 3 whole cells along -x of the grid (0.2 cells stay pending)
    o2 = _oracle(gs=gs, res=0.5)
    _put(o2, xs, ys)
    o2.update_pose(0.0, 0.0, np.float32(np.pi / 2))
    o2.update_pose(0.0, 1.6, np.float32(np.pi / 2))
    s2 = o2.get_particles()["state"]
    np.testing.assert_array_equal(s2[:, 0], xs - 3.0)
    np.testing.assert_array_equal(s2[:, 1], ys)
