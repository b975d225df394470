"""Measurement-grid generation of the checker (the two-level lidar ray cast) against the independent
numpy generator of the benchmark scenes and against closed forms on single beams."""
import numpy as np
import pytest

import oracle_ctypes as oc
import scene as scene_mod
from parity_util import small_scene


def _grid(sc, frame):
    c = sc.cfg
    cs, sn = sc.beam_directions()
    return oc.lidar_grid(c.grid_size, c.resolution, c.max_range, c.occ_mass, c.free_mass, cs, sn, sc.ranges(frame))


@pytest.mark.parametrize("make", [lambda: small_scene(gs=48, res=0.5), lambda: small_scene(gs=96, res=0.25, seed=8),
                                  lambda: scene_mod.make_scene("plain_small"),
                                  lambda: scene_mod.make_scene("plain_full")])
def test_checker_matches_numpy_generator(make):
    sc = make()
    for f in (0, 3):
        got = _grid(sc, f)
        want = sc.measurement(f)
        assert got.tobytes() == want.tobytes(), f"frame {f}"
        assert np.count_nonzero(want["occ_mass"]) > 0 and np.count_nonzero(want["free_mass"]) > 0


def _one_beam(r, ang=0.0, gs=20, res=0.5, max_range=4.0):
    return oc.lidar_grid(gs, res, max_range, 0.7, 0.4, [np.cos(ang)], [np.sin(ang)], [r]).reshape(gs, gs)


def test_single_beam_closed_form():
    # Sensor at (5, 5) m = cell (10, 10); beam along +x, return at 2.3 m -> x = 7.3 m, cell 14.
    g = _one_beam(2.3)
    assert g["occ_mass"][10, 14] == np.float32(0.7) and g["free_mass"][10, 14] == 0
    # samples at 0, 0.25, ..., 2.25 m (< 2.3) cover cells 10..14; cell 14 is the return
    assert [x for x in range(20) if g["free_mass"][10, x] > 0] == [10, 11, 12, 13]
    assert np.count_nonzero(g["occ_mass"]) == 1 and np.count_nonzero(g["free_mass"]) == 4
    assert np.all(g["likelihood"] == 1) and np.all(g["p_A"] == 1)


def test_single_beam_no_return_and_limits():
    free_to_max = [x for x in range(20) if _one_beam(np.inf)["free_mass"][10, x] > 0]
    assert free_to_max == list(range(10, 18))                       # samples < 4 m: up to x = 8.75 m
    g = _one_beam(6.0)                                               # beyond max_range: no return
    assert np.count_nonzero(g["occ_mass"]) == 0
    assert [x for x in range(20) if g["free_mass"][10, x] > 0] == free_to_max
    g = _one_beam(4.0)                                               # exactly max_range: a return
    assert g["occ_mass"][10, 18] == np.float32(0.7)
    g = _one_beam(np.nan)                                            # NaN: nothing
    assert np.count_nonzero(g["occ_mass"]) == 0 and np.count_nonzero(g["free_mass"]) == 0
    g = _one_beam(1.0, ang=np.pi)                                    # towards -x: cells 10, 9, return in 8
    assert [x for x in range(20) if g["free_mass"][10, x] > 0] == [9, 10]
    assert g["occ_mass"][10, 8] == np.float32(0.7)


def test_return_outside_grid_and_occ_wins():
    # 3 m beam on a 2.5 m half-width grid: the return falls outside, samples stop at the edge.
    g = oc.lidar_grid(10, 0.5, 4.0, 0.7, 0.4, [1.0], [0.0], [3.0]).reshape(10, 10)
    assert np.count_nonzero(g["occ_mass"]) == 0
    assert [x for x in range(10) if g["free_mass"][5, x] > 0] == [5, 6, 7, 8, 9]
    # two beams: one passes through the cell where the other returns -> occupied, not free
    g = oc.lidar_grid(20, 0.5, 4.0, 0.7, 0.4, [1.0, 1.0], [0.0, 0.0], [1.2, 3.0]).reshape(20, 20)
    assert g["occ_mass"][10, 12] == np.float32(0.7) and g["free_mass"][10, 12] == 0
