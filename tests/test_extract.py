"""Result extraction of the checker (dogm/demo/utils/image_creation.cpp:19-66, 80-83, 123-125) against
closed forms: pignistic occupancy on dyadic masses, the 8-bit image, v^T S^-1 v on matrices with known
inverses, singular covariances, inclusive thresholds and the demo's row-major order.

The reference evaluates the velocity test with OpenCV (cv::Mat::inv, matrix products), which is not
installed here and has no golden vectors in the reference's tests; its exact f32 rounding is
therefore "parity unpinned" beyond what these closed forms fix (exact results on representable
inputs, ~1e-6 relative agreement with an f64 evaluation elsewhere)."""
from fractions import Fraction

import numpy as np
import pytest

import oracle_ctypes as oc
from dogm_ctypes import GRID_CELL_DTYPE


def test_pignistic_dyadic_masses_are_exact():
    occ, free, want = [], [], []
    for a in range(0, 65):
        for b in range(0, 65 - a):
            occ.append(a / 64)
            free.append(b / 64)
            want.append(Fraction(a, 64) + Fraction(1, 2) * (1 - Fraction(a, 64) - Fraction(b, 64)))
    p, img = oc.occupancy(np.array(occ), np.array(free))
    assert all(Fraction(float(x)) == w for x, w in zip(p, want))
    # 255 - floor(255 p), with 255 p exact in f32 for these p (at most 15 significant bits).
    want_img = [255 - int((255 * w).numerator // (255 * w).denominator) for w in want]
    assert img.tolist() == want_img


def test_image_extremes_and_clamp():
    occ = np.array([1.0, 0.0, 0.0, 1.5, 0.0, np.nan], np.float32)
    free = np.array([0.0, 1.0, 0.0, 0.0, 1.5, 0.0], np.float32)
    p, img = oc.occupancy(occ, free)
    assert p[:5].tolist() == [1.0, 0.0, 0.5, 1.25, -0.25]
    # 1 -> 0 (black), 0 -> 255, unknown 0.5 -> 255 - 127; out of range clamps; NaN -> 255.
    assert img.tolist() == [0, 255, 128, 0, 255, 255]


@pytest.mark.parametrize("v,s,want", [
    ((2.0, 8.0), (4.0, 16.0, 0.0), 5.0),          # diagonal: vx^2/sxx + vy^2/syy
    ((1.0, -1.0), (2.0, 2.0, 1.0), 2.0),          # S^-1 = [[2, -1], [-1, 2]] / 3
    ((3.0, 0.0), (1.0, 1.0, 0.0), 9.0),
    ((0.0, 0.0), (1.0, 1.0, 0.5), 0.0),
    ((5.0, 7.0), (0.0, 0.0, 0.0), 0.0),           # singular: OpenCV returns a zero inverse
    ((5.0, 7.0), (1.0, 1.0, 1.0), 0.0),           # singular, non-zero
])
def test_velocity_closed_forms(v, s, want):
    assert oc.velocity_normalized(v[0], v[1], s[0], s[1], s[2]) == want


def test_velocity_off_diagonal_sign():
    # v = (1, 1) with S = [[2, 1], [1, 2]]: 2/3; a sign error in the off-diagonal would give 2.
    m = oc.velocity_normalized(1.0, 1.0, 2.0, 2.0, 1.0)
    assert abs(m - 2.0 / 3.0) < 1e-6


def test_velocity_random_spd_against_f64():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        a = rng.normal(size=(2, 2))
        s = a @ a.T + 0.1 * np.eye(2)
        s = s.astype(np.float32).astype(np.float64)
        v = rng.normal(scale=5.0, size=2).astype(np.float32).astype(np.float64)
        want = float(v @ np.linalg.inv(s) @ v)
        got = oc.velocity_normalized(v[0], v[1], s[0, 0], s[1, 1], s[0, 1])
        cond = np.linalg.cond(s)
        assert abs(got - want) <= 1e-6 * cond * max(1.0, abs(want)), (got, want, cond)


def _grid(gs):
    cells = np.zeros(gs * gs, dtype=GRID_CELL_DTYPE)
    cells["free_mass"] = 0.25
    cells["occ_mass"] = 0.75            # pignistic 0.75
    cells["var_x_vel"] = 4.0
    cells["var_y_vel"] = 16.0
    return cells


def test_cells_with_velocity_thresholds_and_order():
    gs = 8
    cells = _grid(gs)
    moving = [(1, 0), (6, 2), (0, 5), (7, 7), (3, 5)]   # (x, y)
    for x, y in moving:
        cells[y * gs + x]["mean_x_vel"] = 2.0
        cells[y * gs + x]["mean_y_vel"] = 8.0          # v^T S^-1 v = 5
    slow = 2 * gs + 3
    cells[slow]["mean_x_vel"] = 2.0                    # 1
    want = sorted(y * gs + x for x, y in moving)
    assert oc.cells_with_velocity(gs, cells, 0.75, 5.0).tolist() == want           # both thresholds inclusive
    assert oc.cells_with_velocity(gs, cells, 0.75, np.nextafter(np.float32(5.0), np.float32(9))).size == 0
    assert oc.cells_with_velocity(gs, cells, np.nextafter(np.float32(0.75), np.float32(1)), 0.0).size == 0
    assert oc.cells_with_velocity(gs, cells, 0.75, 1.0).tolist() == sorted(want + [slow])
    # every cell passes with non-positive thresholds (m = 0 for resting cells)
    assert oc.cells_with_velocity(gs, cells, 0.0, 0.0).tolist() == list(range(gs * gs))


def test_cells_with_velocity_matches_f64_filter_away_from_thresholds():
    rng = np.random.default_rng(11)
    gs = 40
    cells = np.zeros(gs * gs, dtype=GRID_CELL_DTYPE)
    cells["occ_mass"] = rng.random(gs * gs).astype(np.float32)
    cells["free_mass"] = ((1 - cells["occ_mass"]) * rng.random(gs * gs)).astype(np.float32)
    cells["mean_x_vel"] = rng.normal(scale=4, size=gs * gs)
    cells["mean_y_vel"] = rng.normal(scale=4, size=gs * gs)
    cells["var_x_vel"] = rng.uniform(0.5, 8, gs * gs)
    cells["var_y_vel"] = rng.uniform(0.5, 8, gs * gs)
    cells["covar_xy_vel"] = rng.uniform(-0.4, 0.4, gs * gs) * np.sqrt(cells["var_x_vel"] * cells["var_y_vel"])
    o = cells["occ_mass"].astype(np.float64)
    f = cells["free_mass"].astype(np.float64)
    p = o + 0.5 * (1 - o - f)
    sxx, syy, sxy = (cells[k].astype(np.float64) for k in ("var_x_vel", "var_y_vel", "covar_xy_vel"))
    vx, vy = cells["mean_x_vel"].astype(np.float64), cells["mean_y_vel"].astype(np.float64)
    det = sxx * syy - sxy * sxy
    m = (vx * vx * syy - 2 * vx * vy * sxy + vy * vy * sxx) / det
    keep = (p >= 0.7) & (m >= 4.0)
    clear = (np.abs(p - 0.7) > 1e-4) & (np.abs(m - 4.0) > 1e-3 * np.maximum(1, m))
    got = set(oc.cells_with_velocity(gs, cells, 0.7, 4.0).tolist())
    idx = np.nonzero(clear)[0]
    assert {int(i) for i in idx if keep[i]} == {int(i) for i in idx if int(i) in got}
    assert keep.sum() > 20
