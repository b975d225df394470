"""Pins the checker's velocity-moment sums (the f32 path's fixed f64 association) against facts that do
not come from the checker: exact integer sums (any association gives them when every partial sum is
exactly representable), math.fsum, and hand-derived results of cancelling inputs whose value
depends on the association.

The association: chunks of 256 particles (one wave64, 4 consecutive particles per lane); a
sequential fold per lane that restarts where a cell starts; a segmented wave scan of the lane folds
as a tree (row shifts 1, 2, 4, 8 inside 16-lane rows, then lane 15 of rows 0 / 2 into rows 1 / 3 and
lane 31 into rows 2 / 3); a cell ending in lane l = (scan of lane l - 1) + its fold unless it starts in
lane l; a cell crossing chunks = left fold over its chunk pieces.  CPU only."""
import math

import numpy as np
import pytest

import oracle_ctypes as oc

BIG = 2.0 ** 53  # f64 spacing 2 here: BIG + 1 rounds to BIG (ties to even), BIG + 2 is exact


def _sums(keys, x):
    keys = np.asarray(keys, np.int32)
    x = np.asarray(x, np.float32)
    if x.ndim == 1:
        x = np.stack([x] * 5)
    return oc.moment_sums(keys, x)


def _cells(keys):
    keys = np.asarray(keys)
    ends = np.nonzero(np.r_[keys[1:] != keys[:-1], True])[0]
    starts = np.r_[0, ends[:-1] + 1]
    return starts, ends


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_integer_values_give_exact_sums(seed):
    """Integer channel values (|sum| < 2^53): every association gives the exact sum, so this pins which
    particles each cell sums -- cells inside a lane, across lanes and rows, across chunks (up to five
    chunks), and a partial last chunk."""
    rng = np.random.default_rng(seed)
    lengths = np.concatenate([rng.integers(1, 6, 300), rng.integers(60, 900, 6), rng.integers(1, 3, 200)])
    rng.shuffle(lengths)
    keys = np.repeat(np.cumsum(rng.integers(1, 4, lengths.size)), lengths)
    n = keys.size
    assert n % 256 != 0
    x = rng.integers(-(1 << 20), 1 << 20, size=(5, n)).astype(np.float32)
    got = _sums(keys, x)
    starts, ends = _cells(keys)
    want = np.add.reduceat(x.astype(np.int64), starts, axis=1)
    assert np.array_equal(got[:, ends], want.astype(np.float64))
    mask = np.ones(n, bool)
    mask[ends] = False
    assert np.all(np.isnan(got[:, mask])), "sums are reported at the cells' last particles only"


def test_random_values_close_to_exact():
    rng = np.random.default_rng(7)
    keys = np.sort(rng.integers(0, 4000, 50_000))
    x = (rng.standard_normal((5, keys.size)) * 10.0 ** rng.uniform(-6, 3, (5, keys.size))).astype(np.float32)
    got = _sums(keys, x)
    starts, ends = _cells(keys)
    for c in range(5):
        for s, e in zip(starts[::37], ends[::37]):
            exact = math.fsum(float(v) for v in x[c, s:e + 1])
            mag = math.fsum(abs(float(v)) for v in x[c, s:e + 1])
            assert abs(got[c, e] - exact) <= 1e-14 * mag + 1e-300


def _single_cell(n, values):
    """keys: one cell over particles [0, n) after a one-particle cell 0 ... (cell 1), values placed."""
    keys = np.ones(n, np.int32)
    x = np.zeros(n, np.float32)
    for k, v in values.items():
        x[k] = v
    return keys, x


def test_lane_fold_is_sequential():
    # One cell in one lane: ((BIG + 1) + 1) - BIG = 0 (a pairwise fold gives 1, the exact sum is 2).
    keys = np.array([3, 3, 3, 3], np.int32)
    got = _sums(keys, [BIG, 1.0, 1.0, -BIG])
    assert got[0, 3] == 0.0


def test_lanes_fold_before_they_combine():
    # Lane 0 folds BIG + 1 + 1 + 1 = BIG; lane 1 folds 1 - BIG = -(BIG - 1); the cell = BIG - (BIG - 1) = 1.
    # (One sequential fold over all six gives 0, the exact sum is 4.)
    keys = np.array([5] * 6 + [6, 6], np.int32)
    got = _sums(keys, [BIG, 1, 1, 1, 1, -BIG, 0, 0])
    assert got[0, 5] == 1.0


def test_wave_scan_is_a_tree_over_row_lanes():
    # Lane 0 folds to BIG, lanes 1..15 fold to 1 each (one particle 1 and three 0).  The row's scan is
    # a balanced tree: only the first pairing BIG + 1 loses its 1, so lane 15 holds BIG + 14; a cell
    # ending in lane 16's first particle (value 0) = BIG + 14.  A sequential fold gives BIG.
    n = 17 * 4
    vals = {0: BIG}
    for l in range(1, 16):
        vals[4 * l] = 1.0
    keys, x = _single_cell(n, vals)
    keys[4 * 16 + 1:] = 2
    got = _sums(keys, x)
    assert got[0, 64] == BIG + 14.0


def test_row_broadcasts():
    # Rows 0 and 2 hold BIG at their first lane and rows 1 and 3 a 1 at their first lane; one cell over
    # all 256 particles.  Scan: row 0 -> BIG, row 1 -> 1, row 2 -> BIG, row 3 -> 1; lane 15 into row 1:
    # BIG + 1 = BIG; lane 47 into row 3: BIG + 1 = BIG; lane 31 into rows 2 and 3: BIG + BIG = 2 BIG.
    # The cell ends at particle 255 (lane 63): 2 BIG (the exact sum is 2 BIG + 2).
    vals = {0: BIG, 64: 1.0, 128: BIG, 192: 1.0}
    keys, x = _single_cell(256, vals)
    got = _sums(keys, x)
    assert got[0, 255] == 2 * BIG
    # Same cell but ending in lane 48 (particle 192): rows 2 and 3 first take lane 47 / 31 before row 3
    # adds its own fold: row 3 lane 48 = 1; + lane 47 (BIG) = BIG; + lane 31 (BIG) = 2 BIG.
    keys2 = keys.copy()
    keys2[193:] = 9
    got2 = _sums(keys2, x)
    assert got2[0, 192] == 2 * BIG


def test_chunks_fold_left():
    # A cell over chunks 0, 1, 2 (768 particles): pieces BIG, 1, 1 -> (BIG + 1) + 1 = BIG (a right fold
    # gives BIG + 2).
    vals = {0: BIG, 256: 1.0, 512: 1.0}
    keys, x = _single_cell(768 + 5, vals)
    keys[768:] = 4
    got = _sums(keys, x)
    assert got[0, 767] == BIG
    # The same cell starting in the middle of chunk 0 and ending inside chunk 2.
    keys3 = keys.copy()
    keys3[:100] = 0
    keys3[600:] = 4
    x3 = np.zeros_like(x)
    x3[100], x3[300], x3[550] = BIG, 1.0, 1.0
    got3 = _sums(keys3, x3)
    assert got3[0, 599] == BIG


def test_signed_zeros():
    """A cell of -0.0 values sums to -0.0 (no +0.0 enters a fold); a cell of one -0.0 after lanes of
    other cells is its own -0.0."""
    keys = np.repeat(np.arange(40), 7)
    x = np.full(keys.size, -0.0, np.float32)
    got = _sums(keys, x)
    _, ends = _cells(keys)
    assert np.all(np.signbit(got[0, ends]))
