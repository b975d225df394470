"""Resampling CDF modes of the checker (exact / FP64 / FP32 scans) against closed forms and a
sequential FP64 CDF computed here with numpy."""
import numpy as np
import pytest

from oracle_ctypes import OracleDOGM, spacing_fractions
from parity_util import small_params

SEED = 21


def _resample_with(mode, weights, births, n, b, epoch=0):
    p = small_params(gs=16, res=0.5, n=n, b=b)
    o = OracleDOGM(p, seed=SEED, instance=3)
    o.set_cdf_mode(mode)
    o.set_weight_array(weights)
    if b:
        o.set_birth_particles(weight=births)
    o.epoch = epoch
    o.resampling()
    anc = o.get_resampled_indices()
    w = o.get_particles()["weight"]
    o.close()
    return anc, w


def _sequential_fp64_ancestors(weights, births, n, epoch=0):
    wj = np.concatenate([weights, births]).astype(np.float32)
    wj = np.where(wj > 0, wj, np.float32(0)).astype(np.float64)
    cdf = np.cumsum(wj)  # sequential left-to-right f64 sums
    u = spacing_fractions(SEED, 3, epoch, n)
    r = cdf[-1] * u
    a = np.searchsorted(cdf, r, side="left")
    return np.minimum(a, len(wj) - 1)


def test_integer_weights_all_modes_equal_exact():
    """Small integer weights sum exactly in f32 and f64 under any association, so every mode must
    reproduce the exact mode's ancestors and weights (pins the float modes' indexing)."""
    rng = np.random.default_rng(1)
    n, b = 5000, 700
    w = rng.integers(0, 4, n).astype(np.float32)
    bw = rng.integers(0, 3, b).astype(np.float32)
    ref, ref_w = _resample_with("exact", w, bw, n, b)
    for mode in ("fp64", "fp32", "fp32c"):
        anc, ww = _resample_with(mode, w, bw, n, b)
        assert np.array_equal(anc, ref), mode
        assert np.array_equal(ww.view(np.uint32), ref_w.view(np.uint32)), mode


def test_exact_mode_equals_sequential_fp64_on_dyadic_weights():
    """Dyadic weights make the sequential f64 CDF exact as well: the exact mode must match it."""
    rng = np.random.default_rng(2)
    n, b = 4000, 400
    w = (rng.integers(0, 1 << 10, n) / 1024.0).astype(np.float32)
    bw = (rng.integers(0, 1 << 10, b) / 1024.0).astype(np.float32)
    anc, _ = _resample_with("exact", w, bw, n, b)
    assert np.array_equal(anc, _sequential_fp64_ancestors(w, bw, n))


def test_log_uniform_weights_fp32_loses_ancestors_fp64_does_not():
    """Weights 10^(-9 U): the CDF spans 9 decades.  FP64, compensated FP32 and exact agree with the
    sequential FP64 CDF almost everywhere; the FP32 scan absorbs the small weights and moves many
    ancestors."""
    rng = np.random.default_rng(3)
    n, b = 200_000, 20_000
    w = (10.0 ** (-9.0 * rng.random(n))).astype(np.float32)
    bw = (10.0 ** (-9.0 * rng.random(b))).astype(np.float32)
    seq = _sequential_fp64_ancestors(w, bw, n)
    agree = {}
    for mode in ("exact", "fp64", "fp32c", "fp32"):
        anc, _ = _resample_with(mode, w, bw, n, b)
        assert np.all(np.diff(anc) >= 0), "ancestors of sorted draws are non-decreasing"
        agree[mode] = float(np.mean(anc == seq))
    assert agree["exact"] >= 0.9999 and agree["fp64"] >= 0.9999 and agree["fp32c"] >= 0.9999, agree
    assert agree["fp32"] < 0.999, agree


def test_weights_follow_mode_total():
    """New weight = f32(total) / N with the mode's total (dogm.cu:417); zero weights count as 0."""
    n, b = 3000, 0
    w = np.full(n, 0.25, np.float32)
    w[::7] = -1.0  # negative weights count as zero in the CDF
    for mode in ("exact", "fp64", "fp32", "fp32c"):
        _, ww = _resample_with(mode, w, np.zeros(0, np.float32), n, b)
        expect = np.float32(np.float32(0.25 * np.sum(w > 0)) / np.float32(n))
        assert np.all(ww == expect), mode


def test_two_sum_is_error_free_and_compensated_add_is_accurate():
    """TwoSum returns s = fl(a + b) and the exact error; the compensated add of (hi, lo) pairs keeps
    the sum to ~2^-44 relative (rational arithmetic)."""
    from fractions import Fraction
    import oracle_ctypes as o
    rng = np.random.default_rng(5)
    for _ in range(3000):
        a = np.float32(rng.normal() * 10.0 ** rng.integers(-8, 8))
        b = np.float32(rng.normal() * 10.0 ** rng.integers(-8, 8))
        s, e = o.two_sum_f32(a, b)
        assert s == np.float32(a + b)
        assert Fraction(float(s)) + Fraction(float(e)) == Fraction(float(a)) + Fraction(float(b))
    for _ in range(3000):
        x = [float(v) for v in rng.uniform(0, 1, 2) * 10.0 ** rng.integers(-9, 1, 2)]
        pa, pb = o.f32c_from_double(x[0]), o.f32c_from_double(x[1])
        hi, lo = o.f32c_add(pa, pb)
        exact = sum(Fraction(float(np.float32(v))) for v in pa + pb)
        assert abs(Fraction(float(hi)) + Fraction(float(lo)) - exact) <= exact * Fraction(1, 2 ** 44)
        assert abs(lo) <= abs(hi) * 2.0 ** -24


@pytest.mark.parametrize("mode", ["exact", "fp64", "fp32c"])
def test_systematic_draws_give_floor_or_ceil_copies(mode):
    """Systematic resampling: particle i gets floor(N w_i / W) or ceil(N w_i / W) copies, ancestors
    are sorted, and the offset changes with the epoch."""
    rng = np.random.default_rng(7)
    n, b = 6000, 600
    w = rng.gamma(0.5, size=n).astype(np.float32)
    bw = rng.gamma(0.5, size=b).astype(np.float32)
    w[::11] = 0.0
    p = small_params(gs=16, res=0.5, n=n, b=b)
    outs = []
    for epoch in (0, 1):
        o = OracleDOGM(p, seed=SEED, instance=3)
        o.set_cdf_mode(mode)
        o.set_draw_mode("systematic")
        o.set_weight_array(w)
        o.set_birth_particles(weight=bw)
        o.epoch = epoch
        o.resampling()
        anc = o.get_resampled_indices()
        o.close()
        assert np.all(np.diff(anc) >= 0)
        wj = np.concatenate([w, bw]).astype(np.float64)
        x = n * wj / wj.sum()
        copies = np.bincount(anc, minlength=n + b)
        assert np.all(copies >= np.floor(x - 1e-9)) and np.all(copies <= np.ceil(x + 1e-9))
        assert copies[::11][: n // 11].sum() == 0  # zero weights are never drawn
        outs.append(anc)
    assert not np.array_equal(outs[0], outs[1])
