"""Known-answer and property tests of the shared numerics (dynamic-occupancy-grid-map_amd/csrc/dogm_numerics.h),
exercised through the checker library.  These pin the primitives that the GPU kernels and the
checker share, independently of both."""
import math
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import oracle_ctypes as oc

# Random123 known-answer vectors for philox4x32 with 10 rounds (kat_vectors of the Random123
# distribution): (counter, key) -> output.
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expected):
    assert oc.philox(ctr, key) == expected


TORCH_PHILOX_SRC = r"""
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdint>
int main() {
  unsigned long long seed, sub, off;
  while (std::scanf("%llu %llu %llu", &seed, &sub, &off) == 3) {
    at::philox_engine e(seed, sub, 0); e.set_offset(off);
    unsigned a = e(), b = e(), c = e(), d = e();
    std::printf("%u %u %u %u\n", a, b, c, d);
  }
}
"""


def test_philox_matches_torch_engine(tmp_path):
    """Cross-check against PyTorch's independent Philox4x32-10 (ATen/core/PhiloxRNGEngine.h):
    key = seed, counter = (offset lo, offset hi, subsequence lo, subsequence hi)."""
    torch_inc = None
    try:
        import torch.utils.cpp_extension as ce
        for p in ce.include_paths():
            if (__import__("pathlib").Path(p) / "ATen/core/PhiloxRNGEngine.h").exists():
                torch_inc = p
    except Exception:
        pass
    if torch_inc is None or shutil.which("g++") is None:
        pytest.skip("torch headers or g++ not available")
    src = tmp_path / "ph.cpp"
    src.write_text(TORCH_PHILOX_SRC)
    exe = tmp_path / "ph"
    subprocess.run(["g++", "-std=c++17", f"-I{torch_inc}", str(src), "-o", str(exe)], check=True)
    rng = np.random.default_rng(7)
    cases = [tuple(int(v) for v in rng.integers(0, 2**64, size=3, dtype=np.uint64)) for _ in range(64)]
    out = subprocess.run([str(exe)], input="\n".join(f"{a} {b} {c}" for a, b, c in cases), text=True,
                         capture_output=True, check=True).stdout.split("\n")
    for (seed, sub, off), line in zip(cases, out):
        ctr = (off & 0xffffffff, off >> 32, sub & 0xffffffff, sub >> 32)
        key = (seed & 0xffffffff, seed >> 32)
        assert oc.philox(ctr, key) == tuple(int(v) for v in line.split())


def test_uniform_maps_ranges():
    assert oc.uniform01(0) == 0.0
    assert oc.uniform01(0xffffffff) == 1.0 - 2.0**-24
    assert oc.uniform01(0x000000ff) == 0.0  # the low 8 bits are dropped
    assert oc.uniform01(0x00000100) == 2.0**-24
    assert oc.uniform01_oc(0) == 2.0**-24
    assert oc.uniform01_oc(0xffffffff) == 1.0
    assert oc.uniform01_oc(0x7fffff00) == 0.5


def _ulp32(x: float) -> float:
    x32 = np.float32(x)
    return float(np.spacing(np.abs(x32)))


def test_log_f32_accuracy():
    """Our log against the correctly rounded math.log on the Box-Muller input grid (0, 1]."""
    rng = np.random.default_rng(1)
    ks = np.concatenate([np.arange(1, 2000), rng.integers(1, 2**24 + 1, size=40000),
                         2**24 - np.arange(0, 2000)])
    worst = 0.0
    for k in ks:
        u = float(k) * 2.0**-24
        got = oc.log_f32(u)
        ref = math.log(u)
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / _ulp32(ref))
    assert worst <= 2.0, worst


def test_sincos_2pi_accuracy():
    rng = np.random.default_rng(2)
    ks = np.concatenate([np.arange(0, 3000), rng.integers(0, 2**24, size=40000),
                         np.array([2**22, 2**23, 3 * 2**22, 2**24 - 1])])
    worst = 0.0
    for k in ks:
        t = float(k) * 2.0**-24
        s, c = oc.sincos_2pi(t)
        worst = max(worst, abs(s - math.sin(2 * math.pi * t)), abs(c - math.cos(2 * math.pi * t)))
    assert worst < 3.0e-7, worst
    # Quadrant boundaries land exactly on the axes up to rounding of the centred polynomials.
    for t, (se, ce) in [(0.0, (0.0, 1.0)), (0.25, (1.0, 0.0)), (0.5, (0.0, -1.0)), (0.75, (-1.0, 0.0))]:
        s, c = oc.sincos_2pi(t)
        assert abs(s - se) < 2e-7 and abs(c - ce) < 2e-7


def test_box_muller_distribution():
    from scipy import stats
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2**32, size=60000, dtype=np.uint64)
    b = rng.integers(0, 2**32, size=60000, dtype=np.uint64)
    z = np.array([oc.box_muller(int(x), int(y)) for x, y in zip(a, b)])
    for col in (z[:, 0], z[:, 1]):
        assert abs(col.mean()) < 0.02
        assert abs(col.var() - 1.0) < 0.03
        assert stats.kstest(col, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.02
    # u1 = 1 (top of the grid) gives radius 0.
    assert oc.box_muller(0xffffffff, 12345) == (0.0, 0.0) or max(map(abs, oc.box_muller(0xffffffff, 12345))) == 0.0


def _fx_exact(v: float) -> int:
    """trunc(v * 2^88) with exact rationals, saturating like the C code."""
    if math.isnan(v) or math.isinf(v) or abs(v) >= 2.0**39:
        mag = 1 << 126
        return -mag if (v < 0 or (math.isinf(v) and v < 0)) else mag
    fr = Fraction(v) * (1 << 88)
    q = abs(fr.numerator) // fr.denominator
    return -q if fr < 0 else q


def test_fx_from_f32_nonneg_equals_general_form():
    """Every float with a clear sign bit (all 2^31 patterns: zero, subnormals, normals, +inf, NaNs
    with a clear sign) converts identically through the sign-free form the merge kernel uses."""
    step = 1 << 24
    for start in range(0, 1 << 31, step):
        bits = np.arange(start, start + step, dtype=np.uint32)
        assert oc.fx_from_f32_nonneg_mismatches(bits) == 0, hex(start)


def test_fx_from_f64_nonneg_equals_general_form():
    """Doubles with a clear sign bit convert identically through the sign-free form of the resampling
    draws: every exponent with random significands, the significand extremes, and the boundaries of
    the shift ranges (sh = E - 987 at -64, 0, 11, 12, 64, 74, 75)."""
    rng = np.random.default_rng(12)
    exps = np.repeat(np.arange(2048, dtype=np.uint64), 512)
    mant = rng.integers(0, 1 << 52, size=exps.size, dtype=np.uint64)
    bits = (exps << np.uint64(52)) | mant
    edge = []
    for e in range(2048):
        for m in (0, 1, (1 << 52) - 1, 1 << 51):
            edge.append((e << 52) | m)
    bits = np.concatenate([bits, np.array(edge, dtype=np.uint64)])
    assert oc.fx_from_f64_nonneg_mismatches(bits) == 0


def test_resampling_exponentials_have_a_clear_sign():
    """All 2^24 exponentials -log(u), u in (0, 1], are >= 0 (or -0) and convert identically through the
    sign-free form the merge kernel uses."""
    assert oc.exponential_sign_violations() == 0


def test_fx_from_f32_exact():
    rng = np.random.default_rng(4)
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 1e-30, 2.0**-88, 2.0**-89, 3 * 2.0**-90, 2.0**38, -(2.0**38) * 1.5,
            float(np.float32(1e-45)), float(np.float32(1.17549435e-38)), float("inf"), float("-inf")]
    exps = rng.uniform(-100, 38, size=3000)
    vals += [float(np.float32(s * 2.0**e)) for s, e in zip(rng.choice([-1.0, 1.0], size=3000), exps)]
    for v in vals:
        v32 = float(np.float32(v))
        assert oc.fx_from_f32(v32) == _fx_exact(v32), v32


def test_fx_from_f64_exact():
    rng = np.random.default_rng(5)
    exps = rng.uniform(-100, 38, size=3000)
    for s, e in zip(rng.choice([-1.0, 1.0], size=3000), exps):
        v = s * 2.0**e * (1 + rng.random())
        assert oc.fx_from_f64(v) == _fx_exact(v)


def _check_rounding(x: int, got: float, dtype):
    """got must be a nearest representable value of x / 2^88 (ties to even)."""
    exact = Fraction(x, 1 << 88)
    g = dtype(got)
    assert Fraction(float(g)) == Fraction(got)
    up = np.nextafter(g, dtype(np.inf))
    dn = np.nextafter(g, dtype(-np.inf))
    e0 = abs(Fraction(float(g)) - exact)
    for nb in (up, dn):
        e1 = abs(Fraction(float(nb)) - exact)
        assert e0 <= e1
        if e0 == e1:  # tie: even significand
            bits = np.array([g]).view(np.uint32 if dtype is np.float32 else np.uint64)[0]
            assert bits % 2 == 0


def test_fx_to_float_correctly_rounded():
    rng = np.random.default_rng(6)
    xs = [1, -1, 3, (1 << 88), -(1 << 88), (1 << 112) + (1 << 88), (1 << 24) + 1, (1 << 25) + 1, (1 << 25) + 3,
          (1 << 53) + 1, (1 << 54) + 2, (1 << 126) - 1]
    for _ in range(3000):
        bits = int(rng.integers(1, 126))
        x = int(rng.integers(0, 2**63)) << max(0, bits - 63)
        x = x | int(rng.integers(0, 2**min(bits, 62)))
        xs.append(x if rng.random() < 0.5 else -x)
    for x in xs:
        if x == 0:
            continue
        _check_rounding(x, oc.fx_to_f32(x), np.float32)
        _check_rounding(x, oc.fx_to_f64(x), np.float64)
    assert oc.fx_to_f32(0) == 0.0


def test_fx_roundtrip_f32():
    rng = np.random.default_rng(8)
    for v in rng.uniform(-1000, 1000, size=2000).astype(np.float32):
        assert np.float32(oc.fx_to_f32(oc.fx_from_f32(float(v)))) == v


def test_fx_scale_u64():
    rng = np.random.default_rng(9)
    for _ in range(2000):
        total = int(rng.integers(0, 2**63)) << int(rng.integers(0, 63))
        u = int(rng.integers(0, 2**64, dtype=np.uint64))
        assert oc.fx_scale_u64(total, u) == (total * u) >> 64
    assert oc.fx_scale_u64(12345, 0) == 0
    assert oc.fx_scale_u64(1 << 100, 2**64 - 1) == (1 << 100) - (1 << 36)


def test_order_end_properties():
    """Particle-order ends are monotone in the prefix, bounded by v, and 0 at prefix 0."""
    rng = np.random.default_rng(10)
    for _ in range(200):
        v = int(rng.integers(1, 3_000_000))
        total = int(rng.integers(1, 2**62)) << int(rng.integers(0, 40))
        prefixes = sorted(total * int(q) // 2**62 for q in rng.integers(0, 2**62, size=50)) + [total]
        ends = [oc.order_end(p, total, v) for p in prefixes]
        assert all(0 <= e <= v for e in ends)
        assert ends == sorted(ends)
        assert oc.order_end(0, total, v) == 0
        # Within one particle of the exact proportional split.
        for p, e in zip(prefixes, ends):
            assert abs(e - Fraction(p * v, total)) <= 1
