// Numerics shared bit-for-bit by the gfx950 kernels and the host-side checker.
//
// Everything in this header is plain integer or IEEE-754 single/double arithmetic
// with an explicit operation order, so that hipcc (device) and g++ (host) produce
// identical bits as long as both compile with -ffp-contract=off, correctly rounded
// f32 division/sqrt and no denormal flushing:
//
//  * Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, "Parallel Random
//    Numbers: As Easy as 1, 2, 3", SC'11).  It replaces the reference's per-thread
//    cuRAND XORWOW states (dogm/src/kernel/init.cu:16-22), whose streams depend on
//    the launch geometry (dogm/src/dogm.cu:45-47, 68).  A draw here is a pure
//    function of (seed, instance, stage tag, epoch, element index).
//  * Uniform and normal maps with the reference's conventions
//    (dogm/include/dogm/cuda_utils.h:31-40): U[min, max) = min + (max - min) * w with
//    w in [0, 1); N(mean, sd) = n * sd + mean.  Normals come from Box-Muller with
//    our own log and sin/cos (range reduction plus Taylor polynomials).
//  * Exact fixed-point sums.  Every per-cell sum and every prefix sum on the filter
//    path is accumulated as a signed 128-bit integer with 88 fractional bits.  Integer
//    addition is associative, so any parallel order gives the same bits; the result
//    is rounded once (round-to-nearest-even) back to f32.  This restates the
//    reference's f32 inclusive-scan-and-subtract per-cell sums
//    (dogm/include/dogm/common.h:11-32) without their order dependence and
//    cancellation.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DOGM_HD __host__ __device__ __forceinline__
#define DOGM_HDM __host__ __device__ __forceinline__
#else
#define DOGM_HD static inline
#define DOGM_HDM inline
#endif

namespace dogm_num
{

// ---------------------------------------------------------------------------------------------
// Bit casts
// ---------------------------------------------------------------------------------------------
DOGM_HD uint32_t f32_bits(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, sizeof(u));
    return u;
}

DOGM_HD float f32_from_bits(uint32_t u)
{
    float f;
    __builtin_memcpy(&f, &u, sizeof(f));
    return f;
}

DOGM_HD uint64_t f64_bits(double f)
{
    uint64_t u;
    __builtin_memcpy(&u, &f, sizeof(u));
    return u;
}

DOGM_HD double f64_from_bits(uint64_t u)
{
    double f;
    __builtin_memcpy(&f, &u, sizeof(f));
    return f;
}

// ---------------------------------------------------------------------------------------------
// Philox4x32-10
// ---------------------------------------------------------------------------------------------
struct u32x4
{
    uint32_t x, y, z, w;
};

// Multipliers and Weyl key increments from the Philox paper (Table 2 / section 5.2).
constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;

DOGM_HD u32x4 philox_round(u32x4 c, uint32_t k0, uint32_t k1)
{
    const uint64_t p0 = (uint64_t)PHILOX_M0 * (uint64_t)c.x;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * (uint64_t)c.z;
    u32x4 r;
    r.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    r.y = (uint32_t)p1;
    r.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    r.w = (uint32_t)p0;
    return r;
}

DOGM_HD u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1)
{
    ctr = philox_round(ctr, k0, k1);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 1; r < 10; ++r)
    {
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
        ctr = philox_round(ctr, k0, k1);
    }
    return ctr;
}

// Stage tags: the third counter word, one per random-consuming stage.
enum RngTag : uint32_t
{
    TAG_PREDICT = 1u,       // 4 normals per persistent particle
    TAG_INIT_PARTICLES = 2u,  // 2 uniforms per particle on the first measurement
    TAG_BIRTH = 3u,         // 2 normals per birth particle
    TAG_RESAMPLE = 4u,      // one 64-bit uniform per resampled slot
};

// Counter layout: (element index, epoch, stage tag, instance); key = 64-bit seed.
DOGM_HD u32x4 draw(uint64_t seed, uint32_t instance, uint32_t tag, uint32_t epoch, uint32_t elem)
{
    u32x4 c;
    c.x = elem;
    c.y = epoch;
    c.z = tag;
    c.w = instance;
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// ---------------------------------------------------------------------------------------------
// Uniform maps (24-bit grids, exactly representable in f32)
// ---------------------------------------------------------------------------------------------
// w in [0, 1): the reference's "1 - u" with u in (0, 1] (cuda_utils.h:31-35).
DOGM_HD float uniform01(uint32_t x)
{
    return (float)(x >> 8) * 0x1p-24f;
}

// u in (0, 1]: input of the logarithm in Box-Muller.
DOGM_HD float uniform01_oc(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 0x1p-24f;
}

// U[min, max) = min + (max - min) * w  (cuda_utils.h:31-35, same expression order).
DOGM_HD float uniform_range(uint32_t x, float lo, float hi)
{
    return lo + (hi - lo) * uniform01(x);
}

// ---------------------------------------------------------------------------------------------
// Elementary functions for Box-Muller
// ---------------------------------------------------------------------------------------------
// Natural logarithm of a positive normal float.
// x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1),
// |s| <= 0.1716, atanh series to s^11 (truncation < 2e-10 relative).
DOGM_HD float log_f32(float x)
{
    const uint32_t b = f32_bits(x);
    int e = (int)((b >> 23) & 0xffu) - 127;
    float m = f32_from_bits((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 1.41421354f)
    {
        m = m * 0.5f;
        e = e + 1;
    }
    const float f = m - 1.0f;  // exact (Sterbenz)
    const float s = f / (m + 1.0f);
    const float s2 = s * s;
    float p = fmaf(s2, 0.0909090936f, 0.111111112f);  // 1/11, 1/9
    p = fmaf(s2, p, 0.142857149f);                    // 1/7
    p = fmaf(s2, p, 0.200000003f);                    // 1/5
    p = fmaf(s2, p, 0.333333343f);                    // 1/3
    const float two_s = s + s;
    const float lnm = fmaf(two_s, s2 * p, two_s);
    const float fe = (float)e;
    // ln 2 split: hi has 16 significant bits so fe * hi is exact for |e| < 256.
    return fmaf(fe, 0.693145752f, fmaf(fe, 1.42860677e-06f, lnm));
}

// sin(2 pi t), cos(2 pi t) for t in [0, 1).
// t * 4 = q + f (q quadrant, f in [0, 1)); angle = q pi/2 + pi/4 + x with x = (f - 1/2) pi/2,
// |x| <= pi/4; Taylor polynomials to x^9 / x^10 (truncation < 2e-9).
DOGM_HD void sincos_2pi_f32(float t, float* s_out, float* c_out)
{
    const float t4 = t * 4.0f;  // exact
    int q = (int)t4;
    if (q > 3)
        q = 3;
    const float f = t4 - (float)q;  // exact
    const float x = (f - 0.5f) * 1.57079637f;
    const float x2 = x * x;
    float ps = fmaf(x2, 2.75573188e-06f, -1.98412701e-04f);  // 1/9!, -1/7!
    ps = fmaf(x2, ps, 8.33333377e-03f);                      // 1/5!
    ps = fmaf(x2, ps, -0.166666672f);                        // -1/3!
    const float sx = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, -2.75573188e-07f, 2.48015876e-05f);  // -1/10!, 1/8!
    pc = fmaf(x2, pc, -1.38888892e-03f);                     // -1/6!
    pc = fmaf(x2, pc, 4.16666679e-02f);                      // 1/4!
    pc = fmaf(x2, pc, -0.5f);                                // -1/2!
    const float cx = fmaf(x2, pc, 1.0f);
    const float sp = (sx + cx) * 0.707106769f;  // sin(pi/4 + x)
    const float cp = (cx - sx) * 0.707106769f;  // cos(pi/4 + x)
    float s, c;
    if (q == 0)
    {
        s = sp;
        c = cp;
    }
    else if (q == 1)
    {
        s = cp;
        c = -sp;
    }
    else if (q == 2)
    {
        s = -sp;
        c = -cp;
    }
    else
    {
        s = -cp;
        c = sp;
    }
    *s_out = s;
    *c_out = c;
}

// Two standard normals from two 32-bit words (Box-Muller).
DOGM_HD void box_muller(uint32_t a, uint32_t b, float* z0, float* z1)
{
    const float u1 = uniform01_oc(a);
    const float t = uniform01(b);
    const float r = sqrtf(-2.0f * log_f32(u1));
    float s, c;
    sincos_2pi_f32(t, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

// N(mean, sd) = n * sd + mean  (cuda_utils.h:37-40).
DOGM_HD float normal_scale(float n, float mean, float sd)
{
    return n * sd + mean;
}

// ---------------------------------------------------------------------------------------------
// Exact fixed-point sums: signed 128-bit integers, value = raw * 2^-88.
// Representable magnitude < 2^39; f32 inputs below 2^-88 contribute their bits above 2^-88
// (truncation toward zero).
// ---------------------------------------------------------------------------------------------
typedef __int128 fx_t;
typedef unsigned __int128 ufx_t;
constexpr int FX_FRAC = 88;

DOGM_HD int clz128(ufx_t v)
{
    const uint64_t hi = (uint64_t)(v >> 64);
    const uint64_t lo = (uint64_t)v;
    if (hi != 0)
        return __builtin_clzll(hi);
    if (lo != 0)
        return 64 + __builtin_clzll(lo);
    return 128;
}

// Exact (truncating toward zero below 2^-88).  |v| >= 2^39, inf and NaN saturate to the
// largest representable magnitude with the sign of v (NaN counts as +).  Written on 64-bit halves:
// the 24-bit significand shifted by sh = E - 62 lands in at most two adjacent words.
DOGM_HD fx_t fx_from_f32(float v)
{
    const uint32_t b = f32_bits(v);
    const int sh = (int)((b >> 23) & 0xffu) - 150 + FX_FRAC;  // value = m * 2^(E - 150), fixed point = m << sh
    const uint64_t m = (uint64_t)((b & 0x007fffffu) | 0x00800000u);
    // Candidates for the three shift ranges (every shift count kept in [0, 63]).  E = 0 (zeros,
    // subnormals) gives sh = -62, where m >> 62 = 0 clears the value without a test of its own.
    const uint64_t hi_a = m << ((sh - 64) & 63);   // 64 <= sh <= 103
    const uint64_t lo_b = m << (sh & 63);          // 0 <= sh < 64
    const uint64_t hi_b = m >> ((64 - sh) & 63);   // 41 <= sh < 64
    const uint64_t lo_c = m >> ((-sh) & 63);       // sh < 0
    uint64_t lo = sh >= 64 ? 0 : (sh >= 0 ? lo_b : lo_c);
    uint64_t hi = sh >= 64 ? hi_a : (sh > 40 ? hi_b : 0);
    const bool sat = sh > 103;  // |v| >= 2^40, infinities and NaNs (E = 255 gives sh = 193)
    lo = sat ? 0 : lo;
    hi = sat ? (1ull << 62) : hi;  // 2^126
    // Two's-complement negation for negative v: (x ^ s) - s with s = all ones, on 128 bits.
    const ufx_t sgn = (ufx_t)0 - (ufx_t)(b >> 31);
    return (fx_t)(((((ufx_t)hi << 64) | (ufx_t)lo) ^ sgn) - sgn);
}

// fx_from_f32 for v with a clear sign bit (+0, positive values, +inf; what cdf_weight returns):
// the same value without the two's-complement step.
DOGM_HD fx_t fx_from_f32_nonneg(float v)
{
    const uint32_t b = f32_bits(v);
    const int sh = (int)((b >> 23) & 0xffu) - 150 + FX_FRAC;
    const uint64_t m = (uint64_t)((b & 0x007fffffu) | 0x00800000u);
    const uint64_t hi_a = m << ((sh - 64) & 63);
    const uint64_t lo_b = m << (sh & 63);
    const uint64_t hi_b = m >> ((64 - sh) & 63);
    const uint64_t lo_c = m >> ((-sh) & 63);
    uint64_t lo = sh >= 64 ? 0 : (sh >= 0 ? lo_b : lo_c);
    uint64_t hi = sh >= 64 ? hi_a : (sh > 40 ? hi_b : 0);
    const bool sat = sh > 103;
    lo = sat ? 0 : lo;
    hi = sat ? (1ull << 62) : hi;
    return (fx_t)(((ufx_t)hi << 64) | (ufx_t)lo);
}

DOGM_HD fx_t fx_from_f64(double v)
{
    const uint64_t b = f64_bits(v);
    const uint32_t E = (uint32_t)((b >> 52) & 0x7ffu);
    if (E == 0)
        return 0;
    const bool neg = (b >> 63) != 0;
    ufx_t mag;
    const int sh = (int)E - 1075 + FX_FRAC;  // value = m * 2^(E - 1075)
    if (E == 0x7ffu || sh > 74)
    {
        mag = (((ufx_t)1) << 126);
    }
    else
    {
        const uint64_t m = (b & 0x000fffffffffffffull) | 0x0010000000000000ull;
        if (sh >= 0)
            mag = ((ufx_t)m) << sh;
        else if (sh > -64)
            mag = (ufx_t)(m >> (-sh));
        else
            mag = 0;
    }
    return neg ? -(fx_t)mag : (fx_t)mag;
}

// fx_from_f64 for v with a clear sign bit (+0, positive values, +inf; the resampling draws): the
// same value, written on 64-bit halves without branches (the 53-bit significand shifted by
// sh = E - 987 lands in at most two adjacent words) and without the two's-complement step.
DOGM_HD fx_t fx_from_f64_nonneg(double v)
{
    const uint64_t b = f64_bits(v);
    const int E = (int)((b >> 52) & 0x7ffu);
    const int sh = E - 1075 + FX_FRAC;
    const uint64_t m = (b & 0x000fffffffffffffull) | 0x0010000000000000ull;
    const uint64_t hi_a = m << ((sh - 64) & 63);  // 64 <= sh <= 74
    const uint64_t lo_b = m << (sh & 63);         // 0 <= sh < 64
    const uint64_t hi_b = m >> ((64 - sh) & 63);  // 12 <= sh < 64
    const uint64_t lo_c = m >> ((-sh) & 63);      // -64 < sh < 0
    // E = 0 (+0, subnormals) gives sh = -987: both words 0.
    uint64_t lo = sh >= 64 ? 0 : (sh >= 0 ? lo_b : (sh > -64 ? lo_c : 0));
    uint64_t hi = sh >= 64 ? hi_a : (sh > 11 ? hi_b : 0);
    const bool sat = E == 0x7ff || sh > 74;  // v >= 2^39, +inf, NaN
    lo = sat ? 0 : lo;
    hi = sat ? (1ull << 62) : hi;
    return (fx_t)(((ufx_t)hi << 64) | (ufx_t)lo);
}

// Round-to-nearest-even of raw * 2^-88 to a float with `mbits` significand bits
// (24 for f32, 53 for f64).  Returns the significand (with the implicit bit) and the
// unbiased exponent of its leading bit; mag must be non-zero.
DOGM_HD void fx_round(ufx_t mag, int mbits, uint64_t* mant_out, int* exp_out)
{
    const int msb = 127 - clz128(mag);
    int e = msb - FX_FRAC;
    uint64_t mant;
    if (msb < mbits)
    {
        mant = (uint64_t)(mag << (mbits - 1 - msb));
    }
    else
    {
        const int sh = msb - (mbits - 1);
        const ufx_t q = mag >> sh;
        const ufx_t rem = mag & ((((ufx_t)1) << sh) - 1);
        const ufx_t half = ((ufx_t)1) << (sh - 1);
        mant = (uint64_t)q;
        if (rem > half || (rem == half && (mant & 1u)))
        {
            mant += 1;
            if (mant == (1ull << mbits))
            {
                mant >>= 1;
                e += 1;
            }
        }
    }
    *mant_out = mant;
    *exp_out = e;
}

// Round-to-nearest-even of raw * 2^-88 to f32: the magnitude is normalised into one 64-bit word
// (leading bit at 63) plus a sticky bit for what falls below it.  Its top 32 bits, with every lower
// bit folded into the lowest one, keep the round bit and a sticky bit below the 24-bit significand,
// so the (round-to-nearest-even) u32 -> f32 conversion rounds them exactly as the whole magnitude;
// the result's exponent is then set on the bit pattern (every non-zero fixed-point value is a
// normal f32: 2^-88 <= |v| < 2^40).
DOGM_HD float fx_to_f32(fx_t v)
{
    const ufx_t u = (ufx_t)v;
    const uint32_t neg = (uint32_t)(u >> 127);
    const ufx_t sgn = (ufx_t)0 - (ufx_t)neg;
    const ufx_t mag = (u ^ sgn) - sgn;
    const uint64_t hi = (uint64_t)(mag >> 64), lo = (uint64_t)mag;
    if ((hi | lo) == 0)
        return 0.0f;
    uint64_t t;
    bool sticky;
    int msb;
    if (hi != 0)
    {
        const int z = __builtin_clzll(hi);
        t = z ? (hi << z) | (lo >> (64 - z)) : hi;
        sticky = (z ? (lo << z) : lo) != 0;
        msb = 127 - z;
    }
    else
    {
        const int z = __builtin_clzll(lo);
        t = lo << z;
        sticky = false;
        msb = 63 - z;
    }
    const uint32_t w = (uint32_t)(t >> 32) | (((uint32_t)t != 0u || sticky) ? 1u : 0u);
    // (float)w lies in [2^31, 2^32]; scale it by 2^(msb - 88 - 31) on the exponent field.
    const uint32_t bits = f32_bits((float)w) + ((uint32_t)(msb - FX_FRAC - 31) << 23);
    return f32_from_bits(bits | (neg << 31));
}

// Round-to-nearest-even of raw * 2^-88 to f64, as fx_to_f32: the magnitude is normalised into one 64-bit
// word with every lower bit folded into its lowest one (11 bits below the 53-bit significand keep the round
// and sticky information), converted with one rounding (hi * 2^32 + lo, both exact), and the exponent is set
// on the bit pattern (every non-zero fixed-point value is a normal f64).
DOGM_HD double fx_to_f64(fx_t v)
{
    const ufx_t u = (ufx_t)v;
    const uint64_t neg = (uint64_t)(u >> 127);
    const ufx_t sgn = (ufx_t)0 - (ufx_t)neg;
    const ufx_t mag = (u ^ sgn) - sgn;
    const uint64_t hi = (uint64_t)(mag >> 64), lo = (uint64_t)mag;
    if ((hi | lo) == 0)
        return 0.0;
    uint64_t t;
    bool sticky;
    int msb;
    if (hi != 0)
    {
        const int z = __builtin_clzll(hi);
        t = z ? (hi << z) | (lo >> (64 - z)) : hi;
        sticky = (z ? (lo << z) : lo) != 0;
        msb = 127 - z;
    }
    else
    {
        const int z = __builtin_clzll(lo);
        t = lo << z;
        sticky = false;
        msb = 63 - z;
    }
    const uint64_t w = t | (sticky ? 1ull : 0ull);
    const double f = (double)(uint32_t)(w >> 32) * 4294967296.0 + (double)(uint32_t)w;  // in [2^63, 2^64]
    const uint64_t bits = f64_bits(f) + ((uint64_t)(int64_t)(msb - FX_FRAC - 63) << 52);
    return f64_from_bits(bits | (neg << 63));
}

// floor(total * u / 2^64) for total >= 0: maps a 64-bit uniform onto [0, total).
DOGM_HD fx_t fx_scale_u64(fx_t total, uint64_t u)
{
    const ufx_t t = (ufx_t)total;
    const uint64_t hi = (uint64_t)(t >> 64);
    const uint64_t lo = (uint64_t)t;
    const ufx_t a = (ufx_t)hi * (ufx_t)u;
    const ufx_t b = ((ufx_t)lo * (ufx_t)u) >> 64;
    return (fx_t)(a + b);
}

// ---------------------------------------------------------------------------------------------
// Multinomial resampling draws (dogm/src/kernel/resampling.cu:17-47 draws N uniforms and sorts
// them).  Sorted uniforms are generated directly from exponential spacings: with E_0..E_N i.i.d.
// Exp(1) and S_k = E_0 + ... + E_k, the vector (S_0 / S_N, ..., S_{N-1} / S_N) has the law of the
// order statistics of N i.i.d. U(0, 1) (Devroye, Non-Uniform Random Variate Generation, 1986,
// ch. V.3).  S is an exact fixed-point prefix; the ratio is one IEEE double division; the result is
// a 64-bit fraction of the CDF total.
// ---------------------------------------------------------------------------------------------
// E_k for k = 4m .. 4m + 3 come from the four words of draw m.
DOGM_HD void resample_exponentials4(uint64_t seed, uint32_t instance, uint32_t epoch, uint32_t m, float* e)
{
    const u32x4 r = draw(seed, instance, TAG_RESAMPLE, epoch, m);
    e[0] = -log_f32(uniform01_oc(r.x));
    e[1] = -log_f32(uniform01_oc(r.y));
    e[2] = -log_f32(uniform01_oc(r.z));
    e[3] = -log_f32(uniform01_oc(r.w));
}

DOGM_HD float resample_exponential(uint64_t seed, uint32_t instance, uint32_t epoch, uint32_t k)
{
    float e[4];
    resample_exponentials4(seed, instance, epoch, k >> 2, e);
    return e[k & 3u];
}

// The draw for prefix S_k on the CDF axis, fx(total * u_k) with u_k = (S_k >> sh) * inv, where
// sh brings S_N below 2^53 (so S_k >> sh converts to double exactly) and inv = 1 / (S_N >> sh).
// Every step is monotone in S_k, so sorted prefixes give sorted draws; u_k is capped below 1.
struct Spacing
{
    int shift;
    double inv;
    double total;  // CDF total as a double
};

DOGM_HD Spacing spacing_setup(fx_t s_n, fx_t cdf_total)
{
    Spacing sp;
    sp.shift = 0;
    sp.inv = 0.0;
    if (s_n > 0)
    {
        const int msb = 127 - clz128((ufx_t)s_n);
        sp.shift = msb > 52 ? msb - 52 : 0;
        sp.inv = 1.0 / (double)(uint64_t)(s_n >> sp.shift);
    }
    // A total wrapped below zero (sums of saturated weights) counts as 0, so draws are never negative.
    sp.total = cdf_total > 0 ? fx_to_f64(cdf_total) : 0.0;
    return sp;
}

// Fraction u_k = S_k / S_N in [0, 1) of the spacing construction (the same for every CDF mode).
DOGM_HD double spacing_fraction(fx_t s_k, const Spacing& sp)
{
    double u = (double)(uint64_t)(s_k >> sp.shift) * sp.inv;
    if (!(u < 1.0))
        u = 0x1.fffffffffffffp-1;
    return u;
}

DOGM_HD fx_t spacing_draw(fx_t s_k, const Spacing& sp)
{
    return fx_from_f64(sp.total * spacing_fraction(s_k, sp));
}

// spacing_draw through the sign-free conversion: the product of the total (never below 0) and
// u_k (in [0, 1)) has a clear sign bit, so the value is the same.
DOGM_HD fx_t spacing_draw_nonneg(fx_t s_k, const Spacing& sp)
{
    return fx_from_f64_nonneg(sp.total * spacing_fraction(s_k, sp));
}

// Systematic resampling (an opt-in draw mode; the reference draws multinomially,
// resampling.cu:17-47): one offset u in [0, 1) per resampling and draw k at (k + u) / N of the
// total.  Written in the spacing form: S_k = k * 2^32 + u32, S_N = N * 2^32, so the fraction, draw
// and ancestor search are those of the multinomial path.  u32 is the .x word of the Philox block
// with counter 0xffffffff of the resampling stream (the exponentials use counters below 2^30).
enum DrawMode
{
    DRAW_MULTINOMIAL = 0,
    DRAW_SYSTEMATIC = 1,
};

DOGM_HD uint32_t systematic_u32(uint64_t seed, uint32_t instance, uint32_t epoch)
{
    return draw(seed, instance, TAG_RESAMPLE, epoch, 0xffffffffu).x;
}

DOGM_HD fx_t systematic_s(int k, uint32_t u32)
{
    return (fx_t)(((ufx_t)(uint32_t)k << 32) | (ufx_t)u32);
}

// ---------------------------------------------------------------------------------------------
// Floating-point CDF modes of the resampling scan (precision study).  The exact mode above is the
// default.  In the float modes the CDF is a scan of the non-negative joint weights in f64 or f32
// with one fixed association (4-element sequential folds per thread, a 64-lane wave scan, and
// sequential carries across waves and tiles; the checker restates it), its running maximum (the
// monotone envelope E), total = E[J - 1], draws r_k = total * u_k in the mode's precision, and the
// ancestor is the first j with E[j] >= r_k.
// ---------------------------------------------------------------------------------------------
enum CdfMode
{
    CDF_EXACT = 0,
    CDF_FP64 = 1,
    CDF_FP32 = 2,
    CDF_FP32C = 3,  // compensated f32: double-float pairs
};

DOGM_HD double cdf_float_draw(double total, double u, int mode)
{
    if (mode == CDF_FP32)
        return (double)((float)total * (float)u);
    return total * u;  // FP64, and compensated FP32 whose totals carry ~48 bits
}

// Error-free sum of two floats (Knuth's TwoSum): s = fl(a + b), e = (a + b) - s exactly.
DOGM_HD void two_sum_f32(float a, float b, float* s, float* e)
{
    const float x = a + b;
    const float bv = x - a;
    const float av = x - bv;
    *e = (a - av) + (b - bv);
    *s = x;
}

// Compensated f32 value ("double-float"): hi + lo with |lo| <= ulp(hi) / 2 after every addition.
// Addition (Dekker / Knuth): TwoSum of the high parts, the low parts and that error added in f32,
// then renormalised with a second TwoSum.  Conversions to and from double are exact for normalised
// pairs.
struct f32c
{
    float hi, lo;
    f32c() = default;
    DOGM_HDM explicit f32c(int i) : hi((float)i), lo(0.0f) {}
    DOGM_HDM explicit f32c(float f) : hi(f), lo(0.0f) {}
    DOGM_HDM explicit f32c(double d) : hi((float)d), lo((float)(d - (double)(float)d)) {}
    DOGM_HDM explicit operator double() const { return (double)hi + (double)lo; }
};

DOGM_HD f32c operator+(f32c a, f32c b)
{
    float s, e;
    two_sum_f32(a.hi, b.hi, &s, &e);
    const float t = (a.lo + b.lo) + e;
    f32c r;
    two_sum_f32(s, t, &r.hi, &r.lo);
    return r;
}

DOGM_HD bool operator>(f32c a, f32c b)
{
    return (double)a > (double)b;
}

// CDF contribution of a joint weight: negative values (from rounding in rho_p = occ - rho_b) and
// NaN count as zero, so the CDF is non-decreasing and the ancestor search is well defined.
DOGM_HD float cdf_weight(float w)
{
    return w > 0.0f ? w : 0.0f;
}

// ---------------------------------------------------------------------------------------------
// Particle-order normalisation (exact restatement of normalize_particle_orders,
// dogm/src/kernel/init_new_particles.cu:30-43, 66-74).
// The reference scales an f32 inclusive scan by v/max and truncates; here the prefix S_j is
// exact and P_j = min(v, floor(double(S_j) * (v / double(S_total)))), with P of the last cell
// forced to v so the counts sum to v exactly.  Cell j owns slots [P_{j-1}, P_j).
// ---------------------------------------------------------------------------------------------
DOGM_HD double order_scale(fx_t total, int v)
{
    return (double)v / fx_to_f64(total);
}

DOGM_HD int order_end(fx_t prefix, double scale, int v)
{
    const double x = floor(fx_to_f64(prefix) * scale);
    if (!(x < (double)v))
        return v;
    if (x <= 0.0)
        return 0;
    return (int)x;
}

// Cells of equal weight when the total mass is zero (the reference divides by zero there).
DOGM_HD int order_end_uniform(int j, int cell_count, int v)
{
    return (int)(((int64_t)(j + 1) * (int64_t)v) / (int64_t)cell_count);
}

}  // namespace dogm_num
