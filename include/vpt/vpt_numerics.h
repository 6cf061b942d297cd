/*
 * vpt_numerics.h — arithmetic definitions shared by the gfx950 kernels and the
 * CPU oracle.  Everything here is the *definition* of an operation whose
 * result must be bit-identical on host and device, so every function is
 * written as a fixed sequence of IEEE-754 binary32 operations (compile with
 * -ffp-contract=off on both sides; no fast-math).
 *
 * Contents
 *   - pcg_hash: integer hash of Jarzynski & Olano, "Hash Functions for GPU
 *     Rendering", JCGT 9(3), 2020 (PCG-RXS-M-XS output on one LCG step).
 *     Replaces the reference's wall-clock driven fract(sin()) hash2() and
 *     blue-noise tables (DiffuseRayTraceFrag.glsl:945-967, :807-820).
 *   - vpt_u32_to_unit: exact map u32 -> [0,1) using the top 24 bits.
 *   - vpt_sincos_2pi: sin/cos of 2*pi*u for u in [0,1).  Quadrant reduction
 *     on u (exact in binary) followed by least-squares polynomials fitted on
 *     Chebyshev nodes of [0,1] for sin(pi/2 f)/f and cos(pi/2 f) in f^2
 *     (coefficients fitted for this project; max abs error ~2e-7).
 *   - vpt_expf: Cody & Waite range reduction (x = k ln2 + r, ln2 split into
 *     a short high part and a low part) followed by a degree-6 polynomial for
 *     exp(r) on [-ln2/2, ln2/2], least-squares fitted on Chebyshev nodes
 *     (max rel error ~1e-7), and exponent reconstruction by bit assembly.
 *   - vpt_powi: x^n for integer n by binary exponentiation (fixed order).
 *   - vpt_logf: natural log.  x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 *     log m = 2 atanh(s), s = (m-1)/(m+1), summed as the odd series
 *     2(s + s^3/3 + ... + s^9/9) (|s| <= 0.1716, truncation < 1e-9), plus
 *     e ln2 with ln2 split as in vpt_expf.
 *   - vpt_powf: x^y = vpt_expf(y vpt_logf(x)) for x > 0 (GLSL pow is
 *     undefined for x < 0; 0^y = 0 for y > 0, x^0 = 1).
 *   - vpt_exp2f: 2^x = vpt_expf(x ln2).
 *   - vpt_f32_to_f16 / vpt_f16_to_f32: IEEE binary16 conversion with
 *     round-to-nearest-even, the rounding of a GL RGBA16F render target.
 *   - vpt_unorm8_encode / vpt_unorm8_decode: GL unorm8 conversions
 *     (c / 255 on read, round(clamp(x,0,1) * 255) on write).
 *   - vpt_df_step: the reference's DF step table
 *     floor(ToConservativeEuclidean(floor-exact unorm8 * 255))
 *     (InitialRayTraceFrag.glsl:89-102, :328-330).
 */
#ifndef VPT_NUMERICS_H
#define VPT_NUMERICS_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define VPT_HD __host__ __device__ inline
#else
#define VPT_HD static inline
#endif

/* Bit casts without type punning through unions (valid on both sides). */
VPT_HD uint32_t vpt_f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
VPT_HD float vpt_u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }

/* ---- random numbers --------------------------------------------------- */

VPT_HD uint32_t vpt_pcg_hash(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

/* Top 24 bits scaled by 2^-24: exact, in [0, 1 - 2^-24]. */
VPT_HD float vpt_u32_to_unit(uint32_t v)
{
    return (float)(v >> 8) * 5.9604644775390625e-08f;
}

/* ---- trigonometry ------------------------------------------------------ */

/* sin(2*pi*u), cos(2*pi*u) for u in [0, 1). */
VPT_HD void vpt_sincos_2pi(float u, float* s_out, float* c_out)
{
    float t = u * 4.0f;                 /* exact */
    float q = floorf(t);
    float f = t - q;                    /* exact, in [0,1) */
    int quadrant = ((int)q) & 3;
    float g = f * f;
    /* sin(pi/2 f) = f * P(f^2) */
    float ps = 0.00015186281234491616f;
    ps = ps * g + -0.004674552474170923f;
    ps = ps * g + 0.07969020307064056f;
    ps = ps * g + -0.6459638476371765f;
    ps = ps * g + 1.5707963705062866f;
    float s = ps * f;
    /* cos(pi/2 f) = Q(f^2) */
    float pc = -2.3848328055464663e-05f;
    pc = pc * g + 0.0009177866159006953f;
    pc = pc * g + -0.020862746983766556f;
    pc = pc * g + 0.2536693513393402f;
    pc = pc * g + -1.2337005138397217f;
    pc = pc * g + 1.0f;
    float c = pc;
    float rs, rc;
    if (quadrant == 0) { rs = s; rc = c; }
    else if (quadrant == 1) { rs = c; rc = -s; }
    else if (quadrant == 2) { rs = -s; rc = -c; }
    else { rs = -c; rc = s; }
    *s_out = rs;
    *c_out = rc;
}

/* ---- exponential ------------------------------------------------------- */

VPT_HD float vpt_expf(float x)
{
    if (!(x == x)) return x;                 /* NaN */
    if (x > 88.72283172607421875f) return vpt_u2f(0x7f800000u);
    if (x < -87.33654022216796875f) return 0.0f;
    const float LOG2E = 1.44269502162933349609375f;
    const float LN2_HI = 0.693145751953125f;          /* 0x3F317200: 16 low bits zero */
    const float LN2_LO = 1.428606765330187045e-06f;
    float kf = floorf(x * LOG2E + 0.5f);
    float r = x - kf * LN2_HI;
    r = r - kf * LN2_LO;
    float p = 0.0013829420786350965f;
    p = p * r + 0.008374771103262901f;
    p = p * r + 0.0416683591902256f;
    p = p * r + 0.1666642129421234f;
    p = p * r + 0.49999991059303284f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k = (int)kf;                          /* in [-126, 128] */
    int k1 = k / 2;
    int k2 = k - k1;
    float s1 = vpt_u2f((uint32_t)(k1 + 127) << 23);
    float s2 = vpt_u2f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

/* Natural log by the atanh series (see the header comment). */
VPT_HD float vpt_logf(float x)
{
    if (!(x == x)) return x;                                  /* NaN */
    if (x < 0.0f) return vpt_u2f(0x7fc00000u);                /* NaN */
    if (x == 0.0f) return vpt_u2f(0xff800000u);               /* -inf */
    if (x == vpt_u2f(0x7f800000u)) return x;                  /* +inf */
    uint32_t u = vpt_f2u(x);
    int e = (int)((u >> 23) & 0xffu) - 127;
    if (e == -127) {                                          /* subnormal: scale by 2^23 */
        u = vpt_f2u(x * 8388608.0f);
        e = (int)((u >> 23) & 0xffu) - 127 - 23;
    }
    float m = vpt_u2f((u & 0x007fffffu) | 0x3f800000u);       /* [1, 2) */
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float LN2_HI = 0.693145751953125f;
    const float LN2_LO = 1.428606765330187045e-06f;
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.111111111111f;                                /* 1/9 */
    p = p * s2 + 0.142857142857f;                             /* 1/7 */
    p = p * s2 + 0.2f;                                        /* 1/5 */
    p = p * s2 + 0.333333333333f;                             /* 1/3 */
    float lm = (s + s) + (s + s) * (s2 * p);
    float fe = (float)e;
    return fe * LN2_HI + (fe * LN2_LO + lm);
}

/* x^y for x >= 0 (see the header comment). */
VPT_HD float vpt_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : vpt_u2f(0x7f800000u);
    if (x == 1.0f) return 1.0f;
    return vpt_expf(y * vpt_logf(x));
}

/* 2^x (see the header comment). */
VPT_HD float vpt_exp2f(float x) { return vpt_expf(x * 0.693147180559945309f); }

/* x^n, n >= 0, by binary exponentiation (squares taken in a fixed order). */
VPT_HD float vpt_powi(float x, int n)
{
    float result = 1.0f;
    float base = x;
    while (n > 0) {
        if (n & 1) result = result * base;
        n >>= 1;
        if (n > 0) base = base * base;
    }
    return result;
}

/* ---- storage formats --------------------------------------------------- */

VPT_HD uint16_t vpt_f32_to_f16(float f)
{
    uint32_t x = vpt_f2u(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t mag = x & 0x7fffffffu;
    if (mag >= 0x7f800000u) {                         /* inf or NaN */
        return (uint16_t)(sign | 0x7c00u | (mag > 0x7f800000u ? 0x200u : 0u));
    }
    if (mag >= 0x477ff000u) {                         /* rounds to >= 65520 -> inf */
        return (uint16_t)(sign | 0x7c00u);
    }
    if (mag < 0x38800000u) {                          /* result is subnormal or zero */
        if (mag < 0x33000000u) return (uint16_t)sign; /* < 2^-25: rounds to zero */
        uint32_t e = mag >> 23;                       /* 102..112 */
        uint32_t m = (mag & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;                    /* 14..24 */
        uint32_t half_m = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1u);
        if (rem > halfway || (rem == halfway && (half_m & 1u))) half_m += 1u;
        return (uint16_t)(sign | half_m);
    }
    uint32_t h = ((mag >> 13) - (112u << 10));        /* rebias exponent */
    uint32_t rem = mag & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
    return (uint16_t)(sign | h);
}

VPT_HD float vpt_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    if (e == 0u) {
        /* subnormal: m * 2^-24, exact in binary32 */
        float v = (float)m * 5.9604644775390625e-08f;
        return sign ? -v : v;
    }
    if (e == 31u) return vpt_u2f(sign | 0x7f800000u | (m << 13));
    return vpt_u2f(sign | ((e + 112u) << 23) | (m << 13));
}

/* Round a binary32 value through binary16 storage. */
VPT_HD float vpt_round_f16(float f) { return vpt_f16_to_f32(vpt_f32_to_f16(f)); }

VPT_HD uint8_t vpt_unorm8_encode(float x)
{
    float c = x;
    if (!(c > 0.0f)) c = 0.0f;               /* also maps NaN to 0 */
    if (c > 1.0f) c = 1.0f;
    return (uint8_t)floorf(c * 255.0f + 0.5f);
}

/* c / 255 correctly rounded, as a product and one fused correction step
 * (q = c * fl(1/255); e = c - q * 255 exactly; q + e * fl(1/255)); equal to
 * the IEEE quotient for every byte (checked exhaustively by the tests) and
 * four instructions instead of a full division on the GPU. */
VPT_HD float vpt_unorm8_decode(uint8_t c)
{
    const float x = (float)c;
    const float r = 0.0039215688593685627f;   /* fl(1/255) */
    float q = x * r;
    float e = fmaf(-q, 255.0f, x);
    return fmaf(e, r, q);
}

/* DF byte (Manhattan distance 0..254) -> conservative step count
 * E = int(floor(M == 1 ? 1 : M * 0.57735026918)) with M read back through
 * the unorm8 texture path (exact for every byte, checked by the tests). */
VPT_HD int vpt_df_step(int df_byte)
{
    float m = ((float)df_byte / 255.0f) * 255.0f;
    float e = (m == 1.0f) ? 1.0f : m * 0.57735026918f;
    return (int)floorf(e);
}

#endif /* VPT_NUMERICS_H */
