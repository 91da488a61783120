/*
 * ORACLE — test infrastructure only.  Nothing in the product links this.
 *
 * Elementary functions of the oracle, written from the mathematics so that
 * the CPU oracle and the gfx950 kernels can evaluate the same sequence of
 * IEEE-754 binary32 operations and agree bit for bit.  The reference calls
 * GLSL log/exp/sin/cos/pow, whose accuracy the Vulkan driver defines
 * (shaders/PathTracer.comp:249, :289-290, :307-308, :338;
 * shaders/ShadowVolume.comp:92); any implementation within a few ULP is a
 * faithful restatement.  Each function below is pinned against double
 * precision libm in tests/test_oracle.py.
 *
 *  o_logf   : x = m * 2^e with m in [sqrt(1/2), sqrt(2)], f = m - 1,
 *             log m = f - f^2/2 + f^3 P(f) with P of degree 7, its
 *             coefficients fitted by the build (least squares weighted to the
 *             relative error of log m, refined by Lawson iterations toward
 *             the minimax fit, on f in [sqrt(1/2)-1, sqrt(2)-1]; no
 *             division), log x = e ln2 + log m with ln2 split in a high and a
 *             low part (Cody & Waite's reduction); <= 1.3 ULP.
 *  o_expf   : x = k ln2 + r, |r| <= ln2/2 (Cody-Waite reduction), Taylor
 *             series of e^r to r^8, scaled by 2^k.
 *  o_sincosf: quadrant reduction by pi/2 in three parts (Cody-Waite; the
 *             split 1.5703125 + 4.837512969970703e-4 + 7.54978995489188e-8
 *             is twice the pi/4 split of Cephes' sinf, DP1..DP3), Taylor
 *             series on [-pi/4, pi/4] (sin to r^11, cos to r^12);
 *             defined for |x| <= 2^17, NaN beyond (the three-part reduction
 *             loses accuracy there, and the quadrant would need a float to
 *             int conversion that C leaves undefined for huge values).
 *  o_pow15  : x^1.5 = x * sqrt(x) (sqrt is correctly rounded in IEEE-754).
 *  o_atan2f : octant reduction to atan(t), t in [0, 1]; above tan(pi/12)
 *             atan(t) = pi/6 + atan((t sqrt3 - 1) / (t + sqrt3)); Taylor series
 *             to t^13 on |t| <= tan(pi/12).  o_acosf(x) = atan2(sqrt((1-x)(1+x)), x).
 *
 * Compile with -ffp-contract=off and without -ffast-math.
 */
#ifndef CR_ORACLE_OMATH_H
#define CR_ORACLE_OMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint32_t o_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float o_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline float o_logf(float x)
{
    if (x == 0.0f) return -INFINITY;
    if (!(x > 0.0f)) return NAN;
    if (x == INFINITY) return INFINITY;
    int32_t e = 0;
    uint32_t b = o_f2u(x);
    if (b < 0x00800000u) {            /* subnormal: scale by 2^23 (exact) */
        x = x * 8388608.0f;
        b = o_f2u(x);
        e = -23;
    }
    e += (int32_t)((b >> 23) & 0xffu) - 127;
    float m = o_u2f((b & 0x007fffffu) | 0x3f800000u); /* [1,2) */
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;  /* exact: m is within a factor 2 of 1 */
    float z = f * f;
    float p = -0.07369332015514374f;
    p = p * f + 0.12623557448387146f;
    p = p * f + -0.1320105493068695f;
    p = p * f + 0.14226537942886353f;
    p = p * f + -0.166214257478714f;
    p = p * f + 0.19999857246875763f;
    p = p * f + -0.2500085234642029f;
    p = p * f + 0.3333335220813751f;
    float logm = (f - 0.5f * z) + (z * f) * p;
    float fe = (float)e;
    return fe * 0.693145751953125f + (logm + fe * 1.42860682030941723212e-6f);
}

static inline float o_expf(float x)
{
    if (x != x) return x;
    if (x > 88.7f) return INFINITY;
    if (x < -87.0f) return 0.0f;      /* results below 2^-126 are returned as 0 */
    float k = floorf(x * 1.44269504f + 0.5f);
    float r = (x - k * 0.693145751953125f) - k * 1.42860682030941723212e-6f;
    float p = 2.48015873e-5f;         /* 1/8! */
    p = p * r + 1.98412698e-4f;       /* 1/7! */
    p = p * r + 1.38888889e-3f;       /* 1/6! */
    p = p * r + 8.33333333e-3f;       /* 1/5! */
    p = p * r + 4.16666667e-2f;       /* 1/4! */
    p = p * r + 0.166666667f;         /* 1/3! */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int32_t ki = (int32_t)k;          /* in [-126, 128] */
    if (ki > 127) return p * 2.0f * o_u2f((uint32_t)(127 + 127) << 23);
    if (ki < -126) return p * o_u2f((uint32_t)(ki + 127 + 64) << 23) * 5.42101086e-20f; /* 2^-64 */
    return p * o_u2f((uint32_t)(ki + 127) << 23);
}

static inline void o_sincosf(float x, float* s_out, float* c_out)
{
    if (!(fabsf(x) <= 131072.0f)) {
        *s_out = NAN;
        *c_out = NAN;
        return;
    }
    float k = floorf(x * 0.636619772f + 0.5f);           /* round(x / (pi/2)) */
    float r = ((x - k * 1.5703125f) - k * 4.83751296997070312e-4f) - k * 7.54978995489188216e-8f;
    float z = r * r;
    float ps = -2.50521084e-8f;       /* -1/11! */
    ps = ps * z + 2.75573192e-6f;     /*  1/9!  */
    ps = ps * z - 1.98412698e-4f;     /* -1/7!  */
    ps = ps * z + 8.33333333e-3f;     /*  1/5!  */
    ps = ps * z - 0.166666667f;       /* -1/3!  */
    float sn = r + r * (z * ps);
    float pc = 2.08767570e-9f;        /*  1/12! */
    pc = pc * z - 2.75573192e-7f;     /* -1/10! */
    pc = pc * z + 2.48015873e-5f;     /*  1/8!  */
    pc = pc * z - 1.38888889e-3f;     /* -1/6!  */
    pc = pc * z + 4.16666667e-2f;     /*  1/4!  */
    pc = pc * z - 0.5f;
    float cs = 1.0f + z * pc;
    int32_t q = ((int32_t)k) & 3;
    if (q == 0) { *s_out = sn;  *c_out = cs; }
    else if (q == 1) { *s_out = cs;  *c_out = -sn; }
    else if (q == 2) { *s_out = -sn; *c_out = -cs; }
    else { *s_out = -cs; *c_out = sn; }
}

/* atan(t) for |t| <= tan(pi/12): Taylor series to t^13 (truncation < 3e-10). */
static inline float o_atan_small(float t)
{
    const float z = t * t;
    float p = 0.0769230769f;      /*  1/13 */
    p = p * z - 0.0909090909f;    /* -1/11 */
    p = p * z + 0.111111111f;     /*  1/9  */
    p = p * z - 0.142857143f;     /* -1/7  */
    p = p * z + 0.2f;             /*  1/5  */
    p = p * z - 0.333333333f;     /* -1/3  */
    return t + t * (z * p);
}

/* atan2 by octant: atan of min/max in [0, 1], above tan(pi/12) reduced by
 * atan(t) = pi/6 + atan((t sqrt3 - 1) / (t + sqrt3)); then pi/2 - a, pi - a and
 * the sign of y.  atan2(0, 0) = 0.  acos and atan2 <= 3 ULP (tests/test_oracle.py). */
static inline float o_atan2f(float y, float x)
{
    const float ay = fabsf(y), ax = fabsf(x);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx == 0.0f ? 0.0f : mn / mx;
    const int big = t > 0.267949194f;
    const float rt = big ? (t * 1.73205081f - 1.0f) / (t + 1.73205081f) : t;
    const float at = o_atan_small(rt);
    float r = big ? 0.523598776f + at : at;
    r = ay > ax ? 1.57079633f - r : r;
    r = x < 0.0f ? 3.14159265f - r : r;
    return y < 0.0f ? -r : r;
}

static inline float o_acosf(float x) { return o_atan2f(sqrtf((1.0f - x) * (1.0f + x)), x); }

static inline float o_pow15(float x) { return x * sqrtf(x); }

#endif /* CR_ORACLE_OMATH_H */
