/*
 * ORACLE — test infrastructure only.
 *
 * Exhaustive CPU check of the rounding correction the kernels apply to the
 * hardware square root (cloudrendering-vulkan_amd/csrc/cr_device.h fsqrt):
 * for an input a in [2^-100, FLT_MAX] and a start value s within one step of
 * the correctly rounded root, s- and s+ its neighbours,
 *     r = (a - s- s <= 0) ? s- : s;   r = (a - s+ s > 0) ? s+ : r;
 * with each residual one fused multiply-add.  The restatement below is the
 * same sequence; sqrtf is IEEE-754's correctly rounded square root.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

extern "C" {

static inline float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static inline float sqrt_fma_correct(float s, float a)
{
    const float sd = bits_f(f_bits(s) - 1u), su = bits_f(f_bits(s) + 1u);
    const float vd = fmaf(-sd, s, a);
    const float vu = fmaf(-su, s, a);
    const float r = vd <= 0.0f ? sd : s;
    return vu > 0.0f ? su : r;
}

/* Number of (a, s) pairs, a over the bit patterns [lo_bits, hi_bits] and s the
 * correctly rounded root or one of its neighbours, for which the correction
 * does not return the correctly rounded root. */
uint64_t or_sqrt_correction_errors(uint32_t lo_bits, uint32_t hi_bits)
{
    uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = (int64_t)lo_bits; i <= (int64_t)hi_bits; i++) {
        const float a = bits_f((uint32_t)i);
        const float rn = sqrtf(a);
        for (int d = -1; d <= 1; d++)
            bad += sqrt_fma_correct(bits_f(f_bits(rn) + (uint32_t)d), a) != rn;
    }
    return bad;
}

} // extern "C"
