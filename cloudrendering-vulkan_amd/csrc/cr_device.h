// Device-side building blocks of the gfx950 kernels: float3 arithmetic in a
// fixed evaluation order, elementary functions, the xorshift32 stream, grid
// fetches and the linear clamp-to-edge texture lookup.
//
// Every float expression is written in the order the reference GLSL states it
// (shaders/PathTracer.comp) and the library is compiled with
// -ffp-contract=off, so results are reproducible bit for bit by any IEEE-754
// binary32 implementation that evaluates the same sequence.  The reference
// calls GLSL log/exp/sin/cos/pow, whose precision the Vulkan driver defines;
// the functions below are accurate to about one ULP:
//   crd_logf    : x = m 2^e, m in [sqrt(1/2), sqrt(2)], f = m - 1,
//                 log m = f - f^2/2 + f^3 P(f), P of degree 7 fitted to the
//                 relative error on that range (no division),
//                 log x = e ln2 + log m with ln2 = hi + lo.
//   crd_expf    : Cody-Waite reduction x = k ln2 + r, Taylor series to r^8.
//   crd_sincosf : reduction by pi/2 in three parts, Taylor series of sin to
//                 r^11 and cos to r^12 on [-pi/4, pi/4]; NaN for |x| > 2^17.
//   x^1.5       : x * sqrt(x) (sqrt correctly rounded).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "../../include/cr/abi.h"

namespace crd {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

__device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// Correctly rounded binary32 square root.  The hardware square root
// (v_sqrt_f32, which __fsqrt_rn lowers to) is accurate to one ULP only, so its
// result s is corrected.
//  * a in [2^-100, FLT_MAX]: with s- and s+ the neighbours of s, the residuals
//    a - s- s and a - s+ s are multiples of ulp(s)^2, as a is, and the squared
//    midpoints lie a quarter of ulp(s)^2 above those products; so the signs of
//    the two FMA residuals (exact: no underflow in this range) decide whether
//    the square root lies below the lower or above the upper midpoint.
//    Checked exhaustively on the CPU for every such a and every start value
//    within one step of the correctly rounded result (the oracle's
//    or_sqrt_correction_errors).
//  * other inputs: s is compared with the midpoints between s and its two
//    neighbours in binary64, where a midpoint (25 significant bits) and its
//    square (50) are exact, and no binary32 input equals such a square.
__device__ __forceinline__ float fsqrt(float a)
{
    const float s = __fsqrt_rn(a);
    if (a >= 0x1p-100f && a <= 3.40282347e38f) {
        const float sd = __uint_as_float(__float_as_uint(s) - 1u);
        const float su = __uint_as_float(__float_as_uint(s) + 1u);
        const float vd = __builtin_fmaf(-sd, s, a);
        const float vu = __builtin_fmaf(-su, s, a);
        const float r = vd <= 0.0f ? sd : s;
        return vu > 0.0f ? su : r;
    }
    if (!(a > 0.0f) || a == __builtin_inff()) return s;
    const double ad = a;
    const float lo = __uint_as_float(__float_as_uint(s) - 1u);
    const double mlo = 0.5 * ((double)lo + (double)s);
    if (ad < mlo * mlo) return lo;
    const float hi = __uint_as_float(__float_as_uint(s) + 1u);
    const double mhi = 0.5 * ((double)s + (double)hi);
    if (ad > mhi * mhi) return hi;
    return s;
}
__device__ __forceinline__ float length(f3 a) { return fsqrt(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a)
{
    float l = fsqrt(dot(a, a));
    return mk(__fdiv_rn(a.x, l), __fdiv_rn(a.y, l), __fdiv_rn(a.z, l));
}
__device__ __forceinline__ f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

__device__ __forceinline__ float fdiv(float a, float b) { return __fdiv_rn(a, b); }

// log x for a normal, finite x > 0 (scaled by 2^e0 before the call): the
// core of crd_logf without its special cases.  xs_float's values, in
// [2^-32, 1], are such x, so the free-flight draws call this directly.
__device__ __forceinline__ float crd_log_normal(float x, int32_t e0 = 0)
{
    int32_t e = e0;
    const uint32_t b = __float_as_uint(x);
    e += (int32_t)((b >> 23) & 0xffu) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;  // exact: m is within a factor 2 of 1
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

__device__ __forceinline__ float crd_logf(float x)
{
    if (x == 0.0f) return -__builtin_inff();
    if (!(x > 0.0f)) return __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    if (__float_as_uint(x) < 0x00800000u) return crd_log_normal(x * 8388608.0f, -23);
    return crd_log_normal(x);
}

__device__ __forceinline__ float crd_expf(float x)
{
    if (x != x) return x;
    if (x > 88.7f) return __builtin_inff();
    if (x < -87.0f) return 0.0f;
    float k = floorf(x * 1.44269504f + 0.5f);
    float r = (x - k * 0.693145751953125f) - k * 1.42860682030941723212e-6f;
    float p = 2.48015873e-5f;
    p = p * r + 1.98412698e-4f;
    p = p * r + 1.38888889e-3f;
    p = p * r + 8.33333333e-3f;
    p = p * r + 4.16666667e-2f;
    p = p * r + 0.166666667f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int32_t ki = (int32_t)k;
    if (ki > 127) return p * 2.0f * __uint_as_float((uint32_t)(127 + 127) << 23);
    if (ki < -126) return p * __uint_as_float((uint32_t)(ki + 127 + 64) << 23) * 5.42101086e-20f;
    return p * __uint_as_float((uint32_t)(ki + 127) << 23);
}

// Written with selects rather than branches (the quadrant and the domain), so
// the lanes of a wave run it together.
__device__ __forceinline__ void crd_sincosf(float x_in, float* s_out, float* c_out)
{
    const bool inDomain = fabsf(x_in) <= 131072.0f;  // false for NaN
    const float x = inDomain ? x_in : 0.0f;
    float k = floorf(x * 0.636619772f + 0.5f);
    float r = ((x - k * 1.5703125f) - k * 4.83751296997070312e-4f) - k * 7.54978995489188216e-8f;
    float z = r * r;
    float ps = -2.50521084e-8f;
    ps = ps * z + 2.75573192e-6f;
    ps = ps * z - 1.98412698e-4f;
    ps = ps * z + 8.33333333e-3f;
    ps = ps * z - 0.166666667f;
    float sn = r + r * (z * ps);
    float pc = 2.08767570e-9f;
    pc = pc * z - 2.75573192e-7f;
    pc = pc * z + 2.48015873e-5f;
    pc = pc * z - 1.38888889e-3f;
    pc = pc * z + 4.16666667e-2f;
    pc = pc * z - 0.5f;
    float cs = 1.0f + z * pc;
    // quadrant q: (sin, cos) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn)
    const int32_t q = ((int32_t)k) & 3;
    const float sa = (q & 1) ? cs : sn;
    const float ca = (q & 1) ? sn : cs;
    const float so = (q & 2) ? -sa : sa;
    const float co = ((q + 1) & 2) ? -ca : ca;
    *s_out = inDomain ? so : __builtin_nanf("");
    *c_out = inDomain ? co : __builtin_nanf("");
}

// atan(t) for |t| <= tan(pi/12): its Taylor series to t^13 (truncation below
// 3e-10 there).
__device__ __forceinline__ float crd_atan_small(float t)
{
    const float z = t * t;
    float p = 0.0769230769f;      //  1/13
    p = p * z - 0.0909090909f;    // -1/11
    p = p * z + 0.111111111f;     //  1/9
    p = p * z - 0.142857143f;     // -1/7
    p = p * z + 0.2f;             //  1/5
    p = p * z - 0.333333333f;     // -1/3
    return t + t * (z * p);
}

// atan2 by octant: atan of min/max in [0, 1], above tan(pi/12) reduced by
// atan(t) = pi/6 + atan((t sqrt3 - 1) / (t + sqrt3)); then pi/2 - a, pi - a
// and the sign of y.  atan2(0, 0) = 0.
__device__ __forceinline__ float crd_atan2f(float y, float x)
{
    const float ay = fabsf(y), ax = fabsf(x);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx == 0.0f ? 0.0f : fdiv(mn, mx);
    const bool big = t > 0.267949194f;                       // tan(pi/12)
    const float rt = big ? fdiv(t * 1.73205081f - 1.0f, t + 1.73205081f) : t;
    const float at = crd_atan_small(rt);
    float r = big ? 0.523598776f + at : at;                  // pi/6
    r = ay > ax ? 1.57079633f - r : r;
    r = x < 0.0f ? 3.14159265f - r : r;
    return y < 0.0f ? -r : r;
}

__device__ __forceinline__ float crd_acosf(float x) { return crd_atan2f(fsqrt((1.0f - x) * (1.0f + x)), x); }

__device__ __forceinline__ float pow15(float x) { return x * fsqrt(x); }

// ------------------------------------------------------------------ RNG
// xorshift32 (13,17,5) — shaders/PathTracer.comp:220-241.
__device__ __forceinline__ uint32_t xs_next(uint32_t& s)
{
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}
__device__ __forceinline__ float xs_float(uint32_t& s) { return (float)xs_next(s) * 2.3283064365386963e-10f; }
__device__ __forceinline__ void xs_init(uint32_t& s, uint32_t seed)
{
    s = seed;
    const uint32_t n = seed % 7u + 2u;
    for (uint32_t i = 0; i < n; i++) xs_next(s);
}

__device__ __forceinline__ float smoothstep_f(float e0, float e1, float x)
{
    float t = fdiv(x - e0, e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------------ grids
// Density storage.  Layout: dense x-fastest, or 4x4x4 bricks (64 voxels,
// 128 B in FP16) stored brick-major, x fastest inside a brick.
struct GridView {
    const void* data;
    int32_t nx, ny, nz;
    int32_t nbx, nby;   // bricks per axis (bricked layout)
};

enum { LAYOUT_DENSE = 0, LAYOUT_BRICK4 = 1 };

template <int LAYOUT>
__device__ __forceinline__ uint64_t voxel_index(const GridView& g, int32_t x, int32_t y, int32_t z)
{
    if (LAYOUT == LAYOUT_DENSE)
        return (uint64_t)x + (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z);
    uint64_t brick = (uint64_t)(x >> 2) + (uint64_t)g.nbx * ((uint64_t)(y >> 2) + (uint64_t)g.nby * (uint64_t)(z >> 2));
    return (brick << 6) + (uint64_t)((x & 3) + ((y & 3) << 2) + ((z & 3) << 4));
}

// Two binary32 values -> two bf16 packed low/high (v_cvt_pk_bf16_f32, round
// to nearest even; the learned network's activations and gradients).
typedef __bf16 crd_v2bf __attribute__((ext_vector_type(2)));
typedef float crd_v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bf16x2_rne(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((crd_v2f{lo, hi}), crd_v2bf));
}

// binary32 -> binary16 bits, round to nearest, ties to even, in integer
// arithmetic.  Handing a float expression straight to a half conversion lets
// the compiler fold its last multiply into v_fma_mixlo_f16, which rounds the
// exact product to half once, where IEEE semantics first round it to
// binary32; the two differ whenever the binary32 value is a tie between two
// halves.
__device__ __forceinline__ uint16_t f32_to_f16_rne(float f)
{
    const uint32_t b = __float_as_uint(f);
    const uint32_t sign = (b >> 16) & 0x8000u;
    const uint32_t a = b & 0x7fffffffu;
    if (a >= 0x7f800000u) return (uint16_t)(sign | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   // >= 65520 rounds to infinity
    uint32_t h, rem, half;
    if (a < 0x38800000u) {                                      // below 2^-14: a subnormal half or zero
        const uint32_t e = a >> 23;
        if (e < 102u) return (uint16_t)sign;                    // below 2^-25 (a tie at 2^-25 goes to 0)
        const uint32_t m = (a & 0x7fffffu) | 0x800000u;         // value = m 2^(e-150) = (m >> (126-e)) 2^-24 + rest
        const uint32_t sh = 126u - e;                           // 14..24
        h = m >> sh;
        rem = m & ((1u << sh) - 1u);
        half = 1u << (sh - 1u);
    } else {
        h = (((a >> 23) - 112u) << 10) | ((a & 0x7fffffu) >> 13);
        rem = a & 0x1fffu;
        half = 0x1000u;
    }
    if (rem > half || (rem == half && (h & 1u))) h += 1u;       // a carry moves into the exponent correctly
    return (uint16_t)(sign | h);
}

template <typename T> __device__ __forceinline__ float load_voxel(const T* p, uint64_t i);
template <> __device__ __forceinline__ float load_voxel<float>(const float* p, uint64_t i) { return p[i]; }
template <> __device__ __forceinline__ float load_voxel<__half>(const __half* p, uint64_t i)
{
    const unsigned short* u = reinterpret_cast<const unsigned short*>(p);
    return __half2float(__ushort_as_half(__ldg(u + i)));
}

__device__ __forceinline__ void texel_axis(float n, int32_t N, int32_t& i0, int32_t& i1, float& a)
{
    float u = n * (float)N - 0.5f;
    u = fminf(fmaxf(u, -1.0f), (float)N);
    float fl = floorf(u);
    a = u - fl;
    int32_t i = (int32_t)fl;
    i0 = clampi(i, 0, N - 1);
    i1 = clampi(i + 1, 0, N - 1);
}

// texture(sampler3D, n).x with VK_FILTER_LINEAR + CLAMP_TO_EDGE
// (Initializers.cpp:408-418), exact-float weights.
template <typename T, int LAYOUT>
__device__ __forceinline__ float trilinear(const GridView& g, f3 n)
{
    int32_t x0, x1, y0, y1, z0, z1;
    float a, b, c;
    texel_axis(n.x, g.nx, x0, x1, a);
    texel_axis(n.y, g.ny, y0, y1, b);
    texel_axis(n.z, g.nz, z0, z1, c);
    const T* p = reinterpret_cast<const T*>(g.data);
    // The voxel index is a sum of one term per axis (grids hold < 2^32 voxels):
    // dense  x + nx y + nx ny z;  bricks ((x>>2)<<6 | x&3) + (y>>2) 64 nbx +
    // (y&3) 4 + (z>>2) 64 nbx nby + (z&3) 16.
    uint32_t ax0, ax1, by0, by1, cz0, cz1;
    if (LAYOUT == LAYOUT_DENSE) {
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        ax0 = (uint32_t)x0; ax1 = (uint32_t)x1;
        by0 = (uint32_t)y0 * sy; by1 = (uint32_t)y1 * sy;
        cz0 = (uint32_t)z0 * sz; cz1 = (uint32_t)z1 * sz;
    } else {
        const uint32_t sy = (uint32_t)g.nbx << 6, sz = ((uint32_t)g.nbx * (uint32_t)g.nby) << 6;
        ax0 = (((uint32_t)x0 >> 2) << 6) | ((uint32_t)x0 & 3u);
        ax1 = (((uint32_t)x1 >> 2) << 6) | ((uint32_t)x1 & 3u);
        by0 = ((uint32_t)y0 >> 2) * sy + (((uint32_t)y0 & 3u) << 2);
        by1 = ((uint32_t)y1 >> 2) * sy + (((uint32_t)y1 & 3u) << 2);
        cz0 = ((uint32_t)z0 >> 2) * sz + (((uint32_t)z0 & 3u) << 4);
        cz1 = ((uint32_t)z1 >> 2) * sz + (((uint32_t)z1 & 3u) << 4);
    }
    const uint32_t yz00 = by0 + cz0, yz10 = by1 + cz0, yz01 = by0 + cz1, yz11 = by1 + cz1;
    float v000 = load_voxel<T>(p, ax0 + yz00);
    float v100 = load_voxel<T>(p, ax1 + yz00);
    float v010 = load_voxel<T>(p, ax0 + yz10);
    float v110 = load_voxel<T>(p, ax1 + yz10);
    float v001 = load_voxel<T>(p, ax0 + yz01);
    float v101 = load_voxel<T>(p, ax1 + yz01);
    float v011 = load_voxel<T>(p, ax0 + yz11);
    float v111 = load_voxel<T>(p, ax1 + yz11);
    float oa = 1.0f - a, ob = 1.0f - b, oc = 1.0f - c;
    float c00 = v000 * oa + v100 * a;
    float c10 = v010 * oa + v110 * a;
    float c01 = v001 * oa + v101 * a;
    float c11 = v011 * oa + v111 * a;
    float c0 = c00 * ob + c10 * b;
    float c1 = c01 * ob + c11 * b;
    return c0 * oc + c1 * c;
}

// sampleCloud's normalisation — PathTracer.comp:115-119.
__device__ __forceinline__ f3 cloud_norm(const cr_cloud_properties& cp, f3 p)
{
    return mk(fdiv(p.x - cp.bounds[0][0], cp.bounds[1][0] - cp.bounds[0][0]),
              fdiv(p.y - cp.bounds[0][1], cp.bounds[1][1] - cp.bounds[0][1]),
              fdiv(p.z - cp.bounds[0][2], cp.bounds[1][2] - cp.bounds[0][2]));
}

// sampleShadowVolume's light-space map — PathTracer.comp:121-127.
__device__ __forceinline__ f3 sv_norm(const cr_shadow_volume_properties& s, f3 p)
{
    const float* M = s.basisChange;
    float v0 = p.x - s.bounds[0][0], v1 = p.y - s.bounds[0][1], v2 = p.z - s.bounds[0][2];
    float v3 = 1.0f - s.bounds[0][3];
    float e0 = s.bounds[1][0] - s.bounds[0][0], e1 = s.bounds[1][1] - s.bounds[0][1];
    float e2 = s.bounds[1][2] - s.bounds[0][2], e3 = s.bounds[1][3] - s.bounds[0][3];
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float num = ((M[0 * 4 + i] * v0 + M[1 * 4 + i] * v1) + M[2 * 4 + i] * v2) + M[3 * 4 + i] * v3;
        float den = ((M[0 * 4 + i] * e0 + M[1 * 4 + i] * e1) + M[2 * 4 + i] * e2) + M[3 * 4 + i] * e3;
        r[i] = fdiv(num, den);
    }
    return mk(r[0], r[1], r[2]);
}

// intersectCloud — PathTracer.comp:154-184.
// ... with the reciprocal direction given (invdir = 1 / dir).
__device__ __forceinline__ bool intersect_box_inv(const float* b0, const float* b1, f3 pos, f3 inv, float& tmin,
                                                  float& tmax)
{
    bool sx = inv.x < 0.0f, sy = inv.y < 0.0f, sz = inv.z < 0.0f;
    tmin = ((sx ? b1[0] : b0[0]) - pos.x) * inv.x;
    tmax = ((sx ? b0[0] : b1[0]) - pos.x) * inv.x;
    float tymin = ((sy ? b1[1] : b0[1]) - pos.y) * inv.y;
    float tymax = ((sy ? b0[1] : b1[1]) - pos.y) * inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = ((sz ? b1[2] : b0[2]) - pos.z) * inv.z;
    float tzmax = ((sz ? b0[2] : b1[2]) - pos.z) * inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return true;
}

__device__ __forceinline__ f3 recip(f3 d) { return mk(fdiv(1.0f, d.x), fdiv(1.0f, d.y), fdiv(1.0f, d.z)); }

__device__ __forceinline__ bool intersect_box(const float* b0, const float* b1, f3 pos, f3 dir, float& tmin, float& tmax)
{
    return intersect_box_inv(b0, b1, pos, recip(dir), tmin, tmax);
}

// isInCloud — ShadowVolume.comp:34-42.
__device__ __forceinline__ bool in_box(const float* b0, const float* b1, f3 p)
{
    return !(p.x < b0[0] || p.y < b0[1] || p.z < b0[2] || p.x > b1[0] || p.y > b1[1] || p.z > b1[2]);
}

constexpr float PI_F = 3.14159265359f;
constexpr float INV_4PI = 1.0f / (4.0f * PI_F);

// sampleBackground — PathTracer.comp:104-113 (mix(a,b,t) = a(1-t) + b t), or the constant sky.
__device__ __forceinline__ f4 background(const cr_render_ext& ext, f3 dir)
{
    f4 col;
    if (ext.skyMode == CR_SKY_CONSTANT) {
        col.x = ext.skyColor[0]; col.y = ext.skyColor[1]; col.z = ext.skyColor[2]; col.w = ext.skyColor[3];
        return col;
    }
    const float C[5][4] = {
        {0.00f, 0.0f, 0.02f, 1.0f}, {0.01f, 0.05f, 0.2f, 1.0f}, {0.7f, 0.9f, 1.0f, 1.0f},
        {0.1f, 0.3f, 1.0f, 1.0f}, {0.01f, 0.1f, 0.7f, 1.0f}};
    const float D[5] = {-1.0f, -0.04f, 0.0f, 0.5f, 1.0f};
    col.x = C[0][0]; col.y = C[0][1]; col.z = C[0][2]; col.w = C[0][3];
#pragma unroll
    for (int i = 1; i < 5; i++) {
        float t = smoothstep_f(D[i - 1], D[i], dir.y);
        float ot = 1.0f - t;
        col.x = col.x * ot + C[i][0] * t;
        col.y = col.y * ot + C[i][1] * t;
        col.z = col.z * ot + C[i][2] * t;
        col.w = col.w * ot + C[i][3] * t;
    }
    return col;
}

__device__ __forceinline__ bool isotropic(const cr_parameters& p) { return (p.isotropic & 0xffu) != 0; }

// calculatePDF / samplePhase — PathTracer.comp:247-274.
__device__ __forceinline__ float phase_pdf(const cr_parameters& p, f3 a, f3 b)
{
    if (isotropic(p)) return INV_4PI;
    float cosT = dot(a, b);
    return fdiv(INV_4PI * p.phaseOneMinusG2, pow15(p.phaseOnePlusG2 - (2.0f * p.phaseG) * cosT));
}

// scatterRay + createOrthonormalBasis — PathTracer.comp:197-206, :277-311.
__device__ __forceinline__ f3 scatter(const cr_parameters& p, f3 n, uint32_t& rng)
{
    if (isotropic(p)) {
        f3 d;
        float xi = xs_float(rng);
        d.z = xi * 2.0f - 1.0f;
        float st = 1.0f - d.z * d.z;
        if (st > 0.0f) {
            st = fsqrt(fmaxf(0.0f, st));
            xi = xs_float(rng);
            float phi = (xi * 2.0f) * PI_F;
            float sp, cp;
            crd_sincosf(phi, &sp, &cp);
            d.x = st * cp;
            d.y = st * sp;
        } else {
            d.x = 0.0f;
            d.y = 0.0f;
        }
        return d;
    }
    float phi = (xs_float(rng) * 2.0f) * PI_F;
    float xi = xs_float(rng);
    float sq = fdiv(p.phaseOneMinusG2, (1.0f - p.phaseG) + (2.0f * p.phaseG) * xi);
    float cosT = p.phaseOneOver2G * (p.phaseOnePlusG2 - sq * sq);
    float sinT = fsqrt(fmaxf(0.0f, 1.0f - cosT * cosT));
    float sz = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = fdiv(n.y, 1.0f + fabsf(n.z));
    float b = n.y * a;
    float c = -n.x * a;
    f3 t0 = mk(n.z + sz * b, sz * c, -n.x);
    f3 t1 = mk(c, 1.0f - b, -sz * n.y);
    float sp, cp;
    crd_sincosf(phi, &sp, &cp);
    float k0 = sinT * cp, k1 = sinT * sp;
    return add(add(mul(t0, k0), mul(t1, k1)), mul(n, cosT));
}

// getCameraRay — PathTracer.comp:144-151; dx, dy are pixel offsets from the centre.
__device__ __forceinline__ void camera_ray(const cr_camera_properties& c, float dx, float dy, f3& pos, f3& dir)
{
    f3 d;
    d.x = (c.forward[0] * c.nearPlane + (c.right[0] * c.pixelSizeX) * dx) - (c.up[0] * c.pixelSizeY) * dy;
    d.y = (c.forward[1] * c.nearPlane + (c.right[1] * c.pixelSizeX) * dx) - (c.up[1] * c.pixelSizeY) * dy;
    d.z = (c.forward[2] * c.nearPlane + (c.right[2] * c.pixelSizeX) * dx) - (c.up[2] * c.pixelSizeY) * dy;
    pos = add(d, ld3(c.position));
    dir = normalize(d);
}

// ----------------------------------------------------------- wave helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Sum of a 64-bit value over the active lanes of the wave, returned to all.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

} // namespace crd
