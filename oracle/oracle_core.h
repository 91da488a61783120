/*
 * ORACLE — test infrastructure only.
 *
 * Shared inline helpers of the oracle: float3 arithmetic in a fixed
 * evaluation order, the xorshift32 stream, grid fetches and the trilinear
 * texture lookup.  Evaluation order is part of the contract with the HIP
 * kernels (both sides are compiled with -ffp-contract=off).
 */
#ifndef CR_ORACLE_CORE_H
#define CR_ORACLE_CORE_H

#include <math.h>
#include <stdint.h>
#include "omath.h"
#include "oracle.h"

namespace orc {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float length(f3 a) { return sqrtf(dot(a, a)); }
static inline f3 normalize(f3 a) { float l = sqrtf(dot(a, a)); return mk(a.x / l, a.y / l, a.z / l); }
static inline f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

/* xorshift32 (13,17,5) — shaders/PathTracer.comp:220-228 */
static inline uint32_t xs_next(uint32_t& s)
{
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}
/* generateRandomNumber — PathTracer.comp:230-233: float(u32) / 2^32 */
static inline float xs_float(uint32_t& s) { return (float)xs_next(s) / 4294967296.0f; }
/* initializeRandom — PathTracer.comp:235-241 */
static inline void xs_init(uint32_t& s, uint32_t seed)
{
    s = seed;
    for (uint32_t i = 0; i < seed % 7u + 2u; i++) xs_next(s);
}

static inline float f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    if (e == 0) {
        float v = (float)m * 5.9604644775390625e-8f; /* m * 2^-24, exact */
        return sign ? -v : v;
    }
    if (e == 31) return o_u2f(sign | 0x7f800000u | (m << 13));
    return o_u2f(sign | ((e + 112u) << 23) | (m << 13));
}

struct Grid {
    const float* f32;
    const uint16_t* f16;
    int32_t nx, ny, nz;
    inline float at(int32_t x, int32_t y, int32_t z) const
    {
        uint64_t i = (uint64_t)x + (uint64_t)nx * ((uint64_t)y + (uint64_t)ny * (uint64_t)z);
        return f32 ? f32[i] : f16_to_f32(f16[i]);
    }
};

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* One axis of the linear, clamp-to-edge texel footprint: u = n*N - 1/2. */
static inline void texel_axis(float n, int32_t N, int32_t& i0, int32_t& i1, float& a)
{
    float u = n * (float)N - 0.5f;
    u = fminf(fmaxf(u, -1.0f), (float)N);
    float fl = floorf(u);
    a = u - fl;
    int32_t i = (int32_t)fl;
    i0 = clampi(i, 0, N - 1);
    i1 = clampi(i + 1, 0, N - 1);
}

/* texture(sampler3D, n).x with VK_FILTER_LINEAR + CLAMP_TO_EDGE
 * (Initializers.cpp:408-418); weights are exact float fractions. */
static inline float trilinear(const Grid& g, f3 n)
{
    int32_t x0, x1, y0, y1, z0, z1;
    float a, b, c;
    texel_axis(n.x, g.nx, x0, x1, a);
    texel_axis(n.y, g.ny, y0, y1, b);
    texel_axis(n.z, g.nz, z0, z1, c);
    float oa = 1.0f - a, ob = 1.0f - b, oc = 1.0f - c;
    float c00 = g.at(x0, y0, z0) * oa + g.at(x1, y0, z0) * a;
    float c10 = g.at(x0, y1, z0) * oa + g.at(x1, y1, z0) * a;
    float c01 = g.at(x0, y0, z1) * oa + g.at(x1, y0, z1) * a;
    float c11 = g.at(x0, y1, z1) * oa + g.at(x1, y1, z1) * a;
    float c0 = c00 * ob + c10 * b;
    float c1 = c01 * ob + c11 * b;
    return c0 * oc + c1 * c;
}

/* sampleCloud — PathTracer.comp:115-119 */
static inline f3 cloud_norm(const cr_cloud_properties& cp, f3 p)
{
    return mk((p.x - cp.bounds[0][0]) / (cp.bounds[1][0] - cp.bounds[0][0]),
              (p.y - cp.bounds[0][1]) / (cp.bounds[1][1] - cp.bounds[0][1]),
              (p.z - cp.bounds[0][2]) / (cp.bounds[1][2] - cp.bounds[0][2]));
}

/* sampleShadowVolume's coordinate map — PathTracer.comp:121-127 */
static inline f3 sv_norm(const cr_shadow_volume_properties& s, f3 p)
{
    const float* M = s.basisChange;
    float v0 = p.x - s.bounds[0][0], v1 = p.y - s.bounds[0][1], v2 = p.z - s.bounds[0][2];
    float v3 = 1.0f - s.bounds[0][3];
    float e0 = s.bounds[1][0] - s.bounds[0][0], e1 = s.bounds[1][1] - s.bounds[0][1];
    float e2 = s.bounds[1][2] - s.bounds[0][2], e3 = s.bounds[1][3] - s.bounds[0][3];
    float r[3];
    for (int i = 0; i < 3; i++) {
        float num = ((M[0 * 4 + i] * v0 + M[1 * 4 + i] * v1) + M[2 * 4 + i] * v2) + M[3 * 4 + i] * v3;
        float den = ((M[0 * 4 + i] * e0 + M[1 * 4 + i] * e1) + M[2 * 4 + i] * e2) + M[3 * 4 + i] * e3;
        r[i] = num / den;
    }
    return mk(r[0], r[1], r[2]);
}

/* intersectCloud — PathTracer.comp:154-184 (scratchapixel slab test). */
/* intersectCloud with the reciprocal direction given (invdir = 1 / dir). */
static inline bool intersect_box_inv(const float* b0, const float* b1, f3 pos, f3 inv, float& tmin, float& tmax)
{
    const float* B[2] = {b0, b1};
    int sx = inv.x < 0.0f, sy = inv.y < 0.0f, sz = inv.z < 0.0f;
    tmin = (B[sx][0] - pos.x) * inv.x;
    tmax = (B[1 - sx][0] - pos.x) * inv.x;
    float tymin = (B[sy][1] - pos.y) * inv.y;
    float tymax = (B[1 - sy][1] - pos.y) * inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (B[sz][2] - pos.z) * inv.z;
    float tzmax = (B[1 - sz][2] - pos.z) * inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return true;
}

static inline f3 recip(f3 d) { return mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z); }

static inline bool intersect_box(const float* b0, const float* b1, f3 pos, f3 dir, float& tmin, float& tmax)
{
    return intersect_box_inv(b0, b1, pos, recip(dir), tmin, tmax);
}

/* isInCloud — ShadowVolume.comp:34-42 */
static inline bool in_box(const float* b0, const float* b1, f3 p)
{
    return !(p.x < b0[0] || p.y < b0[1] || p.z < b0[2] || p.x > b1[0] || p.y > b1[1] || p.z > b1[2]);
}

static const float PI_F = 3.14159265359f;

static inline float smoothstep_f(float e0, float e1, float x)
{
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}

} // namespace orc

#endif
