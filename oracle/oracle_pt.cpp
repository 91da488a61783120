/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of the reference path tracer (shaders/PathTracer.comp),
 * the shadow-volume precompute (shaders/ShadowVolume.comp) and the build's
 * extensions (albedo, bounce cap, constant sky, brick-majorant DDA tracking,
 * ratio-tracked shadow rays).  With cr_render_ext_default() the per-pixel
 * computation is exactly PathTracer.comp main (:366-425).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

using namespace orc;

extern "C" {

float or_logf(float x) { return o_logf(x); }
float or_expf(float x) { return o_expf(x); }
void or_sincosf(float x, float* s, float* c) { o_sincosf(x, s, c); }
void or_math_array(int op, const float* in, uint32_t n, float* out)
{
    for (uint32_t i = 0; i < n; i++) {
        const float x = in[i];
        float sn, cs;
        switch (op) {
        case CR_MATH_SQRT: out[i] = sqrtf(x); break;
        case CR_MATH_LOG: out[i] = o_logf(x); break;
        case CR_MATH_EXP: out[i] = o_expf(x); break;
        case CR_MATH_ACOS: out[i] = o_acosf(x); break;
        case CR_MATH_ATAN2: out[i] = o_atan2f(x, in[(i ^ 1u) < n ? (i ^ 1u) : i]); break;
        default: o_sincosf(x, &sn, &cs); out[i] = op == CR_MATH_SIN ? sn : cs; break;
        }
    }
}

uint32_t or_xorshift32(uint32_t s) { return xs_next(s); }

void or_rng_stream(int32_t seed, uint32_t id, uint32_t count, uint32_t* out)
{
    uint32_t s;
    xs_init(s, (uint32_t)seed * id); /* PathTracer.comp:368: int seed * uint id, mod 2^32 */
    for (uint32_t i = 0; i < count; i++) out[i] = xs_next(s);
}

/* MSVC CRT rand(): state = state * 214013 + 2531011, return (state >> 16) & 0x7fff. */
void or_msvc_rand(uint32_t srand_seed, uint32_t count, int32_t* out)
{
    uint32_t st = srand_seed;
    for (uint32_t i = 0; i < count; i++) {
        st = st * 214013u + 2531011u;
        out[i] = (int32_t)((st >> 16) & 0x7fffu);
    }
}

float or_trilinear_f32(const float* grid, const uint32_t dims[3], const float uvw[3])
{
    Grid g = {grid, nullptr, (int32_t)dims[0], (int32_t)dims[1], (int32_t)dims[2]};
    return trilinear(g, mk(uvw[0], uvw[1], uvw[2]));
}

} // extern "C"

namespace {

const float INV_4PI = 1.0f / (4.0f * PI_F);

/* BG_COLORS / BG_DISTS — PathTracer.comp:22-37 */
const float BG_COLORS[5][4] = {
    {0.00f, 0.0f, 0.02f, 1.0f},
    {0.01f, 0.05f, 0.2f, 1.0f},
    {0.7f, 0.9f, 1.0f, 1.0f},
    {0.1f, 0.3f, 1.0f, 1.0f},
    {0.01f, 0.1f, 0.7f, 1.0f},
};
const float BG_DISTS[5] = {-1.0f, -0.04f, 0.0f, 0.5f, 1.0f};

inline f4 background(const cr_render_ext& ext, f3 dir)
{
    f4 col;
    if (ext.skyMode == CR_SKY_CONSTANT) {
        col.x = ext.skyColor[0]; col.y = ext.skyColor[1]; col.z = ext.skyColor[2]; col.w = ext.skyColor[3];
        return col;
    }
    /* sampleBackground — PathTracer.comp:104-113; mix(a,b,t) = a*(1-t) + b*t */
    col.x = BG_COLORS[0][0]; col.y = BG_COLORS[0][1]; col.z = BG_COLORS[0][2]; col.w = BG_COLORS[0][3];
    for (int i = 1; i < 5; i++) {
        float t = smoothstep_f(BG_DISTS[i - 1], BG_DISTS[i], dir.y);
        float ot = 1.0f - t;
        col.x = col.x * ot + BG_COLORS[i][0] * t;
        col.y = col.y * ot + BG_COLORS[i][1] * t;
        col.z = col.z * ot + BG_COLORS[i][2] * t;
        col.w = col.w * ot + BG_COLORS[i][3] * t;
    }
    return col;
}

inline bool isotropic(const cr_parameters& p) { return (p.isotropic & 0xffu) != 0; }

/* calculatePDF / samplePhase — PathTracer.comp:247-274 */
inline float phase_pdf(const cr_parameters& p, f3 a, f3 b)
{
    if (isotropic(p)) return INV_4PI;
    float cosT = dot(a, b);
    return (INV_4PI * p.phaseOneMinusG2) / o_pow15(p.phaseOnePlusG2 - (2.0f * p.phaseG) * cosT);
}

/* invertCDF — PathTracer.comp:252-256 */
inline float invert_cdf(const cr_parameters& p, float xi)
{
    float sq = p.phaseOneMinusG2 / ((1.0f - p.phaseG) + (2.0f * p.phaseG) * xi);
    return p.phaseOneOver2G * (p.phaseOnePlusG2 - sq * sq);
}

/* scatterRay + createOrthonormalBasis — PathTracer.comp:197-206, :277-311 */
inline f3 scatter(const cr_parameters& p, f3 n, uint32_t& rng)
{
    if (isotropic(p)) {
        f3 d;
        float xi = xs_float(rng);
        d.z = xi * 2.0f - 1.0f;
        float st = 1.0f - d.z * d.z;
        if (st > 0.0f) {
            st = sqrtf(fmaxf(0.0f, st));
            xi = xs_float(rng);
            float phi = (xi * 2.0f) * PI_F;
            float sp, cp;
            o_sincosf(phi, &sp, &cp);
            d.x = st * cp;
            d.y = st * sp;
        } else {
            d.x = 0.0f;
            d.y = 0.0f;
        }
        return d;
    }
    float phi = (xs_float(rng) * 2.0f) * PI_F;
    float cosT = invert_cdf(p, xs_float(rng));
    float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
    float sz = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = n.y / (1.0f + fabsf(n.z));
    float b = n.y * a;
    float c = -n.x * a;
    f3 t0 = mk(n.z + sz * b, sz * c, -n.x);
    f3 t1 = mk(c, 1.0f - b, -sz * n.y);
    float sp, cp;
    o_sincosf(phi, &sp, &cp);
    float k0 = sinT * cp, k1 = sinT * sp;
    return add(add(mul(t0, k0), mul(t1, k1)), mul(n, cosT));
}

/* getCameraRay — PathTracer.comp:144-151 */
inline void camera_ray(const cr_camera_properties& c, int32_t x, int32_t y, f3& pos, f3& dir)
{
    float dx = (float)(x - c.halfWidth);
    float dy = (float)(y - c.halfHeight);
    f3 d;
    d.x = (c.forward[0] * c.nearPlane + (c.right[0] * c.pixelSizeX) * dx) - (c.up[0] * c.pixelSizeY) * dy;
    d.y = (c.forward[1] * c.nearPlane + (c.right[1] * c.pixelSizeX) * dx) - (c.up[1] * c.pixelSizeY) * dy;
    d.z = (c.forward[2] * c.nearPlane + (c.right[2] * c.pixelSizeX) * dx) - (c.up[2] * c.pixelSizeY) * dy;
    pos = add(d, ld3(c.position));
    dir = normalize(d);
}

struct Ctx {
    const or_scene* s;
    Grid g;
    Grid sv;
    const float* b0;
    const float* b1;
    float sigmaBar;   /* global majorant extinction: maxExtinction*densityScaling/baseScaling */
    float extScale;   /* densityScaling / baseScaling is never formed: see findScatterPoint */
    int32_t mx, my, mz;
    float bw[3];      /* world size of a majorant brick per axis */
    float ibw[3];     /* 1 / bw */
    or_counters* cnt;
};

inline float sample_cloud(const Ctx& C, f3 p)
{
    return trilinear(C.g, cloud_norm(C.s->cloud, p));
}

/* Brick-majorant DDA state along a ray from `pos` (extension). */
struct Dda {
    int32_t cell[3];
    int32_t step[3];
    float tNext[3];
    float tDelta[3];
};

/* The walk's set-up multiplies by the reciprocals the box test already
 * formed (inv = 1 / dir) and by 1 / bw, instead of dividing. */
inline void dda_init(const Ctx& C, f3 pos, f3 dir, f3 invd, Dda& d)
{
    const float p[3] = {pos.x, pos.y, pos.z};
    const float v[3] = {dir.x, dir.y, dir.z};
    const float iv[3] = {invd.x, invd.y, invd.z};
    const int32_t mc[3] = {C.mx, C.my, C.mz};
    for (int k = 0; k < 3; k++) {
        float rel = (p[k] - C.b0[k]) * C.ibw[k];
        int32_t c = (int32_t)floorf(fminf(fmaxf(rel, 0.0f), (float)mc[k]));
        c = clampi(c, 0, mc[k] - 1);
        d.cell[k] = c;
        if (v[k] > 0.0f) {
            d.step[k] = 1;
            d.tNext[k] = fmaxf(((C.b0[k] + (float)(c + 1) * C.bw[k]) - p[k]) * iv[k], 0.0f);
            d.tDelta[k] = C.bw[k] * iv[k];
        } else if (v[k] < 0.0f) {
            d.step[k] = -1;
            d.tNext[k] = fmaxf(((C.b0[k] + (float)c * C.bw[k]) - p[k]) * iv[k], 0.0f);
            d.tDelta[k] = C.bw[k] * -iv[k];
        } else {
            d.step[k] = 0;
            d.tNext[k] = INFINITY;
            d.tDelta[k] = INFINITY;
        }
    }
}

inline int dda_axis(const Dda& d)
{
    if (d.tNext[0] <= d.tNext[1] && d.tNext[0] <= d.tNext[2]) return 0;
    if (d.tNext[1] <= d.tNext[2]) return 1;
    return 2;
}

/* returns false when the ray leaves the brick grid */
inline bool dda_advance(const Ctx& C, Dda& d)
{
    int k = dda_axis(d);
    d.cell[k] += d.step[k];
    d.tNext[k] = d.tNext[k] + d.tDelta[k];
    const int32_t mc[3] = {C.mx, C.my, C.mz};
    return d.cell[k] >= 0 && d.cell[k] < mc[k];
}

inline float brick_majorant(const Ctx& C, const Dda& d)
{
    return C.s->majorants[(uint64_t)d.cell[0] + (uint64_t)C.mx * ((uint64_t)d.cell[1] + (uint64_t)C.my * (uint64_t)d.cell[2])];
}

/* findScatterPoint — PathTracer.comp:318-361 (global majorant, reference) or
 * its brick-majorant DDA form (extension).  Returns true at a real collision
 * and moves pos there. */
inline bool find_scatter(const Ctx& C, f3& pos, f3 dir, uint32_t& rng)
{
    float tmax = 0.0f, tmin = 0.0f;
    const f3 invd = recip(dir);
    if (!intersect_box_inv(C.b0, C.b1, pos, invd, tmin, tmax)) return false;
    f3 exitp = add(pos, mul(dir, tmax));
    float dist = length(sub(pos, exitp));
    const cr_cloud_properties& cp = C.s->cloud;
    if (C.s->ext.trackingMode == CR_TRACK_BRICK_DDA) {
        Dda d;
        dda_init(C, pos, dir, invd, d);
        float t = 0.0f;
        float tau = -o_logf(xs_float(rng));
        for (;;) {
            float maj = brick_majorant(C, d);
            float sb = (maj * cp.densityScaling) / cp.baseScaling;
            float tc = fminf(fminf(fminf(d.tNext[0], d.tNext[1]), d.tNext[2]), dist);
            float seg = tc - t;
            C.cnt->brickVisits++;
            if (sb > 0.0f && tau < sb * seg) {
                t = t + tau / sb;
                f3 cur = add(pos, mul(dir, t));
                float e = sample_cloud(C, cur);
                C.cnt->lookups++;
                float xi = xs_float(rng);
                if (xi < e / maj) {
                    pos = cur;
                    return true;
                }
                tau = -o_logf(xs_float(rng));
                continue;
            }
            tau = tau - sb * seg;
            t = tc;
            if (!(tc < dist)) return false;
            if (!dda_advance(C, d)) return false;
        }
    }
    float t = 0.0f;
    f3 cur = mk(0.0f, 0.0f, 0.0f);
    for (;;) {
        float zeta = xs_float(rng);
        t = t + (-o_logf(zeta)) / C.sigmaBar;
        if (!(t < dist)) return false;
        cur = add(pos, mul(dir, t));
        float e = sample_cloud(C, cur);
        C.cnt->lookups++;
        float xi = xs_float(rng);
        if (xi < e / cp.maxExtinction) break;
    }
    pos = cur;
    return true;
}

/* Ratio-tracked transmittance toward the sun (extension): prod(1 - sigma/sigmaBar). */
inline float ratio_track(const Ctx& C, f3 pos, f3 dir, uint32_t& rng)
{
    float tmax = 0.0f, tmin = 0.0f;
    const f3 invd = recip(dir);
    if (!intersect_box_inv(C.b0, C.b1, pos, invd, tmin, tmax)) return 1.0f;
    f3 exitp = add(pos, mul(dir, tmax));
    float dist = length(sub(pos, exitp));
    const cr_cloud_properties& cp = C.s->cloud;
    float T = 1.0f;
    if (C.s->ext.trackingMode == CR_TRACK_BRICK_DDA) {
        Dda d;
        dda_init(C, pos, dir, invd, d);
        float t = 0.0f;
        float tau = -o_logf(xs_float(rng));
        for (;;) {
            float maj = brick_majorant(C, d);
            float sb = (maj * cp.densityScaling) / cp.baseScaling;
            float tc = fminf(fminf(fminf(d.tNext[0], d.tNext[1]), d.tNext[2]), dist);
            float seg = tc - t;
            C.cnt->brickVisits++;
            if (sb > 0.0f && tau < sb * seg) {
                t = t + tau / sb;
                f3 cur = add(pos, mul(dir, t));
                float e = sample_cloud(C, cur);
                C.cnt->lookups++;
                T = T * (1.0f - e / maj);
                if (T == 0.0f) return 0.0f;
                tau = -o_logf(xs_float(rng));
                continue;
            }
            tau = tau - sb * seg;
            t = tc;
            if (!(tc < dist)) return T;
            if (!dda_advance(C, d)) return T;
        }
    }
    float t = 0.0f;
    for (;;) {
        float zeta = xs_float(rng);
        t = t + (-o_logf(zeta)) / C.sigmaBar;
        if (!(t < dist)) return T;
        f3 cur = add(pos, mul(dir, t));
        float e = sample_cloud(C, cur);
        C.cnt->lookups++;
        T = T * (1.0f - e / cp.maxExtinction);
        if (T == 0.0f) return 0.0f;
    }
}

/* A hard ceiling on real collisions per path, identical in the kernels, that
 * only guards against a non-terminating walk; the reference has none. */
const uint32_t COLLISION_CEILING = 1u << 20;

void init_ctx(Ctx& C, const or_scene* s, or_counters* cnt)
{
    C.s = s;
    C.g.f32 = s->grid_f32;
    C.g.f16 = s->grid_f16;
    C.g.nx = (int32_t)s->dims[0];
    C.g.ny = (int32_t)s->dims[1];
    C.g.nz = (int32_t)s->dims[2];
    C.sv.f32 = s->sv;
    C.sv.f16 = nullptr;
    C.sv.nx = (int32_t)s->svdims[0];
    C.sv.ny = (int32_t)s->svdims[1];
    C.sv.nz = (int32_t)s->svdims[2];
    C.b0 = s->cloud.bounds[0];
    C.b1 = s->cloud.bounds[1];
    C.sigmaBar = (s->cloud.maxExtinction * s->cloud.densityScaling) / s->cloud.baseScaling;
    C.extScale = 0.0f;
    C.mx = (int32_t)s->mdims[0];
    C.my = (int32_t)s->mdims[1];
    C.mz = (int32_t)s->mdims[2];
    for (int k = 0; k < 3; k++)
    {
        C.bw[k] = ((s->cloud.bounds[1][k] - s->cloud.bounds[0][k]) * (float)(s->ext.majorantBrick ? s->ext.majorantBrick : 16u)) /
                  (float)s->dims[k];
        C.ibw[k] = 1.0f / C.bw[k];
    }
    C.cnt = cnt;
}

/* PathTracer.comp main (:366-425) minus the accumulation, for one frame seed. */
f4 pt_sample(const Ctx& C, int32_t x, int32_t y, int32_t seed)
{
    const or_scene& S = *C.s;
    const cr_camera_properties& cam = S.cam;
    uint32_t W = (uint32_t)(2 * cam.halfWidth);
    uint32_t padded = 32u * (W / 32u + 1u);            /* gl_WorkGroupSize.x * gl_NumWorkGroups.x */
    uint32_t id = (uint32_t)x + (uint32_t)y * padded;  /* PathTracer.comp:368 */
    uint32_t rng;
    xs_init(rng, (uint32_t)seed * id);

    f3 pos, dir;
    camera_ray(cam, x, y, pos, dir);
    C.cnt->paths++;
    C.cnt->vertices++;
    f4 res = {0.0f, 0.0f, 0.0f, 0.0f};
    float tmax = 0.0f, tmin = 0.0f;
    if (!intersect_box(C.b0, C.b1, pos, dir, tmin, tmax) || tmax < 0.0f || S.cloud.densityScaling <= 0.0f) {
        return background(S.ext, dir);
    }
    pos = add(pos, mul(dir, tmax < tmin ? tmax : tmin));
    f3 ldir = ld3(S.svp.lightDirection);
    f3 sdir = neg(ldir);
    float thr = 1.0f;
    const float albedo = S.ext.albedo;
    const float I = S.params.lightIntensity;
    uint32_t bounces = 0;
    for (;;) {
        if (!find_scatter(C, pos, dir, rng)) {
            f4 bg = background(S.ext, dir);
            res.x = res.x + bg.x * thr;
            res.y = res.y + bg.y * thr;
            res.z = res.z + bg.z * thr;
            res.w = res.w + bg.w * thr;
            break;
        }
        bounces++;
        if (bounces > S.ext.bounceCap || bounces > COLLISION_CEILING) break;
        C.cnt->vertices++;
        float pdf = phase_pdf(S.params, neg(dir), ldir);
        float T = 1.0f;
        if (S.ext.neeMode == CR_NEE_SHADOW_VOLUME) {
            T = trilinear(C.sv, sv_norm(S.svp, pos));
            C.cnt->svLookups++;
        } else if (S.ext.neeMode == CR_NEE_RATIO_TRACKING) {
            T = ratio_track(C, pos, sdir, rng);
        } else {
            T = 0.0f;
        }
        float w = thr * albedo;
        /* SUNLIGHT_COLOR * lightIntensity * pdf * T, SUNLIGHT_COLOR = 1 */
        float contrib = (((1.0f * I) * pdf) * T) * w;
        res.x = res.x + contrib;
        res.y = res.y + contrib;
        res.z = res.z + contrib;
        res.w = res.w + contrib;
        dir = scatter(S.params, dir, rng);
        thr = w;
    }
    return res;
}

} // namespace

extern "C" {

void or_background(const cr_render_ext* ext, const float dir[3], float out[4])
{
    f4 c = background(*ext, ld3(dir));
    out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = c.w;
}

float or_phase_pdf(const cr_parameters* p, const float a[3], const float b[3])
{
    return phase_pdf(*p, ld3(a), ld3(b));
}

void or_scatter_dir(const cr_parameters* p, const float in_dir[3], uint32_t rng_state, float out_dir[3],
                    uint32_t* rng_state_out)
{
    uint32_t r = rng_state;
    f3 d = scatter(*p, ld3(in_dir), r);
    out_dir[0] = d.x; out_dir[1] = d.y; out_dir[2] = d.z;
    if (rng_state_out) *rng_state_out = r;
}

void or_camera_ray(const cr_camera_properties* cam, int32_t x, int32_t y, float pos[3], float dir[3])
{
    f3 p, d;
    camera_ray(*cam, x, y, p, d);
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

int or_intersect_box(const float b0[3], const float b1[3], const float pos[3], const float dir[3], float* tmin,
                     float* tmax)
{
    return intersect_box(b0, b1, ld3(pos), ld3(dir), *tmin, *tmax) ? 1 : 0;
}

void or_sv_dims(const or_scene* s, uint32_t out[3])
{
    if (s->ext.svDims[0] && s->ext.svDims[1] && s->ext.svDims[2]) {
        out[0] = s->ext.svDims[0]; out[1] = s->ext.svDims[1]; out[2] = s->ext.svDims[2];
    } else {
        out[0] = out[1] = out[2] = s->svp.voxelAxisCount;
    }
}

/* ShadowVolume.comp main (:77-93).  With per-axis texel counts (extension)
 * the lattice spacing along axis k is voxelSize * (N-1)/(N_k-1), N =
 * voxelAxisCount; with the cubic default this is the reference exactly. */
void or_sv_build(const or_scene* s, float* out)
{
    uint32_t d[3];
    or_sv_dims(s, d);
    const cr_shadow_volume_properties& P = s->svp;
    const cr_cloud_properties& cp = s->cloud;
    Grid g = {s->grid_f32, s->grid_f16, (int32_t)s->dims[0], (int32_t)s->dims[1], (int32_t)s->dims[2]};
    float k[3];
    for (int a = 0; a < 3; a++)
        k[a] = (d[a] == P.voxelAxisCount) ? 1.0f : (float)(P.voxelAxisCount - 1u) / (float)(d[a] - 1u);
    float hz = P.voxelSize * k[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t y = 0; y < (int64_t)d[1]; y++) {
        for (uint32_t x = 0; x < d[0]; x++) {
            float acc = 0.0f;
            float accDist = 0.0f;
            float fx = (float)x * k[0], fy = (float)y * k[1];
            for (uint32_t z = 0; z < d[2]; z++) {
                float fz = (float)z * k[2];
                f3 p;
                p.x = P.bounds[0][0] + ((fx * P.right[0] + fy * P.up[0]) + fz * P.lightDirection[0]) * P.voxelSize;
                p.y = P.bounds[0][1] + ((fx * P.right[1] + fy * P.up[1]) + fz * P.lightDirection[1]) * P.voxelSize;
                p.z = P.bounds[0][2] + ((fx * P.right[2] + fy * P.up[2]) + fz * P.lightDirection[2]) * P.voxelSize;
                if (in_box(cp.bounds[0], cp.bounds[1], p)) {
                    acc = acc + (trilinear(g, cloud_norm(cp, p)) * cp.densityScaling) / cp.baseScaling;
                }
                accDist = accDist + hz;
                out[(uint64_t)x + (uint64_t)d[0] * ((uint64_t)y + (uint64_t)d[1] * z)] =
                    o_expf(((-accDist) * acc) / (float)(z + 1u));
            }
        }
    }
}

void or_pt_sample(const or_scene* s, int32_t x, int32_t y, int32_t seed, float out[4], or_counters* c)
{
    or_counters local = {0, 0, 0, 0, 0};
    Ctx C;
    init_ctx(C, s, c ? c : &local);
    f4 r = pt_sample(C, x, y, seed);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

/* The pixels [x0, x0+w) x [y0, y0+h) of the full image (same ids, same
 * streams); accum is w x h RGBA.  Used to time a bounded sample of a large
 * render on the host. */
void or_pt_render_region(const or_scene* s, const cr_push_constants* frames, uint32_t n_frames, float* accum,
                         int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t threads, or_counters* c)
{
    uint64_t tp = 0, tv = 0, tl = 0, ts = 0, tb = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tp, tv, tl, ts, tb)
    for (int32_t j = 0; j < h * w; j++) {
        or_counters cnt = {0, 0, 0, 0, 0};
        Ctx C;
        init_ctx(C, s, &cnt);
        int32_t x = x0 + j % w, y = y0 + j / w;
        float* a = accum + 4 * (uint64_t)j;
        for (uint32_t f = 0; f < n_frames; f++) {
            f4 L = pt_sample(C, x, y, frames[f].seed);
            float n = (float)frames[f].frameCount;
            float d = (float)(frames[f].frameCount + 1u);
            a[0] = (L.x + a[0] * n) / d;
            a[1] = (L.y + a[1] * n) / d;
            a[2] = (L.z + a[2] * n) / d;
            a[3] = (L.w + a[3] * n) / d;
        }
        tp += cnt.paths; tv += cnt.vertices; tl += cnt.lookups; ts += cnt.svLookups; tb += cnt.brickVisits;
    }
    if (c) {
        c->paths += tp; c->vertices += tv; c->lookups += tl; c->svLookups += ts; c->brickVisits += tb;
    }
}

void or_pt_render(const or_scene* s, const cr_push_constants* frames, uint32_t n_frames, float* accum,
                  int32_t threads, or_counters* c)
{
    int32_t W = 2 * s->cam.halfWidth, H = 2 * s->cam.halfHeight;
    uint64_t tp = 0, tv = 0, tl = 0, ts = 0, tb = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : tp, tv, tl, ts, tb)
    for (int32_t y = 0; y < H; y++) {
        or_counters cnt = {0, 0, 0, 0, 0};
        Ctx C;
        init_ctx(C, s, &cnt);
        for (int32_t x = 0; x < W; x++) {
            float* a = accum + 4 * ((uint64_t)y * W + x);
            for (uint32_t f = 0; f < n_frames; f++) {
                f4 L = pt_sample(C, x, y, frames[f].seed);
                /* PathTracer.comp:420-424 */
                float n = (float)frames[f].frameCount;
                float d = (float)(frames[f].frameCount + 1u);
                a[0] = (L.x + a[0] * n) / d;
                a[1] = (L.y + a[1] * n) / d;
                a[2] = (L.z + a[2] * n) / d;
                a[3] = (L.w + a[3] * n) / d;
            }
        }
        tp += cnt.paths; tv += cnt.vertices; tl += cnt.lookups; ts += cnt.svLookups; tb += cnt.brickVisits;
    }
    if (c) {
        c->paths += tp; c->vertices += tv; c->lookups += tl; c->svLookups += ts; c->brickVisits += tb;
    }
}

} // extern "C"
