/*
 * ORACLE — test infrastructure only.
 *
 * Photon mapping (RenderTechniquePPM.cpp:229-312) restated on the CPU:
 * photon tracing (shaders/PPM_PT.comp:304-322, :398-488) into a uniform grid
 * and the ray-marched estimate (shaders/PPM_PE.comp:304-375, :443-545), with
 * the corrections of the build (deterministic deposits: emission order, a
 * cell keeps its first 32; the azimuth stored without doubling; sun and sky
 * photons started over the projected area, sky directions uniform; the
 * estimate reads only a cell's kept photons and skips the six density taps
 * whose mean the reference never uses).  The same IEEE binary32 sequence as
 * the gfx950 kernels (kern_ppm.hip), so the two agree bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

using namespace orc;

extern "C" {
typedef struct or_photon_map {
    const uint32_t* arrivals;   /* per cell */
    const uint32_t* start;      /* per cell: first of its run in photons */
    const cr_photon* photons;   /* sorted by cell, emission order within a cell */
} or_photon_map;
}

namespace {

const float INV_4PI_M = 1.0f / (4.0f * PI_F);
const float PI4_3 = (PI_F * 4.0f) / 3.0f;
const float TWO_PI = 6.28318530718f;
const uint32_t TRACK_CEILING = 1u << 16;
const uint32_t MARCH_CEILING = 1u << 16;
const uint32_t PER_CELL = CR_PPM_PHOTONS_PER_CELL;

inline Grid grid_of(const or_scene& s)
{
    Grid g = {s.grid_f32, s.grid_f16, (int32_t)s.dims[0], (int32_t)s.dims[1], (int32_t)s.dims[2]};
    return g;
}
inline float density(const or_scene& s, const Grid& g, f3 p) { return trilinear(g, cloud_norm(s.cloud, p)); }
inline f4 sky(const cr_render_ext& e, f3 d)
{
    float dd[3] = {d.x, d.y, d.z}, o[4];
    or_background(&e, dd, o);
    f4 r = {o[0], o[1], o[2], o[3]};
    return r;
}
inline float phase(const cr_parameters& p, f3 a, f3 b)
{
    float aa[3] = {a.x, a.y, a.z}, bb[3] = {b.x, b.y, b.z};
    return or_phase_pdf(&p, aa, bb);
}
/* scatterRay without the final normalize (PPM_PE.comp's form). */
inline f3 scatter_raw(const cr_parameters& p, f3 n, uint32_t& rng)
{
    float in[3] = {n.x, n.y, n.z}, out[3];
    uint32_t st = 0;
    or_scatter_dir(&p, in, rng, out, &st);
    rng = st;
    return mk(out[0], out[1], out[2]);
}

/* Entry point over the projected area of the box for direction d; returns the area. */
inline float entry_point(const float* b0, const float* b1, f3 d, uint32_t& rng, f3& pos)
{
    const float dx = b1[0] - b0[0], dy = b1[1] - b0[1], dz = b1[2] - b0[2];
    const float ax = (dz * dy) * fabsf(d.x), ay = (dx * dz) * fabsf(d.y), az = (dy * dx) * fabsf(d.z);
    const float total = (ax + ay) + az;
    const float sel = xs_float(rng) * total;
    const f3 c1 = mk(d.x > 0.0f ? b0[0] : b1[0], d.y > 0.0f ? b0[1] : b1[1], d.z > 0.0f ? b0[2] : b1[2]);
    const float u = xs_float(rng), v = xs_float(rng);
    if (sel < ax) pos = mk(c1.x, b0[1] * (1.0f - u) + b1[1] * u, b0[2] * (1.0f - v) + b1[2] * v);
    else if (sel < ax + ay) pos = mk(b0[0] * (1.0f - u) + b1[0] * u, c1.y, b0[2] * (1.0f - v) + b1[2] * v);
    else pos = mk(b0[0] * (1.0f - u) + b1[0] * u, b0[1] * (1.0f - v) + b1[1] * v, c1.z);
    return total;
}

/* depositPhoton's cell (PPM_PT.comp:306-309), clamped to the grid. */
inline uint32_t cell_of(const cr_photon_map_properties& pm, f3 p)
{
    const int32_t ix = clampi((int32_t)((p.x - pm.bounds[0][0]) / pm.voxelSize), 0, pm.voxelCount[0] - 1);
    const int32_t iy = clampi((int32_t)((p.y - pm.bounds[0][1]) / pm.voxelSize), 0, pm.voxelCount[1] - 1);
    const int32_t iz = clampi((int32_t)((p.z - pm.bounds[0][2]) / pm.voxelSize), 0, pm.voxelCount[2] - 1);
    return (uint32_t)ix + ((uint32_t)iy + (uint32_t)iz * (uint32_t)pm.voxelCount[1]) * (uint32_t)pm.voxelCount[0];
}

/* One photon path (PPM_PT.comp main :452-488); with out, deposits are written there. */
uint32_t trace_path(const or_scene& s, const cr_photon_map_properties& pm, const cr_ppm_params& prm, int32_t seed,
                    uint32_t id, cr_photon* out, uint32_t* cells)
{
    uint32_t n = 0;
    const cr_cloud_properties& cp = s.cloud;
    if (!(cp.densityScaling > 0.0f)) return 0;
    const Grid g = grid_of(s);
    const float* b0 = cp.bounds[0];
    const float* b1 = cp.bounds[1];
    uint32_t rng;
    xs_init(rng, (uint32_t)seed * id);
    const float emitted = (float)prm.numPaths;
    f3 pos, dir;
    f4 col;
    if (xs_float(rng) < prm.sunProbability) {
        dir = ld3(pm.lightDirection);
        const float pdf = 1.0f / entry_point(b0, b1, dir, rng, pos);
        const float c = (((1.0f * s.params.lightIntensity) / (pdf * emitted)) / prm.sunProbability) * 0.5f;
        col.x = col.y = col.z = col.w = c;
    } else {
        const float z = 1.0f - 2.0f * xs_float(rng);
        const float ph = TWO_PI * xs_float(rng);
        const float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cph;
        o_sincosf(ph, &sp, &cph);
        dir = mk(st * cph, st * sp, z);
        const float pdf = INV_4PI_M / entry_point(b0, b1, dir, rng, pos);
        const f4 bg = sky(s.ext, neg(dir));
        const float k = pdf * emitted, q = 1.0f - prm.sunProbability;
        col.x = (((bg.x / k) * 0.5f) / q) * 0.5f;
        col.y = (((bg.y / k) * 0.5f) / q) * 0.5f;
        col.z = (((bg.z / k) * 0.5f) / q) * 0.5f;
        col.w = (((bg.w / k) * 0.5f) / q) * 0.5f;
    }
    const float sigmaBar = (cp.maxExtinction * cp.densityScaling) / cp.baseScaling;
    bool first = true;
    for (uint32_t c = 0; c < prm.maxCollisions; c++) {
        /* interactWithMedium — PPM_PT.comp:398-446 */
        float tmin = 0.0f, tmax = 0.0f;
        if (!intersect_box(b0, b1, pos, dir, tmin, tmax)) break;
        const float dist = length(sub(pos, add(pos, mul(dir, tmax))));
        float t = 0.0f;
        f3 cur = pos;
        bool hit = false;
        for (uint32_t step = 0; step < TRACK_CEILING; step++) {
            t = t + (-o_logf(xs_float(rng))) / sigmaBar;
            if (!(t <= dist)) break;
            cur = add(pos, mul(dir, t));
            const float e = density(s, g, cur);
            if (xs_float(rng) < e / cp.maxExtinction) { hit = true; break; }
        }
        if (!hit) break;
        pos = cur;
        if (first) {
            first = false;
        } else {
            if (out) {
                cr_photon& p = out[n];
                p.position[0] = cur.x; p.position[1] = cur.y; p.position[2] = cur.z; p.position[3] = 0.0f;
                p.power[0] = col.x; p.power[1] = col.y; p.power[2] = col.z; p.power[3] = col.w;
                p.theta = o_acosf(fminf(fmaxf(dir.z, -1.0f), 1.0f));
                const float phi = o_atan2f(dir.y, dir.x);
                p.phi = phi < 0.0f ? phi + TWO_PI : phi;
                p._padding[0] = 0.0f;
                p._padding[1] = 0.0f;
                cells[n] = cell_of(pm, cur);
            }
            n++;
        }
        if (!(xs_float(rng) > pm.absorption)) break;
        dir = normalize(scatter_raw(s.params, dir, rng));
    }
    return n;
}

/* samplePhotonMap — PPM_PE.comp:304-375 */
f4 photon_query(const or_scene& s, const cr_photon_map_properties& pm, const or_photon_map& m, f3 p, f3 dir,
                float scatter, float r)
{
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (scatter == 0.0f) return acc;
    const f3 g = sub(p, ld3(pm.bounds[0]));
    const int32_t x0 = std::max((int32_t)((g.x - r) / pm.voxelSize), 0);
    const int32_t x1 = std::min((int32_t)((g.x + r) / pm.voxelSize), pm.voxelCount[0] - 1);
    const int32_t y0 = std::max((int32_t)((g.y - r) / pm.voxelSize), 0);
    const int32_t y1 = std::min((int32_t)((g.y + r) / pm.voxelSize), pm.voxelCount[1] - 1);
    const int32_t z0 = std::max((int32_t)((g.z - r) / pm.voxelSize), 0);
    const int32_t z1 = std::min((int32_t)((g.z + r) / pm.voxelSize), pm.voxelCount[2] - 1);
    const float r2 = r * r;
    for (int32_t z = z0; z <= z1; z++)
        for (int32_t y = y0; y <= y1; y++)
            for (int32_t x = x0; x <= x1; x++) {
                const uint32_t c = (uint32_t)x + ((uint32_t)y + (uint32_t)z * (uint32_t)pm.voxelCount[1]) *
                                                     (uint32_t)pm.voxelCount[0];
                const uint32_t cnt = std::min(m.arrivals[c], PER_CELL);
                for (uint32_t i = 0; i < cnt; i++) {
                    const cr_photon& ph = m.photons[m.start[c] + i];
                    const f3 d = sub(ld3(ph.position), p);
                    if (dot(d, d) <= r2) {
                        float st, ct, sp, cph;
                        o_sincosf(ph.theta, &st, &ct);
                        o_sincosf(ph.phi, &sp, &cph);
                        const float f = phase(s.params, dir, mk(st * cph, st * sp, ct));
                        acc.x = acc.x + f * ph.power[0];
                        acc.y = acc.y + f * ph.power[1];
                        acc.z = acc.z + f * ph.power[2];
                        acc.w = acc.w + f * ph.power[3];
                    }
                }
            }
    const float avg = (((0.0f + acc.x) + acc.y) + acc.z) + acc.w;
    if (avg / 4.0f <= 0.0f) {
        acc.x = acc.y = acc.z = acc.w = 0.0f;
        return acc;
    }
    const float v = ((PI4_3 * r) * r2) * scatter;
    acc.x = acc.x / v; acc.y = acc.y / v; acc.z = acc.z / v; acc.w = acc.w / v;
    return acc;
}

/* PPM_PE.comp main (:443-545) for one pixel and one frame. */
f4 estimate(const or_scene& s, const cr_photon_map_properties& pm, const or_photon_map& m, const cr_ppm_params& prm,
            int32_t x, int32_t y, int32_t seed, float radius)
{
    const cr_cloud_properties& cp = s.cloud;
    const uint32_t W = (uint32_t)(2 * s.cam.halfWidth);
    const uint32_t padded = 32u * (W / 32u + 1u);
    uint32_t rng;
    xs_init(rng, (uint32_t)seed * ((uint32_t)x + (uint32_t)y * padded));
    float pp[3], dd[3];
    or_camera_ray(&s.cam, x, y, pp, dd);
    f3 pos = mk(pp[0], pp[1], pp[2]), dir = mk(dd[0], dd[1], dd[2]);
    f4 res = sky(s.ext, dir);
    float tmin = 0.0f, tmax = 0.0f;
    if (!intersect_box(cp.bounds[0], cp.bounds[1], pos, dir, tmin, tmax) || !(cp.densityScaling > 0.0f)) return res;
    const Grid g = grid_of(s);
    const Grid sv = {s.sv, nullptr, (int32_t)s.svdims[0], (int32_t)s.svdims[1], (int32_t)s.svdims[2]};
    const f3 closest = add(pos, mul(dir, tmin));
    pos = add(pos, mul(dir, tmax - pm.stepSize * xs_float(rng)));
    const float maxDist = length(sub(pos, closest));
    const float step = pm.stepSize;
    const float albedo = 1.0f - pm.absorption;
    const f3 L = ld3(s.svp.lightDirection);
    const float dsbs = cp.densityScaling / cp.baseScaling;
    float cur = 0.0f;
    for (uint32_t k = 0; k < MARCH_CEILING && cur < maxDist; k++) {
        pos = sub(pos, mul(dir, step));
        cur = cur + step;
        const float ext = (density(s, g, pos) * cp.densityScaling) / cp.baseScaling;
        const float sc = albedo * ext;
        const float pdf = phase(s.params, neg(dir), L);
        const float tsun = trilinear(sv, sv_norm(s.svp, pos));
        const float dl = ((1.0f * s.params.lightIntensity) * pdf) * tsun;
        f4 direct = {dl, dl, dl, dl};
        const f3 ad = scatter_raw(s.params, dir, rng);
        f3 q = pos;
        float accD = 0.0f, accL = 0.0f;
        uint32_t samples = 0;
        f4 amb = sky(s.ext, ad);
        for (uint32_t j = 0; j < MARCH_CEILING && in_box(cp.bounds[0], cp.bounds[1], q); j++) {
            q = add(q, mul(ad, prm.ambientStep));
            accD = accD + density(s, g, q);
            accL = accL + prm.ambientStep;
            samples++;
        }
        accD = accD * dsbs;
        const float ta = o_expf((-accL * accD) / (float)samples);
        amb.x = amb.x * ta; amb.y = amb.y * ta; amb.z = amb.z * ta; amb.w = amb.w * ta;
        if (amb.x > 0.0f) {
            direct.x = direct.x + amb.x; direct.y = direct.y + amb.y;
            direct.z = direct.z + amb.z; direct.w = direct.w + amb.w;
        }
        const f4 i0 = photon_query(s, pm, m, pos, neg(dir), sc, radius);
        const f4 ind = {albedo * i0.x, albedo * i0.y, albedo * i0.z, albedo * i0.w};
        const float att = o_expf(-step * ext);
        res.x = res.x * att + (sc * (direct.x + ind.x)) * step;
        res.y = res.y * att + (sc * (direct.y + ind.y)) * step;
        res.z = res.z * att + (sc * (direct.z + ind.z)) * step;
        res.w = res.w * att + (sc * (direct.w + ind.w)) * step;
    }
    return res;
}

} // namespace

extern "C" {

void or_ppm_trace_count(const or_scene* s, const cr_photon_map_properties* pm, const cr_ppm_params* p, int32_t seed,
                        uint32_t* counts)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)p->numPaths; i++)
        counts[i] = trace_path(*s, *pm, *p, seed, (uint32_t)i, nullptr, nullptr);
}

void or_ppm_trace_write(const or_scene* s, const cr_photon_map_properties* pm, const cr_ppm_params* p, int32_t seed,
                        const uint64_t* offsets, cr_photon* photons, uint32_t* cells)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)p->numPaths; i++)
        trace_path(*s, *pm, *p, seed, (uint32_t)i, photons + offsets[i], cells + offsets[i]);
}

/* Stable counting sort of the deposits by cell: arrivals and the first slot
 * of every cell, and the photons in (cell, emission) order. */
void or_ppm_build(const cr_photon* photons, const uint32_t* cells, uint32_t n, uint32_t n_cells, uint32_t* arrivals,
                  uint32_t* start, cr_photon* sorted)
{
    memset(arrivals, 0, (size_t)n_cells * 4);
    for (uint32_t i = 0; i < n; i++) arrivals[cells[i]]++;
    uint32_t acc = 0;
    for (uint32_t c = 0; c < n_cells; c++) {
        start[c] = acc;
        acc += arrivals[c];
    }
    std::vector<uint32_t> fill(start, start + n_cells);
    for (uint32_t i = 0; i < n; i++) sorted[fill[cells[i]]++] = photons[i];
}

void or_ppm_query(const or_scene* s, const cr_photon_map_properties* pm, const or_photon_map* m, const float pos[3],
                  const float dir[3], float scatter, float radius, float out[4])
{
    const f4 r = photon_query(*s, *pm, *m, ld3(pos), ld3(dir), scatter, radius);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

/* n_frames frames of the estimate with cr_pt_render's running mean into accum
 * (W x H RGBA32F); frames[f].pmRadius is the query radius. */
void or_ppm_estimate_render(const or_scene* s, const cr_photon_map_properties* pm, const or_photon_map* m,
                            const cr_ppm_params* p, const cr_push_constants* frames, uint32_t n_frames, float* accum,
                            int32_t threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int32_t W = 2 * s->cam.halfWidth, H = 2 * s->cam.halfHeight;
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t y = 0; y < H; y++)
        for (int32_t x = 0; x < W; x++) {
            float* a = accum + 4 * ((uint64_t)y * W + x);
            for (uint32_t f = 0; f < n_frames; f++) {
                const f4 L = estimate(*s, *pm, *m, *p, x, y, frames[f].seed, frames[f].pmRadius);
                const float fc = (float)frames[f].frameCount;
                const float d = (float)(frames[f].frameCount + 1u);
                a[0] = (L.x + a[0] * fc) / d;
                a[1] = (L.y + a[1] * fc) / d;
                a[2] = (L.z + a[2] * fc) / d;
                a[3] = (L.w + a[3] * fc) / d;
            }
        }
}

} // extern "C"
