/*
 * ORACLE — test infrastructure only.
 *
 * Photon-beam producer (shaders/PPB_PT.comp) and gather (shaders/PPB_PE.comp)
 * restated on the CPU, with these corrections:
 *   - sun photons only, entry face chosen against the light direction and the
 *     face selected with if / else-if (the reference starts photons on
 *     the exit face and always overwrites the yz choice);
 *   - deterministic output order: a counting pass and a writing pass over the
 *     same RNG stream replace the atomic appends (PPB_PT.comp:340, :381);
 *   - the segment's offset along its parent beam is kept in _padding_beam[0]
 *     so the beam transmittance is looked up at the right distance;
 *   - the gather term follows Jarosz et al. 2011 (beam x beam, 1D kernel):
 *     K(d/r)/r * sigma_s * Phi * T_cam * T_beam * f(theta) / sin(theta),
 *     or, with refCompat, PPB_PE.comp:394-403 as written;
 *   - the traversal box test uses half the box size (the reference passes the
 *     full size to iq's intersectBox, PPB_PE.comp:424-431, doubling boxes).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

using namespace orc;

extern "C" {
typedef struct or_beam_scene {
    or_scene base;                   /* grid, cloud, params, camera, SV (along the camera) */
    const cr_photon_beam* beams;     /* sorted order: leaf i -> beams[i] */
    const cr_photon_beam_data* data;
    const cr_tree_node* nodes;
    uint32_t n;
    uint32_t refCompat;
    uint32_t jitter;
    float albedo;
} or_beam_scene;

uint32_t or_morton30(const float c[3], const float b0[3], const float b1[3]);
}

namespace {

const float INV_4PI = 1.0f / (4.0f * PI_F);
const uint32_t NONE = 0xFFFFFFFFu;
/* ceiling on delta-tracking steps of one walk along a beam (same in the kernel) */
const uint32_t TRACK_CEILING = 1u << 16;

inline float phase(const cr_parameters& p, f3 a, f3 b)
{
    if ((p.isotropic & 0xffu) != 0) return INV_4PI;
    float cosT = dot(a, b);
    return (INV_4PI * p.phaseOneMinusG2) / o_pow15(p.phaseOnePlusG2 - (2.0f * p.phaseG) * cosT);
}

inline float sample_cloud(const or_scene& s, const Grid& g, f3 p) { return trilinear(g, cloud_norm(s.cloud, p)); }

inline Grid grid_of(const or_scene& s)
{
    Grid g = {s.grid_f32, s.grid_f16, (int32_t)s.dims[0], (int32_t)s.dims[1], (int32_t)s.dims[2]};
    return g;
}

/* scatterRay of PPB_PT.comp:268-304 (non-isotropic branch plus normalize). */
inline f3 scatter_norm(const cr_parameters& p, f3 n, uint32_t& rng)
{
    f3 d;
    if ((p.isotropic & 0xffu) != 0) {
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
    } else {
        float phi = (xs_float(rng) * 2.0f) * PI_F;
        float xi = xs_float(rng);
        float sq = p.phaseOneMinusG2 / ((1.0f - p.phaseG) + (2.0f * p.phaseG) * xi);
        float cosT = p.phaseOneOver2G * (p.phaseOnePlusG2 - sq * sq);
        float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
        float sz = n.z >= 0.0f ? 1.0f : -1.0f;
        float a = n.y / (1.0f + fabsf(n.z));
        float b = n.y * a;
        float c = -n.x * a;
        f3 t0 = mk(n.z + sz * b, sz * c, -n.x);
        f3 t1 = mk(c, 1.0f - b, -sz * n.y);
        float sp, cp;
        o_sincosf(phi, &sp, &cp);
        d = add(add(mul(t0, sinT * cp), mul(t1, sinT * sp)), mul(n, cosT));
    }
    return normalize(d);
}

/* One delta-tracking walk along a beam: PPB_PT.comp:346-370. */
inline float track_free_path(const or_scene& s, const Grid& g, f3 pos, f3 dir, float len, float sigmaBar,
                             uint32_t& rng)
{
    float t = 0.0f;
    for (uint32_t step = 0; step < TRACK_CEILING; step++) {
        float zeta = xs_float(rng);
        t = t + (-o_logf(zeta)) / sigmaBar;
        if (!(t <= len)) break;
        f3 cur = add(pos, mul(dir, t));
        float e = sample_cloud(s, g, cur);
        float xi = xs_float(rng);
        if (xi < e / s.cloud.maxExtinction) break;
    }
    return t;
}

struct PathOut {
    uint32_t segs = 0, datas = 0;
};

/* One photon path (PPB_PT.comp main :506-539, sun only).  When `beams` is
 * non-null the records are written at seg_base / data_base. */
PathOut trace_path(const or_scene& s, const cr_beam_trace_params& P, uint32_t id, cr_photon_beam* beams,
                   cr_photon_beam_data* data, uint64_t seg_base, uint64_t data_base, uint64_t max_beams,
                   uint64_t max_data)
{
    PathOut o;
    if (s.cloud.densityScaling <= 0.0f) return o;
    Grid g = grid_of(s);
    const float* b0 = s.cloud.bounds[0];
    const float* b1 = s.cloud.bounds[1];
    uint32_t rng;
    xs_init(rng, (uint32_t)P.seed * id);
    f3 L = ld3(P.lightDirection);
    /* generateSunRay — PPB_PT.comp:396-430 (corrected) */
    float dx = b1[0] - b0[0], dy = b1[1] - b0[1], dz = b1[2] - b0[2];
    float ax = (dz * dy) * fabsf(L.x), ay = (dx * dz) * fabsf(L.y), az = (dy * dx) * fabsf(L.z);
    float total = (ax + ay) + az;
    float sel = xs_float(rng) * total;
    f3 c1 = mk(L.x > 0.0f ? b0[0] : b1[0], L.y > 0.0f ? b0[1] : b1[1], L.z > 0.0f ? b0[2] : b1[2]);
    f3 pos;
    if (sel < ax) {
        float u = xs_float(rng), v = xs_float(rng);
        pos = mk(c1.x, b0[1] * (1.0f - u) + b1[1] * u, b0[2] * (1.0f - v) + b1[2] * v);
    } else if (sel < ax + ay) {
        float u = xs_float(rng), v = xs_float(rng);
        pos = mk(b0[0] * (1.0f - u) + b1[0] * u, c1.y, b0[2] * (1.0f - v) + b1[2] * v);
    } else {
        float u = xs_float(rng), v = xs_float(rng);
        pos = mk(b0[0] * (1.0f - u) + b1[0] * u, b0[1] * (1.0f - v) + b1[1] * v, c1.z);
    }
    float pdf = 1.0f / total;
    float power = (1.0f * P.lightIntensity) / ((float)P.numPaths * pdf);
    f3 dir = L;
    const float r = P.radius;
    const float sigmaBar = (s.cloud.maxExtinction * s.cloud.densityScaling) / s.cloud.baseScaling;
    const float absorption = 1.0f - P.albedo;
    for (uint32_t it = 0; it < P.maxInteractions; it++) {
        /* interactWithMedium — PPB_PT.comp:453-501 */
        float tmin = 0.0f, tmax = 0.0f;
        if (!intersect_box(b0, b1, pos, dir, tmin, tmax)) break;
        f3 exitp = add(pos, mul(dir, tmax));
        /* depositPhotonBeam — PPB_PT.comp:334-393 */
        float len = length(sub(pos, exitp));
        uint64_t di = data_base + o.datas;
        o.datas++;
        float tr[16];
        for (int i = 0; i < 16; i++) tr[i] = track_free_path(s, g, pos, dir, len, sigmaBar, rng);
        if (beams && di < max_data) {
            cr_photon_beam_data& D = data[di];
            D.power[0] = D.power[1] = D.power[2] = D.power[3] = power;
            for (int i = 0; i < 16; i++) D.trDistances[i] = tr[i];
        }
        uint32_t segs = (uint32_t)(len / r) + 1u;
        float segSize = len / (float)segs;
        if (beams) {
            for (uint32_t i = 0; i < segs; i++) {
                uint64_t bi = seg_base + o.segs + i;
                if (bi >= max_beams) break;
                cr_photon_beam& B = beams[bi];
                float off = (float)i * segSize;
                f3 st = add(pos, mul(dir, off));
                f3 en = add(st, mul(dir, segSize));
                B.startPos[0] = st.x; B.startPos[1] = st.y; B.startPos[2] = st.z;
                B.endPos[0] = en.x; B.endPos[1] = en.y; B.endPos[2] = en.z;
                B.radius = r;
                float c[3] = {(st.x + en.x) / 2.0f, (st.y + en.y) / 2.0f, (st.z + en.z) / 2.0f};
                B.mortonCode = or_morton30(c, b0, b1);
                B.dataIdx = (uint32_t)di;
                B._padding_beam[0] = o_f2u(off);
                B._padding_beam[1] = 0;
                B._padding_beam[2] = 0;
            }
        }
        o.segs += segs;
        /* next real collision along the ray */
        float dist = len;
        float t = 0.0f;
        f3 cur = pos;
        bool left = true;
        for (uint32_t step = 0; step < TRACK_CEILING; step++) {
            float zeta = xs_float(rng);
            t = t + (-o_logf(zeta)) / sigmaBar;
            if (!(t <= dist)) { left = true; break; }
            cur = add(pos, mul(dir, t));
            float e = sample_cloud(s, g, cur);
            float xi = xs_float(rng);
            if (xi < e / s.cloud.maxExtinction) { left = false; break; }
        }
        if (left) break;
        pos = cur;
        float xi = xs_float(rng);
        if (!(xi > absorption)) break;
        dir = scatter_norm(s.params, dir, rng);
    }
    return o;
}

/* Deviation: a ray with a direction component exactly 0 misses a box whose slab
   on that axis does not hold it (by more than the rounding of the center /
   half-size form).  The slab formula yields NaN there (inf - inf), which fmaxf /
   fminf drop, so that axis would never cull. */
inline bool outside_slab(float ro, float d, float cen, float rad)
{
    return d == 0.0f && fabsf(ro - cen) > rad + 1e-6f * (fabsf(ro) + fabsf(cen) + rad);
}

/* intersectBox — PPB_PE.comp:282-300 with rad = half the box size, plus the
   parallel-axis test above. */
inline bool hit_box(f3 ro, f3 rd, const float* lo, const float* hi)
{
    f3 cen = mk((lo[0] + hi[0]) / 2.0f, (lo[1] + hi[1]) / 2.0f, (lo[2] + hi[2]) / 2.0f);
    f3 rad = mk((hi[0] - lo[0]) * 0.5f, (hi[1] - lo[1]) * 0.5f, (hi[2] - lo[2]) * 0.5f);
    f3 m = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    f3 n = mk(m.x * (ro.x - cen.x), m.y * (ro.y - cen.y), m.z * (ro.z - cen.z));
    f3 k = mk(fabsf(m.x) * rad.x, fabsf(m.y) * rad.y, fabsf(m.z) * rad.z);
    float tN = fmaxf(fmaxf(-n.x - k.x, -n.y - k.y), -n.z - k.z);
    float tF = fminf(fminf(-n.x + k.x, -n.y + k.y), -n.z + k.z);
    bool parallelMiss = outside_slab(ro.x, rd.x, cen.x, rad.x) || outside_slab(ro.y, rd.y, cen.y, rad.y) ||
                        outside_slab(ro.z, rd.z, cen.z, rad.z);
    return !(tN > tF || tF < 0.0f || parallelMiss);
}

/* accumulateRadiance — PPB_PE.comp:362-404 */
inline f4 beam_radiance(const or_beam_scene& S, const Grid& g, const Grid& sv, uint32_t bi, f3 ro, f3 rd)
{
    f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    const cr_photon_beam& B = S.beams[bi];
    f3 st = ld3(B.startPos), en = ld3(B.endPos);
    f3 bd = normalize(sub(en, st));
    if (fabsf(dot(rd, bd)) > 0.9f) return z;
    f3 tangent = cross(rd, bd);
    f3 pn = cross(bd, tangent);
    float t = dot(sub(st, ro), pn) / dot(pn, rd);
    if (!S.refCompat && !(t >= 0.0f)) return z;
    f3 x = add(ro, mul(rd, t));
    float blen = length(sub(en, st));
    float s = dot(sub(x, st), bd);
    if (s < 0.0f || s > blen) return z;
    f3 proj = add(st, mul(bd, s));
    float dist = length(sub(x, proj));
    if (dist > B.radius) return z;
    float q = dist / B.radius;
    float sq = 1.0f - q * q;
    float kern = ((15.0f / 16.0f) * sq) * sq;
    float dens = sample_cloud(S.base, g, x);
    float tcam = trilinear(sv, sv_norm(S.base.svp, x));
    const cr_photon_beam_data& D = S.data[B.dataIdx];
    float soff = S.refCompat ? s : (o_u2f(B._padding_beam[0]) + s);
    uint32_t cnt = 0;
    for (int i = 0; i < 16; i++)
        if (D.trDistances[i] >= soff) cnt++;
    float tb = (float)cnt / 16.0f;
    f3 nrd = neg(rd);
    float ph = phase(S.base.params, bd, nrd);
    float cosA = dot(bd, nrd);
    float sinA, sc;
    if (S.refCompat) {
        sinA = 1.0f - cosA * cosA;
        sc = dens;
    } else {
        kern = kern / B.radius;
        sinA = sqrtf(1.0f - cosA * cosA);
        sc = S.albedo * ((dens * S.base.cloud.densityScaling) / S.base.cloud.baseScaling);
    }
    float w = ph / sinA;
    float ks = kern * sc;
    f4 r;
    r.x = (((ks * D.power[0]) * tcam) * tb) * w;
    r.y = (((ks * D.power[1]) * tcam) * tb) * w;
    r.z = (((ks * D.power[2]) * tcam) * tb) * w;
    r.w = (((ks * D.power[3]) * tcam) * tb) * w;
    return r;
}

inline void acc4(f4& a, f4 b) { a.x = a.x + b.x; a.y = a.y + b.y; a.z = a.z + b.z; a.w = a.w + b.w; }

/* getCameraRay with an optional sub-pixel jitter drawn from the pixel stream. */
inline void gather_ray(const or_beam_scene& S, int32_t x, int32_t y, int32_t seed, f3& pos, f3& dir)
{
    const cr_camera_properties& c = S.base.cam;
    float dx = (float)(x - c.halfWidth);
    float dy = (float)(y - c.halfHeight);
    if (S.jitter) {
        uint32_t W = (uint32_t)(2 * c.halfWidth);
        uint32_t id = (uint32_t)x + (uint32_t)y * (32u * (W / 32u + 1u));
        uint32_t rng;
        xs_init(rng, (uint32_t)seed * id);
        dx = dx + (xs_float(rng) - 0.5f);
        dy = dy + (xs_float(rng) - 0.5f);
    }
    f3 d;
    d.x = (c.forward[0] * c.nearPlane + (c.right[0] * c.pixelSizeX) * dx) - (c.up[0] * c.pixelSizeY) * dy;
    d.y = (c.forward[1] * c.nearPlane + (c.right[1] * c.pixelSizeX) * dx) - (c.up[1] * c.pixelSizeY) * dy;
    d.z = (c.forward[2] * c.nearPlane + (c.right[2] * c.pixelSizeX) * dx) - (c.up[2] * c.pixelSizeY) * dy;
    pos = add(d, ld3(c.position));
    dir = normalize(d);
}

/* traverseTree — PPB_PE.comp:408-466 */
f4 traverse(const or_beam_scene& S, f3 ro, f3 rd, uint64_t* nodes_visited, uint64_t* beams_tested)
{
    f4 rad = {0.0f, 0.0f, 0.0f, 0.0f};
    if (S.n == 0) return rad;
    Grid g = grid_of(S.base);
    Grid sv = {S.base.sv, nullptr, (int32_t)S.base.svdims[0], (int32_t)S.base.svdims[1], (int32_t)S.base.svdims[2]};
    const cr_tree_node* N = S.nodes;
    if (S.n == 1) {
        if (hit_box(ro, rd, N[0].bounds[0], N[0].bounds[1])) {
            acc4(rad, beam_radiance(S, g, sv, N[0].left, ro, rd));
            if (beams_tested) (*beams_tested)++;
        }
        return rad;
    }
    uint32_t stack[64];
    stack[0] = NONE;
    uint32_t sp = 1;
    uint32_t cur = 0;
    do {
        if (nodes_visited) (*nodes_visited)++;
        uint32_t cl = N[cur].left, cr = N[cur].right;
        bool hl = hit_box(ro, rd, N[cl].bounds[0], N[cl].bounds[1]);
        bool hr = hit_box(ro, rd, N[cr].bounds[0], N[cr].bounds[1]);
        bool ll = (N[cl].isLeaf & 0xffu) != 0, lr = (N[cr].isLeaf & 0xffu) != 0;
        if (hl && ll) { acc4(rad, beam_radiance(S, g, sv, N[cl].left, ro, rd)); if (beams_tested) (*beams_tested)++; }
        if (hr && lr) { acc4(rad, beam_radiance(S, g, sv, N[cr].left, ro, rd)); if (beams_tested) (*beams_tested)++; }
        bool tl = hl && !ll, tr = hr && !lr;
        if (!tl && !tr) {
            sp--;
            cur = stack[sp];
        } else {
            cur = tl ? cl : cr;
            if (tl && tr) { stack[sp] = cr; sp++; }
        }
    } while (cur != NONE);
    return rad;
}

} // namespace

extern "C" {

void or_trace_beams_count(const or_scene* s, const cr_beam_trace_params* p, uint32_t* segs, uint32_t* datas)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)p->numPaths; i++) {
        PathOut o = trace_path(*s, *p, (uint32_t)i, nullptr, nullptr, 0, 0, 0, 0);
        segs[i] = o.segs;
        datas[i] = o.datas;
    }
}

void or_trace_beams_write(const or_scene* s, const cr_beam_trace_params* p, const uint64_t* seg_offsets,
                          const uint64_t* data_offsets, cr_photon_beam* beams, cr_photon_beam_data* data,
                          uint64_t max_beams, uint64_t max_data)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)p->numPaths; i++) {
        if (seg_offsets[i] >= max_beams) continue;
        trace_path(*s, *p, (uint32_t)i, beams, data, seg_offsets[i], data_offsets[i], max_beams, max_data);
    }
}

void or_gather_pixel(const or_beam_scene* S, int32_t x, int32_t y, int32_t seed, float out[4],
                     uint64_t* nodes_visited, uint64_t* beams_tested)
{
    f3 ro, rd;
    gather_ray(*S, x, y, seed, ro, rd);
    f4 r = traverse(*S, ro, rd, nodes_visited, beams_tested);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

/* The same estimate as a plain loop over every beam in sorted order. */
void or_gather_pixel_brute(const or_beam_scene* S, int32_t x, int32_t y, int32_t seed, float out[4])
{
    f3 ro, rd;
    gather_ray(*S, x, y, seed, ro, rd);
    Grid g = grid_of(S->base);
    Grid sv = {S->base.sv, nullptr, (int32_t)S->base.svdims[0], (int32_t)S->base.svdims[1], (int32_t)S->base.svdims[2]};
    f4 r = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < S->n; i++) acc4(r, beam_radiance(*S, g, sv, i, ro, rd));
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

/* RenderTechniquePPB::UpdateRadius, RenderTechniquePPB.cpp:530-546, with the
 * C++ conversions the reference's declarations imply: the product of two
 * unsigned ints stored in an int, int + unsigned (unsigned), unsigned + float. */
float or_ppb_update_radius(uint32_t frameNumber, uint32_t beamsPerPass, float alpha, float initialRadius,
                           float pmRadius)
{
    if (frameNumber <= 1) return initialRadius;
    float sum = 0;
    int constant = (int)((frameNumber - 2) * beamsPerPass);
    for (unsigned int j = 1; j <= beamsPerPass; j++) sum += (constant + j + alpha) / (constant + j + 1.0f);
    return pmRadius * sum;
}

void or_gather_render(const or_beam_scene* S, const cr_push_constants* frames, uint32_t n_frames, float* accum,
                      int32_t threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    int32_t W = 2 * S->base.cam.halfWidth, H = 2 * S->base.cam.halfHeight;
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t y = 0; y < H; y++)
        for (int32_t x = 0; x < W; x++) {
            float* a = accum + 4 * ((uint64_t)y * W + x);
            for (uint32_t f = 0; f < n_frames; f++) {
                float L[4];
                or_gather_pixel(S, x, y, frames[f].seed, L, nullptr, nullptr);
                float nn = (float)frames[f].frameCount;
                float d = (float)(frames[f].frameCount + 1u);
                for (int c = 0; c < 4; c++) a[c] = (L[c] + a[c] * nn) / d;
            }
        }
}

} // extern "C"
