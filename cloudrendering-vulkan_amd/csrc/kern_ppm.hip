// Photon mapping (RenderTechniquePPM): photon tracing into a uniform grid
// (shaders/PPM_PT.comp) and the ray-marched radiance estimate
// (shaders/PPM_PE.comp) for gfx950.
//
// Tracing: one lane per photon path, run twice over the same RNG stream -- a
// counting launch, then, after a scan gives every path its slots, a writing
// launch -- so the deposits come out in emission order (path id, then
// collision) instead of the reference's atomic arrival order
// (PPM_PT.comp:318).  A stable radix sort by cell then groups them, and a
// cell keeps its first CR_PPM_PHOTONS_PER_CELL photons in that order: the map
// is a pure function of the inputs, and the CPU oracle reproduces it.
//
// Estimate: one lane per pixel in 8x8-pixel workgroups, marching from the
// far side of the cloud toward the camera (PPM_PE.comp:443-545).  Per step:
// shadow-volume direct light, one scattered sky ray marched to the cloud's
// edge, and the photons of the cells the query sphere overlaps.
#include "cr_device.h"
#include "cr_internal_ppm.h"

using namespace crd;

namespace {

constexpr float PI4_3 = (PI_F * 4.0f) / 3.0f;
constexpr float TWO_PI = 6.28318530718f;
// Ceilings that only bound loops a finite input ends anyway.
constexpr uint32_t TRACK_CEILING = 1u << 16;   // delta-tracking steps of one flight
constexpr uint32_t MARCH_CEILING = 1u << 16;   // estimate steps per pixel, sky-ray steps per estimate step

template <typename T, int LAYOUT>
__device__ __forceinline__ float density(const cr::PpmGrid& g, const cr_cloud_properties& cp, f3 p)
{
    GridView v;
    v.data = g.grid; v.nx = g.nx; v.ny = g.ny; v.nz = g.nz; v.nbx = g.nbx; v.nby = g.nby;
    return trilinear<T, LAYOUT>(v, cloud_norm(cp, p));
}

// Entry point on the faces a ray of direction d enters the box through,
// uniform over the box's projected area perpendicular to d; returns the area.
__device__ __forceinline__ float entry_point(const float* b0, const float* b1, f3 d, uint32_t& rng, f3& pos)
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

// Cell of a deposit (depositPhoton, PPM_PT.comp:306-309), clamped to the grid.
__device__ __forceinline__ uint32_t cell_of(const cr_photon_map_properties& pm, f3 p)
{
    const int32_t ix = clampi((int32_t)fdiv(p.x - pm.bounds[0][0], pm.voxelSize), 0, pm.voxelCount[0] - 1);
    const int32_t iy = clampi((int32_t)fdiv(p.y - pm.bounds[0][1], pm.voxelSize), 0, pm.voxelCount[1] - 1);
    const int32_t iz = clampi((int32_t)fdiv(p.z - pm.bounds[0][2], pm.voxelSize), 0, pm.voxelCount[2] - 1);
    return (uint32_t)ix + ((uint32_t)iy + (uint32_t)iz * (uint32_t)pm.voxelCount[1]) * (uint32_t)pm.voxelCount[0];
}

// One photon path (PPM_PT.comp:452-488).  WRITE: deposits go to
// photons/cells from slot `base`; otherwise only counted.
template <typename T, int LAYOUT, bool WRITE>
__global__ void __launch_bounds__(256) ppm_trace_kernel(cr::PpmTraceArgs a)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= a.numPaths) return;
    uint32_t n = 0;
    const uint32_t base = WRITE ? a.offsets[id] : 0u;
    const cr_cloud_properties& cp = a.cloud;
    if (cp.densityScaling > 0.0f) {
        const float* b0 = cp.bounds[0];
        const float* b1 = cp.bounds[1];
        uint32_t rng;
        xs_init(rng, (uint32_t)a.seed * id);
        const float emitted = (float)a.numPaths;
        f3 pos, dir;
        f4 col;
        if (xs_float(rng) < a.sunProbability) {
            // sun: direction L, entry over the projected area, pdf 1 / area
            dir = ld3(a.pm.lightDirection);
            const float pdf = fdiv(1.0f, entry_point(b0, b1, dir, rng, pos));
            const float c = fdiv(fdiv(1.0f * a.params.lightIntensity, pdf * emitted), a.sunProbability) * 0.5f;
            col.x = col.y = col.z = col.w = c;
        } else {
            // sky: uniform direction, entry over the projected area for it, pdf 1 / (4 pi area)
            const float z = 1.0f - 2.0f * xs_float(rng);
            const float ph = TWO_PI * xs_float(rng);
            const float s = fsqrt(fmaxf(0.0f, 1.0f - z * z));
            float sp, cph;
            crd_sincosf(ph, &sp, &cph);
            dir = mk(s * cph, s * sp, z);
            const float pdf = fdiv(INV_4PI, entry_point(b0, b1, dir, rng, pos));
            const f4 bg = background(a.ext, neg(dir));
            const float k = pdf * emitted, q = 1.0f - a.sunProbability;
            col.x = fdiv(fdiv(bg.x, k) * 0.5f, q) * 0.5f;
            col.y = fdiv(fdiv(bg.y, k) * 0.5f, q) * 0.5f;
            col.z = fdiv(fdiv(bg.z, k) * 0.5f, q) * 0.5f;
            col.w = fdiv(fdiv(bg.w, k) * 0.5f, q) * 0.5f;
        }
        const float sigmaBar = fdiv(cp.maxExtinction * cp.densityScaling, cp.baseScaling);
        bool first = true;
        for (uint32_t c = 0; c < a.maxCollisions; c++) {
            // interactWithMedium — PPM_PT.comp:398-446
            float tmin = 0.0f, tmax = 0.0f;
            if (!intersect_box(b0, b1, pos, dir, tmin, tmax)) break;
            const float dist = length(sub(pos, add(pos, mul(dir, tmax))));
            float t = 0.0f;
            f3 cur = pos;
            bool hit = false;
            for (uint32_t step = 0; step < TRACK_CEILING; step++) {
                t = t + fdiv(-crd_logf(xs_float(rng)), sigmaBar);
                if (!(t <= dist)) break;
                cur = add(pos, mul(dir, t));
                const float e = density<T, LAYOUT>(a.g, cp, cur);
                if (xs_float(rng) < fdiv(e, cp.maxExtinction)) { hit = true; break; }
            }
            if (!hit) break;
            pos = cur;
            if (first) {
                first = false;
            } else {
                if (WRITE) {
                    cr_photon ph;
                    ph.position[0] = cur.x; ph.position[1] = cur.y; ph.position[2] = cur.z; ph.position[3] = 0.0f;
                    ph.power[0] = col.x; ph.power[1] = col.y; ph.power[2] = col.z; ph.power[3] = col.w;
                    ph.theta = crd_acosf(fminf(fmaxf(dir.z, -1.0f), 1.0f));
                    const float phi = crd_atan2f(dir.y, dir.x);
                    ph.phi = phi < 0.0f ? phi + TWO_PI : phi;
                    ph._padding[0] = 0.0f;
                    ph._padding[1] = 0.0f;
                    a.photons[base + n] = ph;
                    a.cells[base + n] = cell_of(a.pm, cur);
                    a.order[base + n] = base + n;
                }
                n++;
            }
            if (!(xs_float(rng) > a.pm.absorption)) break;
            dir = normalize(scatter(a.params, dir, rng));
        }
    }
    if (!WRITE) a.counts[id] = n;
}

// Runs of the cell-sorted deposits: a run's first element records where it
// starts, its last one how long it is (the cell's arrivals).
__global__ void __launch_bounds__(256) ppm_run_start_kernel(const uint32_t* keys, uint32_t n, uint32_t* start)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || keys[i - 1] != keys[i]) start[keys[i]] = i;
}

// The sorted copy of each deposit and its unit direction (sin theta cos phi,
// sin theta sin phi, cos theta) — dirFromPolar, PPM_PE.comp:201-207 — formed once here
// instead of by every estimate query that finds the photon in range.
__global__ void __launch_bounds__(256) ppm_run_end_kernel(const uint32_t* keys, const uint32_t* vals, uint32_t n,
                                                          const uint32_t* start, uint32_t* arrivals,
                                                          const cr_photon* photons, cr_photon* sorted, float4* dirs)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = keys[i];
    if (i + 1 == n || keys[i + 1] != c) arrivals[c] = i + 1 - start[c];
    const cr_photon ph = photons[vals[i]];
    sorted[i] = ph;
    float st, ct, sp, cph;
    crd_sincosf(ph.theta, &st, &ct);
    crd_sincosf(ph.phi, &sp, &cph);
    dirs[i] = make_float4(st * cph, st * sp, ct, 0.0f);
}

// samplePhotonMap — PPM_PE.comp:304-375: photons within r of p, weighted by
// the phase function toward `dir`, over the sphere's volume and sigma_s.
__device__ __forceinline__ f4 photon_query(const cr::PpmEstimateArgs& a, f3 p, f3 dir, float scatter, float r)
{
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (scatter == 0.0f) return acc;
    const cr_photon_map_properties& pm = a.pm;
    const f3 g = sub(p, ld3(pm.bounds[0]));
    const int32_t x0 = max((int32_t)fdiv(g.x - r, pm.voxelSize), 0);
    const int32_t x1 = min((int32_t)fdiv(g.x + r, pm.voxelSize), pm.voxelCount[0] - 1);
    const int32_t y0 = max((int32_t)fdiv(g.y - r, pm.voxelSize), 0);
    const int32_t y1 = min((int32_t)fdiv(g.y + r, pm.voxelSize), pm.voxelCount[1] - 1);
    const int32_t z0 = max((int32_t)fdiv(g.z - r, pm.voxelSize), 0);
    const int32_t z1 = min((int32_t)fdiv(g.z + r, pm.voxelSize), pm.voxelCount[2] - 1);
    const float r2 = r * r;
    for (int32_t z = z0; z <= z1; z++)
        for (int32_t y = y0; y <= y1; y++)
            for (int32_t x = x0; x <= x1; x++) {
                const uint32_t c = (uint32_t)x + ((uint32_t)y + (uint32_t)z * (uint32_t)pm.voxelCount[1]) *
                                                     (uint32_t)pm.voxelCount[0];
                const uint32_t cnt = min(a.arrivals[c], CR_PPM_PHOTONS_PER_CELL);
                const uint32_t s = a.start[c];
                for (uint32_t i = 0; i < cnt; i++) {
                    const cr_photon& ph = a.photons[s + i];
                    const f3 d = sub(ld3(ph.position), p);
                    if (dot(d, d) <= r2) {
                        const float4 dv = a.dirs[s + i];
                        const f3 pd = mk(dv.x, dv.y, dv.z);
                        const float f = phase_pdf(a.params, dir, pd);
                        acc.x = acc.x + f * ph.power[0];
                        acc.y = acc.y + f * ph.power[1];
                        acc.z = acc.z + f * ph.power[2];
                        acc.w = acc.w + f * ph.power[3];
                    }
                }
            }
    const float avg = (((0.0f + acc.x) + acc.y) + acc.z) + acc.w;
    if (fdiv(avg, 4.0f) <= 0.0f) {
        acc.x = acc.y = acc.z = acc.w = 0.0f;
        return acc;
    }
    const float v = ((PI4_3 * r) * r2) * scatter;
    acc.x = fdiv(acc.x, v); acc.y = fdiv(acc.y, v); acc.z = fdiv(acc.z, v); acc.w = fdiv(acc.w, v);
    return acc;
}

// PPM_PE.comp main (:443-545) for one pixel and one frame.
template <typename T, int LAYOUT>
__device__ __forceinline__ f4 estimate(const cr::PpmEstimateArgs& a, int32_t x, int32_t y, int32_t seed, float radius)
{
    const cr_cloud_properties& cp = a.cloud;
    const cr_photon_map_properties& pm = a.pm;
    uint32_t rng;
    xs_init(rng, (uint32_t)seed * ((uint32_t)x + (uint32_t)y * a.padded));
    f3 pos, dir;
    camera_ray(a.cam, (float)(x - a.cam.halfWidth), (float)(y - a.cam.halfHeight), pos, dir);
    f4 res = background(a.ext, dir);
    float tmin = 0.0f, tmax = 0.0f;
    if (!intersect_box(cp.bounds[0], cp.bounds[1], pos, dir, tmin, tmax) || !(cp.densityScaling > 0.0f)) return res;
    const f3 closest = add(pos, mul(dir, tmin));
    pos = add(pos, mul(dir, tmax - pm.stepSize * xs_float(rng)));
    const float maxDist = length(sub(pos, closest));
    const float step = pm.stepSize;
    const float albedo = 1.0f - pm.absorption;
    const f3 L = ld3(a.svp.lightDirection);
    const float dsbs = fdiv(cp.densityScaling, cp.baseScaling);
    // the direct light's phase term: the same for every step of the ray
    const float pdf = phase_pdf(a.params, neg(dir), L);
    GridView sv;
    sv.data = a.sv; sv.nx = a.svx; sv.ny = a.svy; sv.nz = a.svz; sv.nbx = 0; sv.nby = 0;
    float cur = 0.0f;
    for (uint32_t k = 0; k < MARCH_CEILING && cur < maxDist; k++) {
        pos = sub(pos, mul(dir, step));
        cur = cur + step;
        const float ext = fdiv(density<T, LAYOUT>(a.g, cp, pos) * cp.densityScaling, cp.baseScaling);
        const float sc = albedo * ext;
        // direct sunlight through the shadow volume
        const float tsun = trilinear<float, LAYOUT_DENSE>(sv, sv_norm(a.svp, pos));
        const float dl = ((1.0f * a.params.lightIntensity) * pdf) * tsun;
        f4 direct = {dl, dl, dl, dl};
        // one sky direction, its transmittance by a Riemann sum to the cloud's edge
        const f3 ad = scatter(a.params, dir, rng);
        f3 q = pos;
        float accD = 0.0f, accL = 0.0f;
        uint32_t samples = 0;
        f4 amb = background(a.ext, ad);
        // Two samples per iteration, their lookups issued together: the second
        // is taken only if the point before it is still inside the box, as
        // the one-sample loop would (a sample outside the box reads clamped
        // texels, so the speculative lookup stays in bounds).  Same points,
        // same sums in the same order.
        const f3 dq = mul(ad, a.ambientStep);
        for (uint32_t j = 0; j < MARCH_CEILING && in_box(cp.bounds[0], cp.bounds[1], q); j += 2) {
            const f3 q1 = add(q, dq);
            const f3 q2 = add(q1, dq);
            const float v1 = density<T, LAYOUT>(a.g, cp, q1);
            const float v2 = density<T, LAYOUT>(a.g, cp, q2);
            accD = accD + v1;
            accL = accL + a.ambientStep;
            samples++;
            if (!(j + 1 < MARCH_CEILING && in_box(cp.bounds[0], cp.bounds[1], q1))) break;
            accD = accD + v2;
            accL = accL + a.ambientStep;
            samples++;
            q = q2;
        }
        accD = accD * dsbs;
        const float ta = crd_expf(fdiv(-accL * accD, (float)samples));
        amb.x = amb.x * ta; amb.y = amb.y * ta; amb.z = amb.z * ta; amb.w = amb.w * ta;
        if (amb.x > 0.0f) {
            direct.x = direct.x + amb.x; direct.y = direct.y + amb.y;
            direct.z = direct.z + amb.z; direct.w = direct.w + amb.w;
        }
        // photons
        const f4 ind0 = photon_query(a, pos, neg(dir), sc, radius);
        const f4 ind = {albedo * ind0.x, albedo * ind0.y, albedo * ind0.z, albedo * ind0.w};
        const float att = crd_expf(-step * ext);
        res.x = res.x * att + (sc * (direct.x + ind.x)) * step;
        res.y = res.y * att + (sc * (direct.y + ind.y)) * step;
        res.z = res.z * att + (sc * (direct.z + ind.z)) * step;
        res.w = res.w * att + (sc * (direct.w + ind.w)) * step;
    }
    return res;
}

template <typename T, int LAYOUT>
__global__ void __launch_bounds__(64) ppm_estimate_kernel(cr::PpmEstimateArgs a)
{
    const int32_t x = (int32_t)(blockIdx.x * 8u + (threadIdx.x & 7u));
    const int32_t y = (int32_t)(blockIdx.y * 8u + (threadIdx.x >> 3));
    if (x >= a.W || y >= a.H) return;
    float4* px = a.image + (uint64_t)y * (uint32_t)a.W + (uint32_t)x;
    float4 acc = *px;
    for (uint32_t f = 0; f < a.nFrames; f++) {
        const int32_t seed = a.frames[3 * f];
        const float fc = (float)(uint32_t)a.frames[3 * f + 1];
        const float d = (float)((uint32_t)a.frames[3 * f + 1] + 1u);
        const float r = __int_as_float(a.frames[3 * f + 2]);
        const f4 L = estimate<T, LAYOUT>(a, x, y, seed, r);
        acc.x = fdiv(L.x + acc.x * fc, d);
        acc.y = fdiv(L.y + acc.y * fc, d);
        acc.z = fdiv(L.z + acc.z * fc, d);
        acc.w = fdiv(L.w + acc.w * fc, d);
    }
    *px = acc;
}

} // namespace

namespace cr {

#define CR_PPM_DISPATCH(KERNEL, GRID, BLOCK, ...)                                                                   \
    do {                                                                                                            \
        if (dtype == CR_DTYPE_F16) {                                                                                \
            if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((KERNEL<__half, LAYOUT_BRICK4 __VA_ARGS__>), GRID, BLOCK, 0, s, a); \
            else hipLaunchKernelGGL((KERNEL<__half, LAYOUT_DENSE __VA_ARGS__>), GRID, BLOCK, 0, s, a);               \
        } else {                                                                                                    \
            if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((KERNEL<float, LAYOUT_BRICK4 __VA_ARGS__>), GRID, BLOCK, 0, s, a);  \
            else hipLaunchKernelGGL((KERNEL<float, LAYOUT_DENSE __VA_ARGS__>), GRID, BLOCK, 0, s, a);                \
        }                                                                                                           \
    } while (0)

hipError_t launch_ppm_trace(const PpmTraceArgs& a, bool write, int dtype, int layout, hipStream_t s)
{
    if (a.numPaths == 0) return hipSuccess;
    const dim3 grid((a.numPaths + 255u) / 256u);
    if (write) CR_PPM_DISPATCH(ppm_trace_kernel, grid, dim3(256), , true);
    else CR_PPM_DISPATCH(ppm_trace_kernel, grid, dim3(256), , false);
    return hipGetLastError();
}

hipError_t launch_ppm_runs(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t* start, uint32_t* arrivals,
                           const cr_photon* photons, cr_photon* sorted, float4* dirs, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255u) / 256u);
    hipLaunchKernelGGL(ppm_run_start_kernel, grid, dim3(256), 0, s, keys, n, start);
    hipLaunchKernelGGL(ppm_run_end_kernel, grid, dim3(256), 0, s, keys, vals, n, start, arrivals, photons, sorted, dirs);
    return hipGetLastError();
}

hipError_t launch_ppm_estimate(const PpmEstimateArgs& a, int dtype, int layout, hipStream_t s)
{
    if (a.W <= 0 || a.H <= 0 || a.nFrames == 0) return hipSuccess;
    const dim3 grid(((uint32_t)a.W + 7u) / 8u, ((uint32_t)a.H + 7u) / 8u);
    CR_PPM_DISPATCH(ppm_estimate_kernel, grid, dim3(64), );
    return hipGetLastError();
}

} // namespace cr
