// Photon-beam producer (shaders/PPB_PT.comp) and beam gather
// (shaders/PPB_PE.comp) for gfx950.
//
// Producer: one lane per sun photon path.  The reference appends beams with
// atomics (PPB_PT.comp:340, :381), so its output order depends on timing;
// here a counting launch and a writing launch run the same path (same RNG
// stream) and a scan between them gives every path its slots, so the beam
// array is a pure function of the inputs.  Corrections: sun photons start on
// the face the light enters and the face is chosen with if / else-if;
// each segment keeps its offset along the parent beam in _padding_beam[0].
//
// Gather: one lane per pixel; the traversal stack (64 entries, the
// reference's depth, PPB_PE.comp:414) lives in LDS; nodes are 64-B records
// holding both child boxes, so each traversal step reads one line.  The
// visit order and the order in which leaf contributions are added are the
// reference's (PPB_PE.comp:419-463).  The radiance term follows Jarosz et al.
// 2011 (beam x beam, 1D kernel) or, with refCompat, PPB_PE.comp:394-403 as
// written; boxes are tested with half their size (the reference passes the
// full size to intersectBox, PPB_PE.comp:424-431).
#include "cr_device.h"
#include "cr_internal_ppb.h"

using namespace crd;

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
// Ceiling on delta-tracking steps of one walk along a beam (a walk through the
// unit box with majorant 64 takes ~64 steps); it only bounds the loop.
constexpr uint32_t TRACK_CEILING = 1u << 16;

template <typename T, int LAYOUT>
__device__ __forceinline__ float cloud_at(const void* grid, int32_t nx, int32_t ny, int32_t nz, int32_t nbx, int32_t nby,
                                          const cr_cloud_properties& cp, f3 p)
{
    GridView g;
    g.data = grid; g.nx = nx; g.ny = ny; g.nz = nz; g.nbx = nbx; g.nby = nby;
    return trilinear<T, LAYOUT>(g, cloud_norm(cp, p));
}

// scatterRay of PPB_PT.comp:268-304 (with the final normalize).
__device__ __forceinline__ f3 scatter_norm(const cr_parameters& p, f3 n, uint32_t& rng)
{
    f3 d = scatter(p, n, rng);
    return normalize(d);
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton30(f3 c, const float* b0, const float* b1)
{
    const float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float n = fdiv(cc[k] - b0[k], b1[k] - b0[k]);
        q[k] = (uint32_t)fminf(fmaxf(n * 1024.0f, 0.0f), 1023.0f);
    }
    return (expand_bits10(q[0]) << 2) + (expand_bits10(q[1]) << 1) + expand_bits10(q[2]);
}

template <typename T, int LAYOUT, bool WRITE>
__global__ void __launch_bounds__(256) trace_kernel(cr::TraceArgs a)
{
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= a.p.numPaths) return;
    uint32_t segCount = 0, dataCount = 0;
    uint64_t segBase = 0, dataBase = 0;
    if (WRITE) {
        segBase = a.segOffsets[id];
        dataBase = a.dataOffsets[id];
        if (segBase >= a.maxBeams) return;
    }
    const cr_cloud_properties& cp = a.cloud;
    if (cp.densityScaling > 0.0f) {
        const float* b0 = cp.bounds[0];
        const float* b1 = cp.bounds[1];
        uint32_t rng;
        xs_init(rng, (uint32_t)a.p.seed * id);
        const f3 L = ld3(a.p.lightDirection);
        // generateSunRay — PPB_PT.comp:396-430 (corrected)
        const float dx = b1[0] - b0[0], dy = b1[1] - b0[1], dz = b1[2] - b0[2];
        const float ax = (dz * dy) * fabsf(L.x), ay = (dx * dz) * fabsf(L.y), az = (dy * dx) * fabsf(L.z);
        const float total = (ax + ay) + az;
        const float sel = xs_float(rng) * total;
        const f3 c1 = mk(L.x > 0.0f ? b0[0] : b1[0], L.y > 0.0f ? b0[1] : b1[1], L.z > 0.0f ? b0[2] : b1[2]);
        f3 pos;
        {
            const float u = xs_float(rng), v = xs_float(rng);
            if (sel < ax) pos = mk(c1.x, b0[1] * (1.0f - u) + b1[1] * u, b0[2] * (1.0f - v) + b1[2] * v);
            else if (sel < ax + ay) pos = mk(b0[0] * (1.0f - u) + b1[0] * u, c1.y, b0[2] * (1.0f - v) + b1[2] * v);
            else pos = mk(b0[0] * (1.0f - u) + b1[0] * u, b0[1] * (1.0f - v) + b1[1] * v, c1.z);
        }
        const float pdf = fdiv(1.0f, total);
        const float power = fdiv(1.0f * a.p.lightIntensity, (float)a.p.numPaths * pdf);
        f3 dir = L;
        const float r = a.p.radius;
        const float sigmaBar = fdiv(cp.maxExtinction * cp.densityScaling, cp.baseScaling);
        const float absorption = 1.0f - a.p.albedo;
        for (uint32_t it = 0; it < a.p.maxInteractions; it++) {
            float tmin = 0.0f, tmax = 0.0f;
            if (!intersect_box(b0, b1, pos, dir, tmin, tmax)) break;
            const f3 exitp = add(pos, mul(dir, tmax));
            const float len = length(sub(pos, exitp));
            const uint64_t di = dataBase + dataCount;
            dataCount++;
            // 16 transmittance distances — PPB_PT.comp:343-374
            float tr[16];
#pragma unroll 1
            for (int s = 0; s < 16; s++) {
                float t = 0.0f;
                for (uint32_t step = 0; step < TRACK_CEILING; step++) {
                    const float zeta = xs_float(rng);
                    t = t + fdiv(-crd_logf(zeta), sigmaBar);
                    if (!(t <= len)) break;
                    const f3 cur = add(pos, mul(dir, t));
                    const float e = cloud_at<T, LAYOUT>(a.grid, a.nx, a.ny, a.nz, a.nbx, a.nby, cp, cur);
                    const float xi = xs_float(rng);
                    if (xi < fdiv(e, cp.maxExtinction)) break;
                }
                if (WRITE) tr[s] = t;
            }
            if (WRITE && di < a.maxData) {
                cr_photon_beam_data D;
                D.power[0] = D.power[1] = D.power[2] = D.power[3] = power;
#pragma unroll
                for (int s = 0; s < 16; s++) D.trDistances[s] = tr[s];
                a.data[di] = D;
            }
            const uint32_t segs = (uint32_t)fdiv(len, r) + 1u;
            const float segSize = fdiv(len, (float)segs);
            if (WRITE) {
                for (uint32_t i = 0; i < segs; i++) {
                    const uint64_t bi = segBase + segCount + i;
                    if (bi >= a.maxBeams) break;
                    const float off = (float)i * segSize;
                    const f3 st = add(pos, mul(dir, off));
                    const f3 en = add(st, mul(dir, segSize));
                    cr_photon_beam B;
                    B.startPos[0] = st.x; B.startPos[1] = st.y; B.startPos[2] = st.z;
                    B.endPos[0] = en.x; B.endPos[1] = en.y; B.endPos[2] = en.z;
                    B.radius = r;
                    B.mortonCode = morton30(mk(fdiv(st.x + en.x, 2.0f), fdiv(st.y + en.y, 2.0f), fdiv(st.z + en.z, 2.0f)), b0, b1);
                    B.dataIdx = (uint32_t)di;
                    B._padding_beam[0] = __float_as_uint(off);
                    B._padding_beam[1] = 0u;
                    B._padding_beam[2] = 0u;
                    a.beams[bi] = B;
                }
            }
            segCount += segs;
            // next real collision — PPB_PT.comp:473-494
            float t = 0.0f;
            f3 cur = pos;
            bool left = true;
            for (uint32_t step = 0; step < TRACK_CEILING; step++) {
                const float zeta = xs_float(rng);
                t = t + fdiv(-crd_logf(zeta), sigmaBar);
                if (!(t <= len)) { left = true; break; }
                cur = add(pos, mul(dir, t));
                const float e = cloud_at<T, LAYOUT>(a.grid, a.nx, a.ny, a.nz, a.nbx, a.nby, cp, cur);
                const float xi = xs_float(rng);
                if (xi < fdiv(e, cp.maxExtinction)) { left = false; break; }
            }
            if (left) break;
            pos = cur;
            const float xi = xs_float(rng);
            if (!(xi > absorption)) break;
            dir = scatter_norm(a.params, dir, rng);
        }
    }
    if (!WRITE) {
        a.segCounts[id] = segCount;
        a.dataCounts[id] = dataCount;
    }
}

// A ray with a direction component exactly 0 keeps that coordinate for every
// t, so it misses every box whose slab on that axis does not hold it; the
// slab formula cannot tell (m = inf gives inf - inf = NaN, which fmaxf / fminf
// drop), so without this test such a ray culls nothing on that axis (the
// centre pixel's ray, direction (0, 0, 1), would visit the whole tree).  The
// margin covers the rounding of the center / half-size form of the box.
__device__ __forceinline__ bool outside_slab(float ro, float d, float cen, float rad)
{
    return d == 0.0f && fabsf(ro - cen) > rad + 1e-6f * (fabsf(ro) + fabsf(cen) + rad);
}

// intersectBox (PPB_PE.comp:282-300) with rad = half the box size; the box as
// center and half extent, m = 1 / rd formed once per ray; plus, for a ray
// with a zero direction component (par), the test of the axes it is parallel
// to (outside_slab is false on every axis of any other ray).
__device__ __forceinline__ bool hit_box(f3 ro, f3 rd, f3 m, bool par, const float* cen3, const float* rad3)
{
    const f3 cen = mk(cen3[0], cen3[1], cen3[2]);
    const f3 rad = mk(rad3[0], rad3[1], rad3[2]);
    const f3 n = mk(m.x * (ro.x - cen.x), m.y * (ro.y - cen.y), m.z * (ro.z - cen.z));
    const f3 k = mk(fabsf(m.x) * rad.x, fabsf(m.y) * rad.y, fabsf(m.z) * rad.z);
    const float tN = fmaxf(fmaxf(-n.x - k.x, -n.y - k.y), -n.z - k.z);
    const float tF = fminf(fminf(-n.x + k.x, -n.y + k.y), -n.z + k.z);
    bool parallelMiss = false;
    if (par)
        parallelMiss = outside_slab(ro.x, rd.x, cen.x, rad.x) || outside_slab(ro.y, rd.y, cen.y, rad.y) ||
                       outside_slab(ro.z, rd.z, cen.z, rad.z);
    return !(tN > tF || tF < 0.0f || parallelMiss);
}

// accumulateRadiance — PPB_PE.comp:362-404
template <typename T, int LAYOUT>
__device__ __forceinline__ void beam_radiance(const cr::GatherArgs& a, uint32_t bi, f3 ro, f3 rd, f4& acc)
{
    const cr::SortedBeam B = a.beams[bi];
    const f3 st = ld3(B.startPos), en = ld3(B.endPos);
    {
        // Skip a beam whose line passes the camera ray's line farther than its
        // radius plus a margin: the test below measures the distance between a
        // point of the camera ray and a point of the beam's line, never less
        // than the distance between the lines, so it would reject the beam too.
        // The margin (1e-4 of the coordinates' magnitude, plus 1e-3 of the
        // radius) exceeds the rounding of both computations many times over;
        // a NaN anywhere keeps the beam for the full test.
        const f3 u = sub(en, st), w = sub(st, ro);
        const f3 nrm = cross(rd, u);
        const float n2 = dot(nrm, nrm), wn = dot(w, nrm);
        const float mag = fmaxf(fmaxf(fmaxf(fabsf(ro.x), fabsf(ro.y)), fmaxf(fabsf(ro.z), fabsf(st.x))),
                                fmaxf(fmaxf(fabsf(st.y), fabsf(st.z)), fmaxf(fmaxf(fabsf(en.x), fabsf(en.y)), fabsf(en.z))));
        const float r = B.radius * 1.001f + 1e-4f * (mag + 1.0f);
        if (wn * wn > (r * r) * n2) return;
    }
    const f3 bd = normalize(sub(en, st));
    if (fabsf(dot(rd, bd)) > 0.9f) return;
    const f3 tangent = cross(rd, bd);
    const f3 pn = cross(bd, tangent);
    const float t = fdiv(dot(sub(st, ro), pn), dot(pn, rd));
    if (!a.refCompat && !(t >= 0.0f)) return;
    const f3 x = add(ro, mul(rd, t));
    const float blen = length(sub(en, st));
    const float s = dot(sub(x, st), bd);
    if (s < 0.0f || s > blen) return;
    const f3 proj = add(st, mul(bd, s));
    const float dist = length(sub(x, proj));
    if (dist > B.radius) return;
    const float q = fdiv(dist, B.radius);
    const float sq = 1.0f - q * q;
    float kern = ((15.0f / 16.0f) * sq) * sq;
    const float dens = cloud_at<T, LAYOUT>(a.grid, a.nx, a.ny, a.nz, a.nbx, a.nby, a.cloud, x);
    GridView svg;
    svg.data = a.sv; svg.nx = a.svx; svg.ny = a.svy; svg.nz = a.svz; svg.nbx = 0; svg.nby = 0;
    const float tcam = trilinear<float, LAYOUT_DENSE>(svg, sv_norm(a.svp, x));
    const cr_photon_beam_data& D = a.data[B.dataIdx];
    const float soff = a.refCompat ? s : (a.beamOff[bi] + s);
    uint32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (D.trDistances[i] >= soff) cnt++;
    const float tb = fdiv((float)cnt, 16.0f);
    const f3 nrd = neg(rd);
    const float ph = phase_pdf(a.params, bd, nrd);
    const float cosA = dot(bd, nrd);
    float sinA, sc;
    if (a.refCompat) {
        sinA = 1.0f - cosA * cosA;
        sc = dens;
    } else {
        kern = fdiv(kern, B.radius);
        sinA = fsqrt(1.0f - cosA * cosA);
        sc = a.albedo * fdiv(dens * a.cloud.densityScaling, a.cloud.baseScaling);
    }
    const float w = fdiv(ph, sinA);
    const float ks = kern * sc;
    acc.x = acc.x + (((ks * D.power[0]) * tcam) * tb) * w;
    acc.y = acc.y + (((ks * D.power[1]) * tcam) * tb) * w;
    acc.z = acc.z + (((ks * D.power[2]) * tcam) * tb) * w;
    acc.w = acc.w + (((ks * D.power[3]) * tcam) * tb) * w;
}

constexpr int GATHER_THREADS = 128;
constexpr int STACK = 64;

template <typename T, int LAYOUT>
__global__ void __launch_bounds__(GATHER_THREADS) gather_kernel(cr::GatherArgs a)
{
    __shared__ uint32_t stack_mem[STACK * GATHER_THREADS];
    uint32_t* stack = stack_mem + threadIdx.x;  // entry e at stack[e * GATHER_THREADS]: bank-conflict free
    const int32_t x = (int32_t)(blockIdx.x * 16 + (threadIdx.x & 15));
    const int32_t y = (int32_t)(blockIdx.y * 8 + (threadIdx.x >> 4));
    const bool inside = x < a.W && y < a.H;
    uint32_t visited = 0, tested = 0, rays = 0;
    if (inside) {
        float4 o = a.image[(size_t)y * a.W + x];
        f4 acc;
        acc.x = o.x; acc.y = o.y; acc.z = o.z; acc.w = o.w;
        const uint32_t id = (uint32_t)x + (uint32_t)y * a.padded;
        for (uint32_t f = 0; f < a.nFrames; f++) {
            const int32_t seed = a.frames[2 * f];
            const uint32_t fc = (uint32_t)a.frames[2 * f + 1];
            float dx = (float)(x - a.cam.halfWidth), dy = (float)(y - a.cam.halfHeight);
            if (a.jitter) {
                uint32_t rng;
                xs_init(rng, (uint32_t)seed * id);
                dx = dx + (xs_float(rng) - 0.5f);
                dy = dy + (xs_float(rng) - 0.5f);
            }
            f3 ro, rd;
            camera_ray(a.cam, dx, dy, ro, rd);
            rays++;
            f4 L;
            L.x = 0.0f; L.y = 0.0f; L.z = 0.0f; L.w = 0.0f;
            const f3 m = mk(fdiv(1.0f, rd.x), fdiv(1.0f, rd.y), fdiv(1.0f, rd.z));
            const bool par = rd.x == 0.0f || rd.y == 0.0f || rd.z == 0.0f;
            if (a.n == 1) {
                // (lo + hi) / 2: halving is exact, so * 0.5 is the same value
                const float cen[3] = {(a.leafLo[0] + a.leafHi[0]) * 0.5f, (a.leafLo[1] + a.leafHi[1]) * 0.5f,
                                      (a.leafLo[2] + a.leafHi[2]) * 0.5f};
                const float rad[3] = {(a.leafHi[0] - a.leafLo[0]) * 0.5f, (a.leafHi[1] - a.leafLo[1]) * 0.5f,
                                      (a.leafHi[2] - a.leafLo[2]) * 0.5f};
                if (hit_box(ro, rd, m, par, cen, rad)) {
                    beam_radiance<T, LAYOUT>(a, 0, ro, rd, L);
                    tested++;
                }
            } else if (a.n > 1) {
                const uint32_t inner = a.n - 1;
                stack[0] = NONE;
                uint32_t sp = 1;
                uint32_t cur = 0;
                // every internal node is visited at most once per ray; the
                // bound only protects against a malformed tree
                uint32_t budget = a.n;
                do {
                    if (budget-- == 0u || cur >= inner) {
                        atomicOr(a.overflow, 2u);
                        break;
                    }
                    visited++;
                    const cr::TravNode nd = a.nodes[cur];
                    const bool hl = hit_box(ro, rd, m, par, nd.cen[0], nd.rad[0]);
                    const bool hr = hit_box(ro, rd, m, par, nd.cen[1], nd.rad[1]);
                    const bool ll = (nd.leafMask & 1u) != 0, lr = (nd.leafMask & 2u) != 0;
                    // The node's hit leaves, left before right, through one copy
                    // of the beam test: a wave whose lanes hit only left leaves
                    // and only right leaves runs it once, not once per side.
                    uint32_t first = (hl && ll) ? nd.child[0] : ((hr && lr) ? nd.child[1] : NONE);
                    const uint32_t second = (hl && ll && hr && lr) ? nd.child[1] : NONE;
#pragma unroll 1
                    for (int k = 0; k < 2; k++) {
                        if (first != NONE) {
                            beam_radiance<T, LAYOUT>(a, first - inner, ro, rd, L);
                            tested++;
                        }
                        first = second;
                    }
                    const bool tl = hl && !ll, tr = hr && !lr;
                    if (!tl && !tr) {
                        sp--;
                        cur = stack[sp * GATHER_THREADS];
                    } else {
                        cur = tl ? nd.child[0] : nd.child[1];
                        if (tl && tr) {
                            if (sp < STACK) {
                                stack[sp * GATHER_THREADS] = nd.child[1];
                                sp++;
                            } else {
                                atomicOr(a.overflow, 1u);
                            }
                        }
                    }
                } while (cur != NONE);
            }
            // running mean — PPB_PE.comp:482-486
            const float nn = (float)fc;
            const float d = (float)(fc + 1u);
            acc.x = fdiv(L.x + acc.x * nn, d);
            acc.y = fdiv(L.y + acc.y * nn, d);
            acc.z = fdiv(L.z + acc.z * nn, d);
            acc.w = fdiv(L.w + acc.w * nn, d);
        }
        float4 o2;
        o2.x = acc.x; o2.y = acc.y; o2.z = acc.z; o2.w = acc.w;
        a.image[(size_t)y * a.W + x] = o2;
    }
    const unsigned long long s0 = wave_sum_u64(rays), s1 = wave_sum_u64(visited), s2 = wave_sum_u64(tested);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(a.counters + 0, s0);
        atomicAdd(a.counters + 1, s1);
        atomicAdd(a.counters + 2, s2);
    }
}

} // namespace

namespace cr {

hipError_t launch_trace_count(const TraceArgs& a, int dtype, int layout, hipStream_t s)
{
    dim3 g((a.p.numPaths + 255) / 256);
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((trace_kernel<__half, LAYOUT_BRICK4, false>), g, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((trace_kernel<__half, LAYOUT_DENSE, false>), g, dim3(256), 0, s, a);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((trace_kernel<float, LAYOUT_BRICK4, false>), g, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((trace_kernel<float, LAYOUT_DENSE, false>), g, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_trace_write(const TraceArgs& a, int dtype, int layout, hipStream_t s)
{
    dim3 g((a.p.numPaths + 255) / 256);
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((trace_kernel<__half, LAYOUT_BRICK4, true>), g, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((trace_kernel<__half, LAYOUT_DENSE, true>), g, dim3(256), 0, s, a);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((trace_kernel<float, LAYOUT_BRICK4, true>), g, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((trace_kernel<float, LAYOUT_DENSE, true>), g, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs& a, int dtype, int layout, hipStream_t s)
{
    dim3 g((unsigned)((a.W + 15) / 16), (unsigned)((a.H + 7) / 8));
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((gather_kernel<__half, LAYOUT_BRICK4>), g, dim3(GATHER_THREADS), 0, s, a);
        else hipLaunchKernelGGL((gather_kernel<__half, LAYOUT_DENSE>), g, dim3(GATHER_THREADS), 0, s, a);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((gather_kernel<float, LAYOUT_BRICK4>), g, dim3(GATHER_THREADS), 0, s, a);
        else hipLaunchKernelGGL((gather_kernel<float, LAYOUT_DENSE>), g, dim3(GATHER_THREADS), 0, s, a);
    }
    return hipGetLastError();
}

} // namespace cr
