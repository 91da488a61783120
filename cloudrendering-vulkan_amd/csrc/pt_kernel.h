// The persistent path-tracing kernel for gfx950 (device code shared by the
// render, batched-render and learned-renderer launch files).
//
// Computes, per pixel and per frame, what shaders/PathTracer.comp main
// (:366-425) computes, and blends the frames into the RGBA32F accumulation
// image with the reference's running mean (:420-424).
//
// Shape of the computation on CDNA4:
//  * Persistent grid: a few waves per SIMD loop until a global pixel queue is
//    drained.  Lanes claim pixels individually: the wave ballots its idle
//    lanes, one lane advances the queue counter by the popcount, and each
//    lane takes base + its rank among the idle lanes.  Claimed indices walk
//    8x8 pixel tiles so neighbouring lanes trace neighbouring pixels.
//  * Path regeneration: a lane renders all frames of its pixel back to back
//    (frame f uses initializeRandom(seed_f * id) exactly as the reference),
//    keeping the running mean in registers: 32 B of image traffic per pixel
//    per launch.
//  * Flattened state machine: every loop iteration performs at most one
//    density lookup per lane (a delta-tracking step of the camera path, or a
//    ratio-tracking step of a shadow ray); collisions, escapes, frame and
//    pixel changes are handled in a divergent tail.  Lanes never wait at the
//    end of an inner tracking loop for the slowest lane of the wave.
//  * Optional brick-majorant DDA (16^3 bricks with a one-voxel apron) skips
//    empty space and lowers the majorant in thin regions.
// The sequence of floating-point operations and RNG draws of every path is
// the one fixed by the reference shader (in the order the GLSL states it),
// so for the same inputs each pixel's value is reproducible bit for bit.
#pragma once

#include "cr_device.h"
#include "cr_internal.h"

using namespace crd;

namespace {

constexpr uint32_t COLLISION_CEILING = 1u << 20;  // guards a non-terminating walk only
#ifndef PT_DDA_STEPS_PER_PASS
#define PT_DDA_STEPS_PER_PASS 6
#endif
constexpr uint32_t DDA_STEPS_PER_PASS = PT_DDA_STEPS_PER_PASS;  // brick-walk steps per tracking call
// Skip-grid entries the walk fetches together (the current brick and the next
// ones along the ray), so their load latencies overlap.
constexpr int DDA_PREFETCH = 2;
#ifndef PT_TAIL_MIN_LANES
#define PT_TAIL_MIN_LANES 16
#endif
// The event tail runs in a pass only when at least this many lanes of the
// wave (or all of its active lanes) have an event pending; the others wait.
constexpr uint32_t TAIL_MIN_LANES = PT_TAIL_MIN_LANES;

enum : uint32_t { ST_IDLE = 0, ST_ACTIVE = 1, ST_DEAD = 2 };
enum : uint32_t { MODE_PRIMARY = 0, MODE_SHADOW = 1 };

struct Lane {
    uint32_t job;  // batched launch: index into the job table
    uint32_t cost0;  // the lane's work counters when it took its pixel
    int32_t px, py;
    uint32_t id;
    uint32_t f;
    // the pixel's camera ray, the same for every frame: its entry point into
    // the box (or the camera position when it misses the box) and direction
    f3 camPos, camDir;
    bool camMiss;
    // (seed, frameCount) of the next frame, loaded one frame ahead so that
    // starting and blending a frame does not wait on memory
    int32_t seedNext, countNext;
    uint32_t count;  // frameCount of the current frame
    f4 acc;
    uint32_t rng;
    f4 res;
    float thr;
    uint32_t bounces;
    f3 pos, dir;
    float t, dist;
    uint32_t mode;
    float T;      // shadow-ray transmittance
    float ipdf;   // lightIntensity * phase pdf of the pending shadow ray
    float w;      // thr * albedo of the pending shadow ray
    // brick DDA
    int32_t cell[3];
    uint32_t brick;  // cell[0] + mx (cell[1] + my cell[2])
    float tNext[3];
    float tau;
};

struct Counters {
    uint32_t paths, vertices, lookups, svLookups, brickVisits;
};

__device__ __forceinline__ const float* B0(const cr::PTArgs& a) { return a.cloud.bounds[0]; }
__device__ __forceinline__ const float* B1(const cr::PTArgs& a) { return a.cloud.bounds[1]; }

template <typename T, int LAYOUT>
__device__ __forceinline__ float sample_cloud(const cr::PTArgs& a, f3 p)
{
    GridView g;
    g.data = a.grid; g.nx = a.nx; g.ny = a.ny; g.nz = a.nz; g.nbx = a.nbx; g.nby = a.nby;
    return trilinear<T, LAYOUT>(g, cloud_norm(a.cloud, p));
}

__device__ __forceinline__ float sample_sv(const cr::PTArgs& a, f3 p)
{
    GridView g;
    g.data = a.sv; g.nx = a.svx; g.ny = a.svy; g.nz = a.svz; g.nbx = 0; g.nby = 0;
    return trilinear<float, LAYOUT_DENSE>(g, sv_norm(a.svp, p));
}

// Per lane and axis, for the lane's current segment (written once per segment
// by dda_init), kept in LDS to spare registers: the step in t between brick
// boundaries (bw / |direction|, the same quotient every brick step would
// form) and the step in the linear brick index (+stride, -stride or 0; its
// sign is the step in the cell index).  One 8-byte read per brick step.
struct DdaAxis {
    float td;
    int32_t brickStep;
};

__device__ __forceinline__ DdaAxis* dda_lds()
{
    __shared__ DdaAxis d[256 * 3];
    return d;
}

// The walk's set-up multiplies by the reciprocals the box test already formed
// (inv = 1 / dir) and by 1 / bw (a.ibw), instead of dividing.
__device__ __forceinline__ void dda_init(const cr::PTArgs& a, f3 pos, f3 dir, f3 invd, Lane& L)
{
    const float p[3] = {pos.x, pos.y, pos.z};
    const float v[3] = {dir.x, dir.y, dir.z};
    const float iv[3] = {invd.x, invd.y, invd.z};
    const int32_t mc[3] = {a.mx, a.my, a.mz};
    const int32_t stride[3] = {1, a.mx, a.mx * a.my};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float rel = (p[k] - B0(a)[k]) * a.ibw[k];
        int32_t c = (int32_t)floorf(fminf(fmaxf(rel, 0.0f), (float)mc[k]));
        c = clampi(c, 0, mc[k] - 1);
        L.cell[k] = c;
        DdaAxis d;
        if (v[k] > 0.0f) {
            L.tNext[k] = fmaxf(((B0(a)[k] + (float)(c + 1) * a.bw[k]) - p[k]) * iv[k], 0.0f);
            d.td = a.bw[k] * iv[k];
            d.brickStep = stride[k];
        } else if (v[k] < 0.0f) {
            L.tNext[k] = fmaxf(((B0(a)[k] + (float)c * a.bw[k]) - p[k]) * iv[k], 0.0f);
            d.td = a.bw[k] * -iv[k];
            d.brickStep = -stride[k];
        } else {
            L.tNext[k] = __builtin_inff();
            d.td = __builtin_inff();
            d.brickStep = 0;
        }
        dda_lds()[threadIdx.x * 3 + k] = d;
    }
    L.brick = (uint32_t)L.cell[0] + (uint32_t)a.mx * ((uint32_t)L.cell[1] + (uint32_t)a.my * (uint32_t)L.cell[2]);
}

// Advance to the neighbouring brick along the axis of the nearest boundary
// (ties: x before y before z); false when the ray leaves the brick grid.
// Written with selects rather than branches: every lane of a wave takes the
// same instructions whichever axis it steps.
__device__ __forceinline__ bool dda_step(const cr::PTArgs& a, int32_t (&cell)[3], float (&tNext)[3], uint32_t& brick)
{
    const bool sx = tNext[0] <= tNext[1] && tNext[0] <= tNext[2];
    const bool sy = !sx && tNext[1] <= tNext[2];
    const bool sz = !sx && !sy;
    const DdaAxis d = dda_lds()[threadIdx.x * 3 + (sx ? 0 : (sy ? 1 : 2))];
    const float tn = (sx ? tNext[0] : (sy ? tNext[1] : tNext[2])) + d.td;
    const int32_t step = min(max(d.brickStep, -1), 1);
    cell[0] += sx ? step : 0;
    cell[1] += sy ? step : 0;
    cell[2] += sz ? step : 0;
    tNext[0] = sx ? tn : tNext[0];
    tNext[1] = sy ? tn : tNext[1];
    tNext[2] = sz ? tn : tNext[2];
    brick += (uint32_t)d.brickStep;
    return ((uint32_t)cell[0] < (uint32_t)a.mx) & ((uint32_t)cell[1] < (uint32_t)a.my) &
           ((uint32_t)cell[2] < (uint32_t)a.mz);
}

__device__ __forceinline__ bool dda_advance(const cr::PTArgs& a, Lane& L) { return dda_step(a, L.cell, L.tNext, L.brick); }

// The skip grid, two floats per brick: (majorant * densityScaling) /
// baseScaling of a non-empty brick, -d for an empty brick d bricks from the
// nearest non-empty one; then the brick's majorant.  Brick tables hold < 2^29
// entries, so the byte offset fits 32 bits.
__device__ __forceinline__ float2 brick_skip(const cr::PTArgs& a, uint32_t brick)
{
    return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(a.skip) + (brick << 3));
}

// Start a tracking segment from L.pos along `dir` (primary or shadow); invd =
// 1 / dir.  Returns false when the segment misses the box (no RNG draw is made
// then, as findScatterPoint returns before tracking, PathTracer.comp:321-324).
template <int TRACK>
__device__ __forceinline__ bool begin_segment(const cr::PTArgs& a, Lane& L, f3 dir, f3 invd)
{
    float tmax = 0.0f, tmin = 0.0f;
    if (!intersect_box_inv(B0(a), B1(a), L.pos, invd, tmin, tmax)) return false;
    f3 exitp = add(L.pos, mul(dir, tmax));
    L.dist = length(sub(L.pos, exitp));
    L.t = 0.0f;
    if (TRACK == CR_TRACK_BRICK_DDA) {
        dda_init(a, L.pos, dir, invd, L);
        L.tau = -crd_log_normal(xs_float(L.rng));
    }
    return true;
}

enum : uint32_t { EV_NONE = 0, EV_COLLIDE = 1, EV_EXIT = 2 };

// One tracking step: at most one density lookup.  Primary mode: delta
// tracking (PathTracer.comp:335-355); shadow mode: ratio tracking.
template <typename T, int LAYOUT, int TRACK>
__device__ __forceinline__ uint32_t track_step(const cr::PTArgs& a, Lane& L, f3 sdir, Counters& c)
{
    const bool shadow = L.mode == MODE_SHADOW;
    const f3 dir = shadow ? sdir : L.dir;
    if (TRACK == CR_TRACK_BRICK_DDA) {
        float maj = 0.0f, sbHit = 0.0f;
        // Empty bricks (skip entry <= 0) are crossed with no lookup.  Every
        // brick's entry is fetched, two at a time: with the loads overlapped,
        // that is cheaper than testing whether the walk is still inside a cube
        // of bricks the skip grid's distances prove empty.
        // At most DDA_STEPS_PER_PASS bricks per call: a lane crossing a long
        // stretch of thin or empty bricks resumes in the next pass instead of
        // holding its wave's other lanes idle (the walk's state is in Lane).
        bool collided = false;
        for (uint32_t it = 0; !collided; it += DDA_PREFETCH) {
            if (it >= DDA_STEPS_PER_PASS) return EV_NONE;
            // The entries of this brick and the next DDA_PREFETCH - 1 along the
            // ray, loaded back to back (a copy of the walk's state steps ahead;
            // past the grid's edge brick 0 is loaded and never used).
            float2 pre[DDA_PREFETCH];
            {
                int32_t cell[3] = {L.cell[0], L.cell[1], L.cell[2]};
                float tNext[3] = {L.tNext[0], L.tNext[1], L.tNext[2]};
                uint32_t brick = L.brick;
                bool in = true;
                pre[0] = brick_skip(a, brick);
#pragma unroll
                for (int k = 1; k < DDA_PREFETCH; k++) {
                    in = in & dda_step(a, cell, tNext, brick);
                    pre[k] = brick_skip(a, in ? brick : 0u);
                }
            }
#pragma unroll
            for (int k = 0; k < DDA_PREFETCH; k++) {
                // sb = (majorant * densityScaling) / baseScaling, precomputed per
                // brick in the skip grid (the same IEEE division as here); 0 for
                // an empty brick
                const float sb = fmaxf(pre[k].x, 0.0f);
                maj = pre[k].y;
                float tc = fminf(fminf(fminf(L.tNext[0], L.tNext[1]), L.tNext[2]), L.dist);
                float seg = tc - L.t;
                c.brickVisits++;
                if (sb > 0.0f && L.tau < sb * seg) {
                    sbHit = sb;
                    collided = true;
                    break;
                }
                L.tau = L.tau - sb * seg;
                L.t = tc;
                if (!(tc < L.dist)) return EV_EXIT;
                if (!dda_advance(a, L)) return EV_EXIT;
            }
        }
        // the loop leaves only from a brick whose entry it fetched (sb > 0);
        // the division is made once here, not at each unrolled visit
        L.t = L.t + fdiv(L.tau, sbHit);
        f3 cur = add(L.pos, mul(dir, L.t));
        float e = sample_cloud<T, LAYOUT>(a, cur);
        c.lookups++;
        // shadow and primary lanes of a wave share the division and the next
        // free-flight draw (one copy of each instead of one per branch)
        const float q = fdiv(e, maj);
        uint32_t ev = EV_NONE;
        if (shadow) {
            L.T = L.T * (1.0f - q);
            if (L.T == 0.0f) ev = EV_EXIT;
        } else {
            float xi = xs_float(L.rng);
            if (xi < q) {
                L.pos = cur;
                ev = EV_COLLIDE;
            }
        }
        if (ev == EV_NONE) L.tau = -crd_log_normal(xs_float(L.rng));
        return ev;
    }
    float zeta = xs_float(L.rng);
    L.t = L.t + fdiv(-crd_log_normal(zeta), a.sigmaBar);
    if (!(L.t < L.dist)) return EV_EXIT;
    f3 cur = add(L.pos, mul(dir, L.t));
    float e = sample_cloud<T, LAYOUT>(a, cur);
    c.lookups++;
    const float q = fdiv(e, a.cloud.maxExtinction);
    if (shadow) {
        L.T = L.T * (1.0f - q);
        if (L.T == 0.0f) return EV_EXIT;
        return EV_NONE;
    }
    float xi = xs_float(L.rng);
    if (xi < q) {
        L.pos = cur;
        return EV_COLLIDE;
    }
    return EV_NONE;
}

__device__ __forceinline__ void add_scaled(f4& r, const f4& v, float s)
{
    r.x = r.x + v.x * s;
    r.y = r.y + v.y * s;
    r.z = r.z + v.z * s;
    r.w = r.w + v.w * s;
}

__device__ __forceinline__ void add_scalar(f4& r, float s)
{
    r.x = r.x + s;
    r.y = r.y + s;
    r.z = r.z + s;
    r.w = r.w + s;
}

// Blend the finished sample into the running mean (PathTracer.comp:420-424).
__device__ __forceinline__ void blend(const cr::PTArgs& a, Lane& L)
{
    const uint32_t fc = L.count;
    const float n = (float)fc;
    const float d = (float)(fc + 1u);
    L.acc.x = fdiv(L.res.x + L.acc.x * n, d);
    L.acc.y = fdiv(L.res.y + L.acc.y * n, d);
    L.acc.z = fdiv(L.res.z + L.acc.z * n, d);
    L.acc.w = fdiv(L.res.w + L.acc.w * n, d);
}

// What a launch computes per work item.
//   K_RENDER   : pixels, frames blended into the image (the reference's pass).
//   K_RENDER_FEW: the same for launches of a few frames per pixel (the
//                reference's interactive loop renders one), with claims gated
//                as in K_RECORD.
//   K_RECORD   : pixels; each frame's camera path stops at its first real
//                collision, which is recorded as (position, direction, RNG
//                state, flag 1), or, for a path that escapes before any
//                collision, as (radiance, 0, 0, 0, flag 0).
//   K_CONTINUE : recorded collisions; `spp` continuations of the path from the
//                collision (NEE, scattering, further bounces); the first
//                resumes the recorded RNG state, so RECORD + CONTINUE(1)
//                reproduces K_RENDER's sample exactly.  Output: the mean.
enum { K_RENDER = 0, K_RECORD = 1, K_CONTINUE = 2, K_RENDER_FEW = 3 };

constexpr bool renders(int kmode) { return kmode == K_RENDER || kmode == K_RENDER_FEW; }

// Passes whose pixels carry little work (a few frames, camera paths cut at
// their first collision, or the small jobs of a batched launch) claim pixels
// only once CLAIM_MIN_LANES lanes of the wave (or all of its live ones) are
// idle: lanes finish such pixels at scattered times, and the claim (queue
// atomic, job lookup, pixel set-up, camera ray and box test) would otherwise
// run for one or two lanes nearly every pass.  Single renders of many frames
// per pixel claim as soon as one lane is idle.
constexpr bool gated_claims(int kmode, bool batch) { return batch || kmode == K_RECORD || kmode == K_RENDER_FEW; }
constexpr uint32_t CLAIM_MIN_LANES = 16;

__device__ __forceinline__ float* record_of(const cr::PTArgs& a, const Lane& L)
{
    return a.records + ((size_t)L.f * (size_t)a.W * (size_t)a.H + (size_t)L.py * (size_t)a.W + (size_t)L.px) * 8u;
}

// RNG state of continuation s > 0 of a recorded collision (a PCG hash of the
// recorded state and s; never 0, the fixed point of xorshift).
__device__ __forceinline__ uint32_t continuation_seed(uint32_t s0, uint32_t s)
{
    uint32_t state = (s0 ^ (s * 0x9E3779B9u)) * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return ((word >> 22u) ^ word) | 1u;
}

// End of a sample (frame or continuation).
template <int KMODE>
__device__ __forceinline__ void finish_sample(const cr::PTArgs& a, Lane& L)
{
    if (renders(KMODE)) {
        blend(a, L);
    } else if (KMODE == K_CONTINUE) {
        L.acc.x = L.acc.x + L.res.x;
        L.acc.y = L.acc.y + L.res.y;
        L.acc.z = L.acc.z + L.res.z;
        L.acc.w = L.acc.w + L.res.w;
    } else {
        float4* r = reinterpret_cast<float4*>(record_of(a, L));
        r[0] = make_float4(L.res.x, L.res.y, L.res.z, L.res.w);
        r[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// The job table of a batched launch lives in dynamic LDS.
template <bool BATCH> __device__ __forceinline__ cr::PTArgs* job_table();
template <> __device__ __forceinline__ cr::PTArgs* job_table<false>() { return nullptr; }
template <> __device__ __forceinline__ cr::PTArgs* job_table<true>()
{
    extern __shared__ uint64_t pt_job_lds[];
    return reinterpret_cast<cr::PTArgs*>(pt_job_lds);
}

// Last job whose first work index is <= idx (jobs are stored by increasing base).
__device__ __forceinline__ uint32_t find_job(const cr::PTArgs* jobs, uint32_t n, uint32_t idx)
{
    uint32_t lo = 0, hi = n - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (jobs[mid].base <= idx) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// BATCH = false: one job, `a0`.  BATCH = true: the jobs of `b`, one shared
// work queue over the concatenated pixel ranges; every lane follows its own
// job's arguments (read from LDS), so lanes of one wave may render different
// jobs.  Per pixel the arithmetic is the same in both forms.
template <typename T, int LAYOUT, int TRACK, int NEE, bool BATCH, int KMODE>
__global__ void __launch_bounds__(256) pt_kernel(cr::PTArgs a0, cr::BatchArgs b)
{
    cr::PTArgs* const jobs = job_table<BATCH>();
    if (BATCH) {
        const uint32_t words = b.nJobs * (uint32_t)(sizeof(cr::PTArgs) / 4);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(b.jobs);
        uint32_t* dst = reinterpret_cast<uint32_t*>(jobs);
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    uint32_t* const queue = BATCH ? b.queue : a0.queue;
    const uint32_t totalIdx = BATCH ? b.total : a0.totalIdx;
    Lane L;
    Counters c = {0, 0, 0, 0, 0};
    uint32_t state = ST_IDLE;
    const uint32_t lane = lane_id();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    L.mode = MODE_PRIMARY;
    L.job = 0;
    uint32_t ev = EV_NONE;  // the lane's pending event, resolved in the tail

    for (;;) {
        // ---- claim pixels for idle lanes (wave-wide, one atomic per wave)
        const unsigned long long idle = __ballot(state == ST_IDLE);
        if (idle && (!gated_claims(KMODE, BATCH) ||
                     (uint32_t)__popcll(idle) >= min(CLAIM_MIN_LANES, (uint32_t)__popcll(__ballot(state != ST_DEAD))))) {
            uint32_t base = 0;
            const uint32_t leader = (uint32_t)__ffsll((long long)idle) - 1u;
            if (lane == leader) base = atomicAdd(queue, (uint32_t)__popcll(idle));
            base = __shfl(base, (int)leader, 64);
            if (state == ST_IDLE) {
                const uint32_t idx = base + (uint32_t)__popcll(idle & lt_mask);
                if (idx >= totalIdx) {
                    state = ST_DEAD;
                } else if (KMODE == K_CONTINUE) {
                    // one recorded collision; its samples start in the event loop
                    L.px = (int32_t)idx;
                    L.py = 0;
                    L.id = 0;
                    L.acc.x = 0.0f; L.acc.y = 0.0f; L.acc.z = 0.0f; L.acc.w = 0.0f;
                    L.f = 0;
                    state = ST_ACTIVE;
                    ev = EV_NONE;
                    L.mode = 0xFFu;
                } else {
                    const uint32_t j = BATCH ? find_job(jobs, b.nJobs, idx) : 0u;
                    const cr::PTArgs& J = BATCH ? jobs[j] : a0;
                    const uint32_t local = BATCH ? idx - J.base : idx;
                    const uint32_t slot = local >> 6, within = local & 63u;
                    // tiles in the order of the previous launch's cost, dearest first, so the
                    // queue ends on cheap tiles and the last lanes finish together
                    const uint32_t tile = J.tileOrder ? J.tileOrder[slot] : slot;
                    const uint32_t tl = J.tileLog2W, tw = 1u << tl;
                    const int32_t x = (int32_t)((tile % J.tilesX) * tw + (within & (tw - 1u)));
                    const int32_t y = (int32_t)((tile / J.tilesX) * (64u >> tl) + (within >> tl));
                    if (x < J.W && y < J.H) {
                        L.job = j;
                        L.px = x;
                        L.py = y;
                        L.id = (uint32_t)x + (uint32_t)y * J.padded;
                        if (renders(KMODE)) {
                            float4 v = J.image[(size_t)y * J.W + x];
                            L.acc.x = v.x; L.acc.y = v.y; L.acc.z = v.z; L.acc.w = v.w;
                        }
                        if (!BATCH) {
                            f3 cp, cd;
                            camera_ray(J.cam, (float)(x - J.cam.halfWidth), (float)(y - J.cam.halfHeight), cp, cd);
                            float tmax = 0.0f, tmin = 0.0f;
                            L.camMiss = !intersect_box(B0(J), B1(J), cp, cd, tmin, tmax) || tmax < 0.0f ||
                                        J.cloud.densityScaling <= 0.0f;
                            L.camPos = L.camMiss ? cp : add(cp, mul(cd, tmax < tmin ? tmax : tmin));
                            L.camDir = cd;
                        }
                        if (J.nFrames > 0u) {
                            L.seedNext = J.frames[0];
                            L.countNext = J.frames[1];
                        }
                        L.f = 0;
                        L.cost0 = c.lookups + c.brickVisits;
                        state = ST_ACTIVE;
                        ev = EV_NONE;
                        // start frame 0: handled by the event loop below
                        L.mode = 0xFFu;  // marker: needs a new sample
                    }
                }
            }
        }
        if (__ballot(state != ST_DEAD) == 0ull) break;
        if (state != ST_ACTIVE) continue;

        const cr::PTArgs& a = BATCH ? jobs[L.job] : a0;
        const f3 ldir = ld3(a.svp.lightDirection);
        const f3 sdir = neg(ldir);
        const float I = a.params.lightIntensity;
        const float albedo = a.ext.albedo;
        const uint32_t cap = a.ext.bounceCap;

        if (ev == EV_NONE && L.mode != 0xFFu) ev = track_step<T, LAYOUT, TRACK>(a, L, sdir, c);

        // ---- divergent tail: events that need no lookup are resolved here,
        // possibly several in a row (e.g. a camera ray that misses the box).
        // Each pass resolves one event per lane; the work several events end
        // with (sampling a new direction, starting a tracking segment) is done
        // once per pass for every lane that needs it, so lanes handling
        // different events share those instructions.  A wave enters the tail
        // once enough of its lanes have an event (TAIL_MIN_LANES), so its cost
        // is shared by more lanes; until then those lanes keep their event.
        bool needSample = (L.mode == 0xFFu);
        const uint32_t nWant = (uint32_t)__popcll(__ballot(ev != EV_NONE || needSample));
        const uint32_t nActive = (uint32_t)__popcll(__ballot(true));
        if (nWant < min(TAIL_MIN_LANES, nActive)) continue;
        while (ev != EV_NONE || needSample) {
            bool doScatter = false;  // sample a new direction along which to continue
            uint32_t seg = 0;        // start a segment: 1 along L.dir, 2 toward the sun
            // The ends of a sample share one copy of the background and of the
            // blend: sky = the path left the box (camera rays that miss it take
            // the background as their radiance), fin = the sample is complete.
            bool sky = false, missed = false, fin = false;
            bool cam = false;  // batched: a camera ray, its box test after the shared reciprocal below
            if (needSample) {
                needSample = false;
                if (L.f >= (KMODE == K_CONTINUE ? a.spp : a.nFrames)) {
                    if (renders(KMODE)) {
                        float4 o;
                        o.x = L.acc.x; o.y = L.acc.y; o.z = L.acc.z; o.w = L.acc.w;
                        a.image[(size_t)L.py * a.W + L.px] = o;
                        if (a.tileCost) {
                            const uint32_t tl = a.tileLog2W;
                            const uint32_t tile = ((uint32_t)L.py >> (6u - tl)) * a.tilesX + ((uint32_t)L.px >> tl);
                            atomicAdd(a.tileCost + tile, (c.lookups + c.brickVisits) - L.cost0);
                        }
                    } else if (KMODE == K_CONTINUE) {
                        const float n = (float)a.spp;
                        a.out[L.px] = make_float4(fdiv(L.acc.x, n), fdiv(L.acc.y, n), fdiv(L.acc.z, n), fdiv(L.acc.w, n));
                    }
                    state = ST_IDLE;
                    break;
                }
                if (KMODE == K_CONTINUE) {
                    // continuation L.f of recorded collision L.px
                    const float4* r = reinterpret_cast<const float4*>(a.points + (size_t)L.px * 8u);
                    const float4 r0 = r[0], r1 = r[1];
                    if (!(r1.w > 0.5f)) {
                        // not a collision: the path escaped, nothing to continue
                        a.out[L.px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        state = ST_IDLE;
                        break;
                    }
                    L.pos = mk(r0.x, r0.y, r0.z);
                    L.dir = mk(r0.w, r1.x, r1.y);
                    const uint32_t s0 = __float_as_uint(r1.z);
                    L.rng = L.f == 0 ? s0 : continuation_seed(s0, L.f);
                    L.res.x = 0.0f; L.res.y = 0.0f; L.res.z = 0.0f; L.res.w = 0.0f;
                    L.thr = 1.0f;
                    L.bounces = 0;
                    L.mode = MODE_PRIMARY;
                    c.paths++;
                    ev = EV_COLLIDE;
                    continue;
                }
                // start frame f: PathTracer.comp:368-392 (a single render made the
                // camera ray and its box test when the lane took the pixel; the
                // batched kernel, whose lanes take pixels of small jobs with few
                // frames, measures faster forming them per frame)
                const int32_t seed = L.seedNext;
                L.count = (uint32_t)L.countNext;
                if (L.f + 1u < a.nFrames) {
                    L.seedNext = a.frames[2 * (L.f + 1u)];
                    L.countNext = a.frames[2 * (L.f + 1u) + 1];
                }
                xs_init(L.rng, (uint32_t)seed * L.id);
                if (BATCH) {
                    camera_ray(a.cam, (float)(L.px - a.cam.halfWidth), (float)(L.py - a.cam.halfHeight), L.pos, L.dir);
                } else {
                    L.pos = L.camPos;
                    L.dir = L.camDir;
                }
                c.paths++;
                c.vertices++;
                L.res.x = 0.0f; L.res.y = 0.0f; L.res.z = 0.0f; L.res.w = 0.0f;
                L.thr = 1.0f;
                L.bounces = 0;
                L.mode = MODE_PRIMARY;
                if (BATCH) {
                    cam = true;
                    seg = 1;
                } else if (L.camMiss) {
                    sky = true;
                    missed = true;
                } else {
                    seg = 1;
                }
            } else if (ev == EV_EXIT && L.mode == MODE_PRIMARY) {
                // escape: PathTracer.comp:399-403
                sky = true;
            } else if (ev == EV_EXIT && L.mode == MODE_SHADOW) {
                // shadow ray done: add NEE, then scatter
                add_scalar(L.res, ((L.ipdf * L.T) * L.w));
                L.thr = L.w;
                L.mode = MODE_PRIMARY;
                doScatter = true;
            } else {  // EV_COLLIDE
                if (KMODE == K_RECORD && L.bounces == 0) {
                    // first real collision: record it and end the frame here
                    float4* r = reinterpret_cast<float4*>(record_of(a, L));
                    r[0] = make_float4(L.pos.x, L.pos.y, L.pos.z, L.dir.x);
                    r[1] = make_float4(L.dir.y, L.dir.z, __uint_as_float(L.rng), 1.0f);
                    L.f++;
                    ev = EV_NONE;
                    needSample = true;
                    continue;
                }
                L.bounces++;
                if (L.bounces > cap || L.bounces > COLLISION_CEILING) {
                    fin = true;
                } else {
                    c.vertices++;
                    const float pdf = phase_pdf(a.params, neg(L.dir), ldir);
                    const float w = L.thr * albedo;
                    if (NEE == CR_NEE_RATIO_TRACKING) {
                        // a shadow ray that misses the box keeps T = 1 and draws nothing
                        L.ipdf = (1.0f * I) * pdf;
                        L.w = w;
                        L.T = 1.0f;
                        L.mode = MODE_SHADOW;
                        seg = 2;
                    } else {
                        float Tv;
                        if (NEE == CR_NEE_SHADOW_VOLUME) {
                            Tv = sample_sv(a, L.pos);
                            c.svLookups++;
                        } else {
                            Tv = 0.0f;
                        }
                        add_scalar(L.res, (((1.0f * I) * pdf) * Tv) * w);
                        L.thr = w;
                        doScatter = true;
                    }
                }
            }
            if (doScatter) {
                L.dir = scatter(a.params, L.dir, L.rng);
                seg = 1;
            }
            // the direction of the segment each lane starts and its reciprocal
            // (the batched kernel's camera rays take their box test with it)
            const f3 sd = seg == 2 ? sdir : L.dir;
            const f3 invd = recip(sd);
            if (BATCH) {
                if (cam) {
                    float tmax = 0.0f, tmin = 0.0f;
                    if (!intersect_box_inv(B0(a), B1(a), L.pos, invd, tmin, tmax) || tmax < 0.0f ||
                        a.cloud.densityScaling <= 0.0f) {
                        sky = true;
                        missed = true;
                    } else {
                        L.pos = add(L.pos, mul(L.dir, tmax < tmin ? tmax : tmin));
                    }
                }
            }
            if (sky) {
                const f4 bg = background(a.ext, L.dir);
                if (missed) {
                    L.res = bg;
                } else {
                    add_scaled(L.res, bg, L.thr);
                }
                fin = true;
            }
            if (fin) {
                finish_sample<KMODE>(a, L);
                L.f++;
                ev = EV_NONE;
                needSample = true;
                continue;
            }
            // seg != 0 here: every branch that falls through starts a segment
            ev = begin_segment<TRACK>(a, L, sd, invd) ? EV_NONE : EV_EXIT;
        }
    }

    // ---- counters: one atomic per wave (all lanes have reconverged here)
    unsigned long long s0 = wave_sum_u64(c.paths), s1 = wave_sum_u64(c.vertices), s2 = wave_sum_u64(c.lookups);
    unsigned long long s3 = wave_sum_u64(c.svLookups), s4 = wave_sum_u64(c.brickVisits);
    if (lane == 0) {
        unsigned long long* const cnt = BATCH ? b.counters : a0.counters;
        atomicAdd(cnt + 0, s0);
        atomicAdd(cnt + 1, s1);
        atomicAdd(cnt + 2, s2);
        atomicAdd(cnt + 3, s3);
        atomicAdd(cnt + 4, s4);
    }
}

// ------------------------------------------------------------ launching
using PtFn = void (*)(cr::PTArgs, cr::BatchArgs);

template <typename T, int LAYOUT, bool BATCH, int KMODE>
PtFn pick_tn(uint32_t tr, uint32_t ne)
{
    if (tr == CR_TRACK_BRICK_DDA) {
        if (ne == CR_NEE_RATIO_TRACKING) return pt_kernel<T, LAYOUT, CR_TRACK_BRICK_DDA, CR_NEE_RATIO_TRACKING, BATCH, KMODE>;
        if (ne == CR_NEE_NONE) return pt_kernel<T, LAYOUT, CR_TRACK_BRICK_DDA, CR_NEE_NONE, BATCH, KMODE>;
        return pt_kernel<T, LAYOUT, CR_TRACK_BRICK_DDA, CR_NEE_SHADOW_VOLUME, BATCH, KMODE>;
    }
    if (ne == CR_NEE_RATIO_TRACKING) return pt_kernel<T, LAYOUT, CR_TRACK_GLOBAL_MAJORANT, CR_NEE_RATIO_TRACKING, BATCH, KMODE>;
    if (ne == CR_NEE_NONE) return pt_kernel<T, LAYOUT, CR_TRACK_GLOBAL_MAJORANT, CR_NEE_NONE, BATCH, KMODE>;
    return pt_kernel<T, LAYOUT, CR_TRACK_GLOBAL_MAJORANT, CR_NEE_SHADOW_VOLUME, BATCH, KMODE>;
}

// The kernel instantiation for the grid storage and the render modes.
template <bool BATCH, int KMODE>
PtFn pick(int dtype, int layout, uint32_t tr, uint32_t ne)
{
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) return pick_tn<__half, LAYOUT_BRICK4, BATCH, KMODE>(tr, ne);
        return pick_tn<__half, LAYOUT_DENSE, BATCH, KMODE>(tr, ne);
    }
    if (layout == LAYOUT_BRICK4) return pick_tn<float, LAYOUT_BRICK4, BATCH, KMODE>(tr, ne);
    return pick_tn<float, LAYOUT_DENSE, BATCH, KMODE>(tr, ne);
}

hipError_t launch(PtFn fn, cr::PTArgs a0, cr::BatchArgs b, int blocks, int threads, size_t lds, hipStream_t s)
{
    void* args[] = {&a0, &b};
    return hipLaunchKernel(reinterpret_cast<const void*>(fn), dim3((unsigned)blocks), dim3((unsigned)threads), args, lds, s);
}

int occupancy(PtFn fn, int threads, size_t lds)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, lds) != hipSuccess) return 1;
    return n > 0 ? n : 1;
}

} // namespace
