// Photon mapping (RenderTechniquePPM, RenderTechniquePPM.cpp:229-312) over the
// context's grid, shadow volume and uniforms: photon tracing into a uniform
// photon grid, then the ray-marched radiance estimate into the accumulation
// image.  Host side of kern_ppm.hip.
#include <cmath>
#include <cstring>
#include <vector>

#include "cr_internal_ppb.h"
#include "cr_internal_ppm.h"
#include "cr_runtime.h"

using namespace cr;

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)
#define CR_TRY(call) \
    do { cr_status _s = (call); if (_s != CR_OK) return _s; } while (0)

struct Ppm {
    uint32_t *pathCounts = nullptr, *pathOffsets = nullptr;
    size_t pathCountsCap = 0, pathOffsetsCap = 0;
    cr_photon *deposits = nullptr, *sorted = nullptr;
    size_t depositsCap = 0, sortedCap = 0;
    float4* dirs = nullptr;  // the sorted photons' unit directions
    size_t dirsCap = 0;
    uint32_t *keys = nullptr, *vals = nullptr, *keysTmp = nullptr, *valsTmp = nullptr, *sortCounts = nullptr;
    size_t keysCap = 0, valsCap = 0, keysTmpCap = 0, valsTmpCap = 0, sortCountsCap = 0;
    uint32_t *arrivals = nullptr, *start = nullptr;
    size_t arrivalsCap = 0, startCap = 0;
    int32_t* frames = nullptr;
    size_t framesCap = 0;
    uint32_t nDeposits = 0, nCells = 0;
    bool valid = false;
};

namespace cr {

void ppm_free(cr_context* ctx)
{
    Ppm* p = ctx->ppm;
    if (!p) return;
    void* bufs[] = {p->pathCounts, p->pathOffsets, p->deposits, p->sorted, p->dirs, p->keys, p->vals, p->keysTmp,
                    p->valsTmp, p->sortCounts, p->arrivals, p->start, p->frames};
    for (void* b : bufs)
        if (b) hipFree(b);
    delete p;
    ctx->ppm = nullptr;
}

} // namespace cr

static Ppm* ppm_get(cr_context* ctx)
{
    if (!ctx->ppm) ctx->ppm = new Ppm();
    return ctx->ppm;
}

static const uint32_t MAX_DEPOSITS = 1u << 28;
static const uint64_t MAX_CELLS = 1ull << 27;

extern "C" {

void cr_ppm_params_default(cr_ppm_params* p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->numPaths = 200u * 64u;
    p->sunProbability = 0.5f;
    p->ambientStep = 100.0f;
    p->maxCollisions = 1u << 16;
}

void cr_photon_map_default(cr_photon_map_properties* p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    const float s = 1.0f / sqrtf(2.0f);   // glm::normalize(vec4(1, -1, 0, 0))
    p->lightDirection[0] = s;
    p->lightDirection[1] = -s;
    for (int k = 0; k < 4; k++) {
        p->bounds[0][k] = 0.0f;
        p->bounds[1][k] = 100.0f;
    }
    p->voxelCount[0] = p->voxelCount[1] = p->voxelCount[2] = 100;
    p->voxelSize = (p->bounds[1][0] - p->bounds[0][0]) / (float)p->voxelCount[0];
    p->photonSize = (uint32_t)sizeof(cr_photon);
    p->stepSize = 10.0f;
    p->absorption = 0.0f;
}

// PhotonMapProperties::SetBounds (UniformBuffers.h:225-243): a cube from the
// lower corner with the largest extent as side.
void cr_photon_map_set_bounds(cr_photon_map_properties* p, const float b0[4], const float b1[4])
{
    if (!p || !b0 || !b1) return;
    float size[4];
    for (int k = 0; k < 4; k++) {
        p->bounds[0][k] = b0[k];
        size[k] = b1[k] - b0[k];
    }
    int hi = 0;
    float hv = 1.175494351e-38f;   // FLT_MIN
    for (int i = 0; i < 3; i++)
        if (hv < fabsf(size[i])) {
            hi = i;
            hv = fabsf(size[i]);
        }
    for (int k = 0; k < 4; k++) p->bounds[1][k] = p->bounds[0][k] + size[hi];
    p->voxelSize = (p->bounds[1][0] - p->bounds[0][0]) / (float)p->voxelCount[0];
}

cr_status cr_set_photon_map(cr_context* ctx, const cr_photon_map_properties* p)
{
    CR_CHECK_CTX(ctx);
    if (!p) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null photon map properties");
    if (p->voxelCount[0] <= 0 || p->voxelCount[1] <= 0 || p->voxelCount[2] <= 0)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "voxelCount must be positive");
    if ((uint64_t)p->voxelCount[0] * (uint64_t)p->voxelCount[1] * (uint64_t)p->voxelCount[2] > MAX_CELLS)
        return fail(ctx, CR_ERR_CAPACITY, "a photon grid holds at most 2^27 cells");
    if (!(p->voxelSize > 0.0f) || !(p->stepSize > 0.0f) || !std::isfinite(p->voxelSize) || !std::isfinite(p->stepSize))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "voxelSize and stepSize must be positive and finite");
    if (!(p->absorption >= 0.0f && p->absorption <= 1.0f))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "absorption must be in [0, 1]");
    ctx->pm = *p;
    ctx->hasPm = true;
    ctx->sceneVersion++;
    if (ctx->ppm) ctx->ppm->valid = false;
    return CR_OK;
}

float cr_ppm_update_radius(uint32_t frame_number, float alpha, float initial_radius, float previous_radius)
{
    if (frame_number <= 1) return initial_radius;
    // (frameNumber - 1 + alpha) / frameNumber with the unsigned-to-float conversions of the reference
    return previous_radius * powf(((float)(frame_number - 1u) + alpha) / (float)frame_number, 0.33333f);
}

static cr_status check_params(cr_context* ctx, const cr_ppm_params* p)
{
    if (!p) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null photon-mapping parameters");
    if (p->numPaths == 0 || p->numPaths > (1u << 24)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "numPaths must be in [1, 2^24]");
    if (!(p->sunProbability >= 0.0f && p->sunProbability <= 1.0f))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "sunProbability must be in [0, 1]");
    if (!(p->ambientStep > 0.0f) || !std::isfinite(p->ambientStep))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "ambientStep must be positive and finite");
    return CR_OK;
}

static PpmGrid grid_args(const cr_context* ctx)
{
    const DeviceGrid& g = ctx->grid;
    PpmGrid a;
    a.grid = g.data; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nbx = g.nbx; a.nby = g.nby;
    return a;
}

cr_status cr_ppm_trace(cr_context* ctx, int32_t seed, const cr_ppm_params* prm, uint32_t* deposits_out)
{
    CR_CHECK_CTX(ctx);
    CR_TRY(check_params(ctx, prm));
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!ctx->hasPm) return fail(ctx, CR_ERR_STATE, "no photon map properties (cr_set_photon_map)");
    hipSetDevice(ctx->device);
    Ppm* p = ppm_get(ctx);
    p->valid = false;
    const uint32_t P = prm->numPaths;
    const uint64_t nCells = (uint64_t)ctx->pm.voxelCount[0] * ctx->pm.voxelCount[1] * ctx->pm.voxelCount[2];
    hipStream_t s = ctx->stream;
    CR_TRY(ensure_buffer(ctx, (void**)&p->pathCounts, &p->pathCountsCap, (size_t)P * 4, "path counts"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->pathOffsets, &p->pathOffsetsCap, (size_t)P * 4, "path offsets"));
    PpmTraceArgs a;
    memset(&a, 0, sizeof(a));
    a.g = grid_args(ctx);
    a.cloud = ctx->cloud;
    a.params = ctx->params;
    a.pm = ctx->pm;
    a.ext = ctx->ext;
    a.numPaths = P;
    a.seed = seed;
    a.sunProbability = prm->sunProbability;
    a.maxCollisions = prm->maxCollisions;
    a.counts = p->pathCounts;
    const DeviceGrid& g = ctx->grid;
    CR_HIP(ctx, launch_ppm_trace(a, false, g.dtype, g.layout, s), "photon trace (count)");
    std::vector<uint32_t> counts(P), offsets(P);
    CR_HIP(ctx, hipMemcpyAsync(counts.data(), p->pathCounts, (size_t)P * 4, hipMemcpyDeviceToHost, s), "copy counts");
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    uint64_t n = 0;
    for (uint32_t i = 0; i < P; i++) {
        offsets[i] = (uint32_t)n;
        n += counts[i];
        if (n > MAX_DEPOSITS) return fail(ctx, CR_ERR_CAPACITY, "more than 2^28 photon deposits in one trace");
    }
    const uint32_t nd = (uint32_t)n;
    CR_TRY(ensure_buffer(ctx, (void**)&p->arrivals, &p->arrivalsCap, (size_t)nCells * 4, "cell arrivals"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->start, &p->startCap, (size_t)nCells * 4, "cell starts"));
    CR_HIP(ctx, zero_async(p->arrivals, (size_t)nCells * 4, s), "clear arrivals");
    if (nd > 0) {
        CR_TRY(ensure_buffer(ctx, (void**)&p->deposits, &p->depositsCap, (size_t)nd * sizeof(cr_photon), "deposits"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->sorted, &p->sortedCap, (size_t)nd * sizeof(cr_photon), "sorted photons"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->dirs, &p->dirsCap, (size_t)nd * sizeof(float4), "photon directions"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->keys, &p->keysCap, (size_t)nd * 4, "deposit cells"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->vals, &p->valsCap, (size_t)nd * 4, "deposit order"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->keysTmp, &p->keysTmpCap, (size_t)nd * 4, "sort keys"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->valsTmp, &p->valsTmpCap, (size_t)nd * 4, "sort values"));
        CR_TRY(ensure_buffer(ctx, (void**)&p->sortCounts, &p->sortCountsCap, (size_t)sort_counts_words(nd) * 4,
                             "sort histograms"));
        CR_HIP(ctx, hipMemcpyAsync(p->pathOffsets, offsets.data(), (size_t)P * 4, hipMemcpyHostToDevice, s), "copy offsets");
        a.counts = nullptr;
        a.offsets = p->pathOffsets;
        a.photons = p->deposits;
        a.cells = p->keys;
        a.order = p->vals;
        CR_HIP(ctx, launch_ppm_trace(a, true, g.dtype, g.layout, s), "photon trace (write)");
        CR_HIP(ctx, launch_radix_sort(p->keys, p->vals, p->keysTmp, p->valsTmp, p->sortCounts, nd, s), "sort by cell");
        CR_HIP(ctx, launch_ppm_runs(p->keys, p->vals, nd, p->start, p->arrivals, p->deposits, p->sorted, p->dirs, s),
               "cell runs");
    }
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    p->nDeposits = nd;
    p->nCells = (uint32_t)nCells;
    p->valid = true;
    if (deposits_out) *deposits_out = nd;
    return CR_OK;
}

cr_status cr_ppm_estimate(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, const cr_ppm_params* prm)
{
    CR_CHECK_CTX(ctx);
    CR_TRY(check_params(ctx, prm));
    if (!frames && n_frames) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames");
    if (n_frames == 0) return CR_OK;
    Ppm* p = ctx->ppm;
    if (!p || !p->valid) return fail(ctx, CR_ERR_STATE, "no photon map (cr_ppm_trace)");
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    if (!ctx->svValid) return fail(ctx, CR_ERR_STATE, "the estimate's direct light needs cr_sv_build");
    for (uint32_t f = 0; f < n_frames; f++)
        if (!(frames[f].pmRadius > 0.0f) || !std::isfinite(frames[f].pmRadius))
            return fail(ctx, CR_ERR_INVALID_ARGUMENT, "pmRadius must be positive and finite");
    hipSetDevice(ctx->device);
    std::vector<int32_t> fr(3 * (size_t)n_frames);
    for (uint32_t f = 0; f < n_frames; f++) {
        fr[3 * f] = frames[f].seed;
        fr[3 * f + 1] = (int32_t)frames[f].frameCount;
        memcpy(&fr[3 * f + 2], &frames[f].pmRadius, 4);
    }
    hipStream_t s = ctx->stream;
    CR_TRY(ensure_buffer(ctx, (void**)&p->frames, &p->framesCap, fr.size() * 4, "frames"));
    CR_HIP(ctx, hipMemcpyAsync(p->frames, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, s), "copy frames");
    PpmEstimateArgs a;
    memset(&a, 0, sizeof(a));
    a.g = grid_args(ctx);
    a.sv = ctx->sv;
    a.svx = (int32_t)ctx->svDimsBuilt[0]; a.svy = (int32_t)ctx->svDimsBuilt[1]; a.svz = (int32_t)ctx->svDimsBuilt[2];
    a.cam = ctx->cam;
    a.cloud = ctx->cloud;
    a.params = ctx->params;
    a.svp = ctx->svp;
    a.pm = ctx->pm;
    a.ext = ctx->ext;
    a.ambientStep = prm->ambientStep;
    a.arrivals = p->arrivals;
    a.start = p->start;
    a.photons = p->sorted;
    a.dirs = p->dirs;
    a.frames = p->frames;
    a.nFrames = n_frames;
    a.image = (float4*)ctx->image;
    a.W = ctx->W;
    a.H = ctx->H;
    a.padded = 32u * ((uint32_t)ctx->W / 32u + 1u);
    const DeviceGrid& g = ctx->grid;
    CR_HIP(ctx, launch_ppm_estimate(a, g.dtype, g.layout, s), "photon estimate");
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    return CR_OK;
}

cr_status cr_ppm_download(cr_context* ctx, uint32_t* n_cells, uint32_t* n_deposits, uint32_t* arrivals, cr_photon* photons)
{
    CR_CHECK_CTX(ctx);
    Ppm* p = ctx->ppm;
    if (!p || !p->valid) return fail(ctx, CR_ERR_STATE, "no photon map (cr_ppm_trace)");
    if (n_cells) *n_cells = p->nCells;
    if (n_deposits) *n_deposits = p->nDeposits;
    hipStream_t s = ctx->stream;
    if (arrivals) CR_HIP(ctx, hipMemcpyAsync(arrivals, p->arrivals, (size_t)p->nCells * 4, hipMemcpyDeviceToHost, s), "copy arrivals");
    if (photons && p->nDeposits)
        CR_HIP(ctx, hipMemcpyAsync(photons, p->sorted, (size_t)p->nDeposits * sizeof(cr_photon), hipMemcpyDeviceToHost, s),
               "copy photons");
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    return CR_OK;
}

} // extern "C"
