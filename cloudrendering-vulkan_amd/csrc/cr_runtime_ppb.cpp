// Host runtime of the photon-beam technique (RenderTechniquePPB,
// RenderTechniquePPB.cpp): beam buffers, the deterministic beam producer, the
// LBVH build and the gather.  The reference allocates beam storage for 4096
// beams with no overflow guard (RenderTechniquePPB.h:117, PPB_PT.comp:381);
// here capacities are explicit and exceeding one is an error.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "cr_internal_ppb.h"
#include "cr_runtime.h"

using namespace cr;

struct Ppb {
    uint32_t n = 0, nData = 0;
    cr_photon_beam* beams = nullptr;
    size_t beamsCap = 0;
    cr_photon_beam_data* data = nullptr;
    size_t dataCap = 0;
    SortedBeam* sorted = nullptr;
    size_t sortedCap = 0;
    float* sortedOff = nullptr;
    size_t sortedOffCap = 0;
    uint32_t *keys = nullptr, *vals = nullptr, *keysTmp = nullptr, *valsTmp = nullptr, *counts = nullptr;
    size_t keysCap = 0, valsCap = 0, keysTmpCap = 0, valsTmpCap = 0, countsCap = 0;
    uint4* meta = nullptr;      // per internal node: left, right, range end, parent
    uint32_t *leafParent = nullptr, *arrivals = nullptr, *exits = nullptr;
    size_t metaCap = 0, leafParentCap = 0, arrivalsCap = 0, exitsCap = 0;
    float *lo = nullptr, *hi = nullptr;
    size_t loCap = 0, hiCap = 0;
    TravNode* trav = nullptr;
    size_t travCap = 0;
    unsigned long long* gcounters = nullptr;
    size_t gcountersCap = 0;
    uint32_t* overflow = nullptr;
    size_t overflowCap = 0;
    bool built = false;
    uint32_t builtN = 0;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool gatherPending = false;
    float gatherMs = 0.0f;
};

namespace cr {

void ppb_free(cr_context* ctx)
{
    Ppb* p = ctx->ppb;
    if (!p) return;
    void* bufs[] = {p->beams, p->data, p->sorted, p->sortedOff, p->keys, p->vals, p->keysTmp, p->valsTmp, p->counts, p->meta,
                    p->leafParent, p->arrivals, p->exits, p->lo, p->hi, p->trav, p->gcounters, p->overflow};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (hipEvent_t e : p->ev)
        if (e) hipEventDestroy(e);
    delete p;
    ctx->ppb = nullptr;
}

static cr_status ppb_get(cr_context* ctx, Ppb** out)
{
    if (!ctx->ppb) {
        Ppb* p = new Ppb();
        for (hipEvent_t& e : p->ev) hipEventCreate(&e);
        ctx->ppb = p;
        cr_status st = ensure_buffer(ctx, (void**)&p->gcounters, &p->gcountersCap, 8 * sizeof(unsigned long long), "gather counters");
        if (st != CR_OK) return st;
        st = ensure_buffer(ctx, (void**)&p->overflow, &p->overflowCap, 64, "overflow flag");
        if (st != CR_OK) return st;
        zero_async(p->gcounters, 8 * sizeof(unsigned long long), ctx->stream);
        zero_async(p->overflow, 4, ctx->stream);
    }
    *out = ctx->ppb;
    return CR_OK;
}

} // namespace cr

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)
#define CR_TRY(call) \
    do { cr_status _s = (call); if (_s != CR_OK) return _s; } while (0)

static const uint32_t MAX_BEAMS = 1u << 28;

// Sort, tree and traversal buffers for n beams (grown, never shrunk).
static cr_status lbvh_buffers(cr_context* ctx, Ppb* p, uint32_t n)
{
    const size_t nodes = 2 * (size_t)n - 1;
    CR_TRY(ensure_buffer(ctx, (void**)&p->sorted, &p->sortedCap, (size_t)n * sizeof(SortedBeam), "sorted beams"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->sortedOff, &p->sortedOffCap, (size_t)n * 4, "sorted beam offsets"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->keys, &p->keysCap, (size_t)n * 4, "keys"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->vals, &p->valsCap, (size_t)n * 4, "vals"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->keysTmp, &p->keysTmpCap, (size_t)n * 4, "keys tmp"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->valsTmp, &p->valsTmpCap, (size_t)n * 4, "vals tmp"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->counts, &p->countsCap, (size_t)sort_counts_words(n) * 4, "histograms"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->meta, &p->metaCap, (size_t)n * 16, "node records"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->leafParent, &p->leafParentCap, (size_t)n * 4, "leaf parents"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->arrivals, &p->arrivalsCap, (size_t)n * 4, "arrivals"));
    // refit exit lists (each at most two entries per internal node) and their counts at the end
    CR_TRY(ensure_buffer(ctx, (void**)&p->exits, &p->exitsCap, (size_t)refit_exit_words(n) * 4, "refit exits"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->lo, &p->loCap, nodes * 12, "box lo"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->hi, &p->hiCap, nodes * 12, "box hi"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->trav, &p->travCap, (size_t)n * sizeof(TravNode), "traversal nodes"));
    return CR_OK;
}

extern "C" {

cr_status cr_beams_upload(cr_context* ctx, const cr_photon_beam* beams, uint32_t n, const cr_photon_beam_data* data,
                          uint32_t n_data)
{
    CR_CHECK_CTX(ctx);
    if ((n && !beams) || (n_data && !data)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null beams or data");
    if (n > MAX_BEAMS || n_data > MAX_BEAMS) return fail(ctx, CR_ERR_CAPACITY, "beam count exceeds 2^28");
    for (uint32_t i = 0; i < n; i++)
        if (beams[i].dataIdx >= n_data) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "beam dataIdx out of range");
    hipSetDevice(ctx->device);
    Ppb* p;
    CR_TRY(ppb_get(ctx, &p));
    CR_TRY(ensure_buffer(ctx, (void**)&p->beams, &p->beamsCap, (size_t)n * sizeof(cr_photon_beam), "beams"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->data, &p->dataCap, (size_t)n_data * sizeof(cr_photon_beam_data), "beam data"));
    if (n) CR_HIP(ctx, hipMemcpyAsync(p->beams, beams, (size_t)n * sizeof(cr_photon_beam), hipMemcpyHostToDevice, ctx->stream), "copy beams");
    if (n_data)
        CR_HIP(ctx, hipMemcpyAsync(p->data, data, (size_t)n_data * sizeof(cr_photon_beam_data), hipMemcpyHostToDevice, ctx->stream),
               "copy beam data");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    p->n = n;
    p->nData = n_data;
    p->built = false;
    return CR_OK;
}

cr_status cr_beams_count(cr_context* ctx, uint32_t* n, uint32_t* n_data)
{
    CR_CHECK_CTX(ctx);
    if (n) *n = ctx->ppb ? ctx->ppb->n : 0;
    if (n_data) *n_data = ctx->ppb ? ctx->ppb->nData : 0;
    return CR_OK;
}

cr_status cr_beams_download(cr_context* ctx, cr_photon_beam* beams, cr_photon_beam_data* data)
{
    CR_CHECK_CTX(ctx);
    Ppb* p = ctx->ppb;
    if (!p) return fail(ctx, CR_ERR_STATE, "no beams");
    if (beams && p->n)
        CR_HIP(ctx, hipMemcpyAsync(beams, p->beams, (size_t)p->n * sizeof(cr_photon_beam), hipMemcpyDeviceToHost, ctx->stream), "copy");
    if (data && p->nData)
        CR_HIP(ctx, hipMemcpyAsync(data, p->data, (size_t)p->nData * sizeof(cr_photon_beam_data), hipMemcpyDeviceToHost, ctx->stream),
               "copy");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_ppb_trace_beams(cr_context* ctx, const cr_beam_trace_params* prm, uint32_t max_beams, uint32_t* out_beams,
                             uint32_t* out_data)
{
    CR_CHECK_CTX(ctx);
    if (!prm) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null params");
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (prm->numPaths == 0 || prm->numPaths > (1u << 26)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "numPaths must be in [1, 2^26]");
    if (!(prm->radius > 0.0f)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "radius must be positive");
    if (max_beams == 0 || max_beams > MAX_BEAMS) return fail(ctx, CR_ERR_CAPACITY, "max_beams must be in [1, 2^28]");
    hipSetDevice(ctx->device);
    Ppb* p;
    CR_TRY(ppb_get(ctx, &p));
    const uint32_t P = prm->numPaths;
    const DeviceGrid& g = ctx->grid;
    TraceArgs a;
    memset(&a, 0, sizeof(a));
    a.grid = g.data; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nbx = g.nbx; a.nby = g.nby;
    a.cloud = ctx->cloud;
    a.params = ctx->params;
    a.p = *prm;
    uint32_t* counts = nullptr;
    CR_HIP(ctx, hipMalloc((void**)&counts, (size_t)P * 8), "hipMalloc counts");
    a.segCounts = counts;
    a.dataCounts = counts + P;
    hipError_t e = launch_trace_count(a, g.dtype, g.layout, ctx->stream);
    std::vector<uint32_t> hc(2 * (size_t)P);
    if (e == hipSuccess) e = hipMemcpyAsync(hc.data(), counts, (size_t)P * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(counts);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace count");
    std::vector<uint64_t> off(2 * (size_t)P);
    uint64_t sb = 0, sd = 0;
    for (uint32_t i = 0; i < P; i++) {
        off[i] = sb;
        off[P + i] = sd;
        if (sb < max_beams) sd += hc[P + i];  // data of paths that keep at least one beam
        sb += hc[i];
    }
    const uint64_t nb = sb < max_beams ? sb : max_beams;
    const uint64_t nd = sd;
    if (nd > MAX_BEAMS) return fail(ctx, CR_ERR_CAPACITY, "beam data exceeds 2^28 records");
    CR_TRY(ensure_buffer(ctx, (void**)&p->beams, &p->beamsCap, (size_t)nb * sizeof(cr_photon_beam), "beams"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->data, &p->dataCap, (size_t)nd * sizeof(cr_photon_beam_data), "beam data"));
    uint64_t* doff = nullptr;
    CR_HIP(ctx, hipMalloc((void**)&doff, (size_t)P * 16), "hipMalloc offsets");
    e = hipMemcpyAsync(doff, off.data(), (size_t)P * 16, hipMemcpyHostToDevice, ctx->stream);
    a.segOffsets = doff;
    a.dataOffsets = doff + P;
    a.beams = p->beams;
    a.data = p->data;
    a.maxBeams = nb;
    a.maxData = nd;
    if (e == hipSuccess) e = launch_trace_write(a, g.dtype, g.layout, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(doff);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace write");
    p->n = (uint32_t)nb;
    p->nData = (uint32_t)nd;
    p->built = false;
    if (out_beams) *out_beams = (uint32_t)nb;
    if (out_data) *out_data = (uint32_t)nd;
    return CR_OK;
}

// RenderTechniquePPB::AllocateResources (RenderTechniquePPB.cpp:282-355): the
// beam, beam-data, sort, tree and traversal buffers for up to max_beams beams
// (and as many beam-data records), so that traces and builds of at most that
// many beams allocate nothing more.
cr_status cr_ppb_allocate(cr_context* ctx, uint32_t max_beams)
{
    CR_CHECK_CTX(ctx);
    if (max_beams < 2 || max_beams > MAX_BEAMS) return fail(ctx, CR_ERR_CAPACITY, "max_beams must be in [2, 2^28]");
    hipSetDevice(ctx->device);
    Ppb* p;
    CR_TRY(ppb_get(ctx, &p));
    CR_TRY(ensure_buffer(ctx, (void**)&p->beams, &p->beamsCap, (size_t)max_beams * sizeof(cr_photon_beam), "beams"));
    CR_TRY(ensure_buffer(ctx, (void**)&p->data, &p->dataCap, (size_t)max_beams * sizeof(cr_photon_beam_data), "beam data"));
    return lbvh_buffers(ctx, p, max_beams);
}

// RenderTechniquePPB::FreeResources: releases every beam and tree buffer (and
// the beams themselves); the next trace or upload allocates again.
cr_status cr_ppb_free(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ppb_free(ctx);
    return CR_OK;
}

cr_status cr_lbvh_build(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    Ppb* p = ctx->ppb;
    if (!p || p->n == 0) return fail(ctx, CR_ERR_STATE, "no beams");
    hipSetDevice(ctx->device);
    const uint32_t n = p->n;
    CR_TRY(lbvh_buffers(ctx, p, n));
    BoxArg box;
    for (int k = 0; k < 3; k++) {
        box.b0[k] = ctx->cloud.bounds[0][k];
        box.b1[k] = ctx->cloud.bounds[1][k];
    }
    hipStream_t s = ctx->stream;
    CR_HIP(ctx, zero_async(p->arrivals, (size_t)n * 4, s), "zero arrivals");
    CR_HIP(ctx, zero_async(p->exits + refit_exit_words(n) - 1, 4, s), "zero exit count");
    CR_HIP(ctx, hipEventRecord(p->ev[0], s), "event");
    CR_HIP(ctx, launch_morton(p->beams, n, box, p->keys, p->vals, s), "morton kernel");
    CR_HIP(ctx, hipEventRecord(p->ev[1], s), "event");
    CR_HIP(ctx, launch_radix_sort(p->keys, p->vals, p->keysTmp, p->valsTmp, p->counts, n, s), "radix sort");
    CR_HIP(ctx, hipEventRecord(p->ev[2], s), "event");
    CR_HIP(ctx, launch_hierarchy(p->keys, n, p->meta, p->leafParent, s), "hierarchy kernel");
    CR_HIP(ctx, hipEventRecord(p->ev[3], s), "event");
    CR_HIP(ctx, launch_leaves(p->beams, p->vals, n, p->sorted, p->sortedOff, p->lo, p->hi, s), "leaf kernel");
    CR_HIP(ctx, launch_refit(n, p->meta, p->leafParent, p->lo, p->hi, p->arrivals, p->exits, s), "refit kernels");
    CR_HIP(ctx, hipEventRecord(p->ev[4], s), "event");
    CR_HIP(ctx, launch_pack(n, p->meta, p->lo, p->hi, p->trav, s), "pack kernel");
    CR_HIP(ctx, hipEventRecord(p->ev[5], s), "event");
    if (n == 1) {
        uint32_t none = 0xFFFFFFFFu;
        CR_HIP(ctx, hipMemcpyAsync(p->leafParent, &none, 4, hipMemcpyHostToDevice, s), "root parent");
    }
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    p->built = true;
    p->builtN = n;
    return CR_OK;
}

cr_status cr_lbvh_timings(cr_context* ctx, double out[5])
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    Ppb* p = ctx->ppb;
    if (!p || !p->built) return fail(ctx, CR_ERR_STATE, "no build");
    for (int i = 0; i < 5; i++) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]);
        out[i] = ms;
    }
    return CR_OK;
}

cr_status cr_lbvh_download(cr_context* ctx, uint32_t* morton_sorted, uint32_t* perm, cr_tree_node* nodes)
{
    CR_CHECK_CTX(ctx);
    Ppb* p = ctx->ppb;
    if (!p || !p->built) return fail(ctx, CR_ERR_STATE, "no build");
    const uint32_t n = p->builtN;
    hipStream_t s = ctx->stream;
    if (morton_sorted) CR_HIP(ctx, hipMemcpyAsync(morton_sorted, p->keys, (size_t)n * 4, hipMemcpyDeviceToHost, s), "copy keys");
    if (perm) CR_HIP(ctx, hipMemcpyAsync(perm, p->vals, (size_t)n * 4, hipMemcpyDeviceToHost, s), "copy perm");
    if (nodes) {
        cr_tree_node* tmp = nullptr;
        const size_t bytes = (2 * (size_t)n - 1) * sizeof(cr_tree_node);
        CR_HIP(ctx, hipMalloc((void**)&tmp, bytes), "hipMalloc nodes");
        hipError_t e = launch_treenodes(n, p->meta, p->leafParent, p->lo, p->hi, tmp, s);
        if (e == hipSuccess) e = hipMemcpyAsync(nodes, tmp, bytes, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        hipFree(tmp);
        if (e != hipSuccess) return hip_fail(ctx, e, "tree nodes");
    }
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    return CR_OK;
}

cr_status cr_ppb_gather(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, const cr_gather_ext* ext)
{
    CR_CHECK_CTX(ctx);
    if (!frames && n_frames) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames");
    if (n_frames == 0) return CR_OK;
    Ppb* p = ctx->ppb;
    if (!p || !p->built) return fail(ctx, CR_ERR_STATE, "no LBVH built");
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    if (!ctx->svValid) return fail(ctx, CR_ERR_STATE, "gather needs a shadow volume (camera transmittance)");
    hipSetDevice(ctx->device);
    cr_gather_ext ge;
    if (ext) ge = *ext;
    else cr_gather_ext_default(&ge);
    CR_TRY(upload_frames(ctx, frames, n_frames));
    GatherArgs a;
    memset(&a, 0, sizeof(a));
    const DeviceGrid& g = ctx->grid;
    a.grid = g.data; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nbx = g.nbx; a.nby = g.nby;
    a.sv = ctx->sv;
    a.svx = (int32_t)ctx->svDimsBuilt[0]; a.svy = (int32_t)ctx->svDimsBuilt[1]; a.svz = (int32_t)ctx->svDimsBuilt[2];
    a.cam = ctx->cam;
    a.cloud = ctx->cloud;
    a.params = ctx->params;
    a.svp = ctx->svp;
    a.beams = p->sorted;
    a.beamOff = p->sortedOff;
    a.data = p->data;
    a.nodes = p->trav;
    a.leafLo = p->lo;
    a.leafHi = p->hi;
    a.n = p->builtN;
    a.refCompat = ge.refCompat;
    a.jitter = ge.jitter;
    a.albedo = ge.albedo;
    a.frames = ctx->dFrames;
    a.nFrames = n_frames;
    a.image = (float4*)ctx->image;
    a.W = ctx->W;
    a.H = ctx->H;
    a.padded = 32u * ((uint32_t)ctx->W / 32u + 1u);
    a.counters = p->gcounters;
    a.overflow = p->overflow;
    CR_HIP(ctx, hipEventRecord(ctx->gev0, ctx->stream), "event");
    CR_HIP(ctx, launch_gather(a, g.dtype, g.layout, ctx->stream), "gather kernel");
    CR_HIP(ctx, hipEventRecord(ctx->gev1, ctx->stream), "event");
    p->gatherPending = true;
    uint32_t ovf = 0;
    CR_HIP(ctx, hipMemcpyAsync(&ovf, p->overflow, 4, hipMemcpyDeviceToHost, ctx->stream), "copy overflow flag");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (ovf) {
        zero_async(p->overflow, 4, ctx->stream);
        if (ovf & 2u) return fail(ctx, CR_ERR_STATE, "malformed tree: traversal exceeded its node budget");
        return fail(ctx, CR_ERR_CAPACITY, "traversal stack (64 entries) overflowed; result incomplete");
    }
    return CR_OK;
}

cr_status cr_gather_stats_get(cr_context* ctx, cr_gather_stats* out)
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    Ppb* p;
    CR_TRY(ppb_get(ctx, &p));
    unsigned long long c[3];
    CR_HIP(ctx, hipMemcpyAsync(c, p->gcounters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream), "copy counters");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    out->rays = c[0];
    out->nodesVisited = c[1];
    out->beamsTested = c[2];
    if (p->gatherPending) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ctx->gev0, ctx->gev1);
        p->gatherMs = ms;
        p->gatherPending = false;
    }
    out->lastKernelMs = p->gatherMs;
    return CR_OK;
}

} // extern "C"
