// Batched rendering: many path-tracing jobs in one persistent launch, the
// launch recorded into a HIP graph, and the autotuner for its shape.
//
// A job is a complete RenderTechniquePT set-up (RenderTechniquePT.cpp:11-168:
// grid, the four uniform blocks, shadow volume, image, frame seeds).  Its
// device buffers are moved out of the context by cr_batch_add, so building a
// batch of 64 clouds never copies a grid.  The job table (one PTArgs per job,
// ordered by decreasing frame count so the longest pixels start first) lives
// in device memory and is copied to LDS by every workgroup of the launch; the
// graph bakes in only the launch shape and the table's address, so frame lists
// can change between replays without re-recording.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "cr_internal.h"
#include "cr_runtime.h"

using namespace cr;

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)

struct BatchJob {
    DeviceGrid grid;
    float* sv = nullptr;
    uint32_t svDims[3] = {0, 0, 0};
    cr_camera_properties cam;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_shadow_volume_properties svp;
    cr_render_ext ext;
    float* image = nullptr;
    int W = 0, H = 0;
    int32_t* dFrames = nullptr;
    size_t framesCap = 0;
    uint32_t nFrames = 0;
};

struct Batch {
    std::vector<BatchJob> jobs;
    cr_batch_tuning tuning;
    int dtype = 0, layout = 0;
    uint32_t tracking = 0, nee = 0;
    PTArgs* dTable = nullptr;       // render table
    PTArgs* dTuneTable = nullptr;   // autotuner trial table
    uint32_t* dQueue = nullptr;     // [0] render queue, [16] trial queue
    unsigned long long* dCounters = nullptr;   // [0..7] render, [8..15] trials
    float* dScratch = nullptr;      // trial images
    size_t scratchCap = 0;
    uint32_t total = 0;
    int blocks = 0;
    bool tableDirty = true, graphDirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t eventsUsed = 0;
    double totalMs = 0.0;
    float lastMs = 0.0f;
    uint64_t launches = 0;
};

namespace {

void free_job(BatchJob& j)
{
    void* bufs[] = {j.grid.data, j.grid.majorants, j.grid.skip, j.sv, j.image, j.dFrames};
    for (void* b : bufs)
        if (b) hipFree(b);
    j = BatchJob();
}

void drop_graph(Batch* b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graphDirty = true;
}

cr_status ensure_batch(cr_context* ctx)
{
    if (ctx->batch) return CR_OK;
    Batch* b = new Batch();
    cr_batch_tuning_default(&b->tuning);
    hipError_t e = hipMalloc((void**)&b->dTable, CR_BATCH_MAX_JOBS * sizeof(PTArgs));
    if (e == hipSuccess) e = hipMalloc((void**)&b->dTuneTable, CR_BATCH_MAX_JOBS * sizeof(PTArgs));
    if (e == hipSuccess) e = hipMalloc((void**)&b->dQueue, 32 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&b->dCounters, 16 * sizeof(unsigned long long));
    if (e == hipSuccess) e = zero_async(b->dCounters, 16 * sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->batch = b;
    if (e != hipSuccess) {
        batch_free(ctx);
        return hip_fail(ctx, e, "batch buffers");
    }
    return CR_OK;
}

// Job indices in table order: decreasing frame count, then insertion order.
std::vector<uint32_t> table_order(const Batch* b)
{
    std::vector<uint32_t> order(b->jobs.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return b->jobs[x].nFrames > b->jobs[y].nFrames; });
    return order;
}

uint32_t tile_log2(uint32_t tileWidth) { return tileWidth == 32 ? 5u : (tileWidth == 16 ? 4u : 3u); }

// Job table for `which` jobs (table order); images at `images[k]` if given
// (trial renders), each job's frame count divided by `frame_div` (at least 1
// frame; 0 = all frames).
std::vector<PTArgs> build_table(const Batch* b, const std::vector<uint32_t>& which, uint32_t tileWidth,
                                const std::vector<float*>* images, uint32_t frame_div, uint32_t* total)
{
    std::vector<PTArgs> t(which.size());
    uint32_t base = 0;
    for (size_t k = 0; k < which.size(); k++) {
        const BatchJob& j = b->jobs[which[k]];
        const uint32_t nf = frame_div ? std::max(1u, j.nFrames / frame_div) : j.nFrames;
        fill_job_args(t[k], j.grid, j.sv, j.svDims, j.cam, j.cloud, j.params, j.svp, j.ext, j.dFrames, nf,
                      images ? (*images)[k] : j.image, j.W, j.H, tile_log2(tileWidth));
        t[k].base = base;
        base += t[k].totalIdx;
    }
    *total = base;
    return t;
}

int launch_blocks(cr_context* ctx, const Batch* b, const cr_batch_tuning& t, uint32_t n_jobs, uint32_t total)
{
    int per_cu = (int)t.blocksPerCU;
    if (per_cu == 0)
        per_cu = pt_batch_blocks_per_cu(b->dtype, b->layout, b->tracking, b->nee, (int)t.blockThreads, n_jobs);
    int blocks = ctx->numCUs * per_cu;
    const uint32_t needed = (total + t.blockThreads - 1u) / t.blockThreads;
    if ((uint32_t)blocks > needed) blocks = (int)needed;
    return blocks > 0 ? blocks : 1;
}

cr_status upload_table(cr_context* ctx, PTArgs* dst, const std::vector<PTArgs>& t)
{
    CR_HIP(ctx, hipMemcpyAsync(dst, t.data(), t.size() * sizeof(PTArgs), hipMemcpyHostToDevice, ctx->stream),
           "copy job table");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync job table");
    return CR_OK;
}

// One trial: `which` jobs, each with 1/TRIAL_FRAME_DIV of its frames (so the
// trial has the real mix of long and short pixels), into the scratch images;
// returns the best of `reps` timed launches (after one untimed one).
constexpr uint32_t TRIAL_FRAME_DIV = 32;

cr_status trial(cr_context* ctx, Batch* b, const std::vector<uint32_t>& which, const cr_batch_tuning& t, int reps,
                float* best_ms)
{
    std::vector<float*> images(which.size());
    size_t off = 0;
    for (size_t k = 0; k < which.size(); k++) {
        images[k] = b->dScratch + off;
        off += (size_t)b->jobs[which[k]].W * b->jobs[which[k]].H * 4;
    }
    uint32_t total = 0;
    std::vector<PTArgs> table = build_table(b, which, t.tileWidth, &images, TRIAL_FRAME_DIV, &total);
    cr_status st = upload_table(ctx, b->dTuneTable, table);
    if (st != CR_OK) return st;
    BatchArgs a;
    a.jobs = b->dTuneTable;
    a.nJobs = (uint32_t)which.size();
    a.total = total;
    a.queue = b->dQueue + 16;
    a.counters = b->dCounters + 8;
    const int blocks = launch_blocks(ctx, b, t, a.nJobs, total);
    hipEvent_t e0, e1;
    CR_HIP(ctx, hipEventCreate(&e0), "hipEventCreate");
    hipError_t e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        hipEventDestroy(e0);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    float best = 1e30f;
    for (int r = 0; r <= reps && e == hipSuccess; r++) {
        e = zero_async(a.queue, 4, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
        if (e == hipSuccess)
            e = launch_pt_batch(a, b->dtype, b->layout, b->tracking, b->nee, blocks, (int)t.blockThreads, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && r > 0) best = std::min(best, ms);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(ctx, e, "autotune trial");
    *best_ms = best;
    return CR_OK;
}

cr_status fold_events(cr_context* ctx, Batch* b)
{
    if (b->eventsUsed == 0) return CR_OK;
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    for (size_t i = 0; i < b->eventsUsed; i++) {
        float ms = 0.0f;
        CR_HIP(ctx, hipEventElapsedTime(&ms, b->events[i].first, b->events[i].second), "hipEventElapsedTime");
        b->totalMs += ms;
        b->lastMs = ms;
        b->launches++;
    }
    b->eventsUsed = 0;
    return CR_OK;
}

bool valid_tuning(const cr_batch_tuning& t)
{
    return (t.blockThreads == 64 || t.blockThreads == 128 || t.blockThreads == 256) && t.blocksPerCU <= 32 &&
           (t.tileWidth == 8 || t.tileWidth == 16 || t.tileWidth == 32);
}

} // namespace

namespace cr {
void batch_free(cr_context* ctx)
{
    Batch* b = ctx->batch;
    if (!b) return;
    drop_graph(b);
    for (BatchJob& j : b->jobs) free_job(j);
    void* bufs[] = {b->dTable, b->dTuneTable, b->dQueue, b->dCounters, b->dScratch};
    for (void* p : bufs)
        if (p) hipFree(p);
    for (auto& ev : b->events) {
        hipEventDestroy(ev.first);
        hipEventDestroy(ev.second);
    }
    delete b;
    ctx->batch = nullptr;
}
} // namespace cr

extern "C" {

void cr_batch_tuning_default(cr_batch_tuning* t)
{
    if (!t) return;
    t->blockThreads = 256;
    t->blocksPerCU = 0;
    t->tileWidth = 8;
    t->_reserved = 0;
}

cr_status cr_batch_add(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, uint32_t* job)
{
    CR_CHECK_CTX(ctx);
    if (!frames || n_frames == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "a job needs at least one frame");
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    if (ctx->ext.neeMode == CR_NEE_SHADOW_VOLUME && !ctx->svValid)
        return fail(ctx, CR_ERR_STATE, "shadow-volume NEE needs cr_sv_build after the last grid/shadow-volume change");
    cr_status st = ensure_batch(ctx);
    if (st != CR_OK) return st;
    Batch* b = ctx->batch;
    if (b->jobs.size() >= CR_BATCH_MAX_JOBS) return fail(ctx, CR_ERR_CAPACITY, "batch holds CR_BATCH_MAX_JOBS jobs");
    if (!b->jobs.empty() && (ctx->grid.dtype != b->dtype || ctx->grid.layout != b->layout ||
                             ctx->ext.trackingMode != b->tracking || ctx->ext.neeMode != b->nee))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT,
                    "jobs of a batch share grid storage type, grid layout, tracking mode and NEE mode");
    hipSetDevice(ctx->device);
    BatchJob j;
    j.W = ctx->W;
    j.H = ctx->H;
    const size_t px = (size_t)j.W * j.H;
    CR_HIP(ctx, hipMalloc((void**)&j.image, px * 16), "hipMalloc job image");
    std::vector<int32_t> fr(2 * (size_t)n_frames);
    for (uint32_t i = 0; i < n_frames; i++) {
        fr[2 * i] = frames[i].seed;
        fr[2 * i + 1] = (int32_t)frames[i].frameCount;
    }
    hipError_t e = hipMalloc((void**)&j.dFrames, fr.size() * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(j.dFrames, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = zero_async(j.image, px * 16, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_job(j);
        return hip_fail(ctx, e, "job buffers");
    }
    j.framesCap = fr.size() * 4;
    j.nFrames = n_frames;
    j.cam = ctx->cam;
    j.cloud = ctx->cloud;
    j.params = ctx->params;
    j.svp = ctx->svp;
    j.ext = ctx->ext;
    // move the grid and the shadow volume out of the context
    j.grid = ctx->grid;
    ctx->grid = DeviceGrid();
    ctx->hasGrid = false;
    ctx->sceneVersion++;
    if (ctx->svValid) {
        j.sv = ctx->sv;
        memcpy(j.svDims, ctx->svDimsBuilt, sizeof(j.svDims));
        ctx->sv = nullptr;
        ctx->svCap = 0;
    }
    ctx->svValid = false;
    if (b->jobs.empty()) {
        b->dtype = j.grid.dtype;
        b->layout = j.grid.layout;
        b->tracking = j.ext.trackingMode;
        b->nee = j.ext.neeMode;
    }
    b->jobs.push_back(j);
    b->tableDirty = true;
    drop_graph(b);
    if (job) *job = (uint32_t)(b->jobs.size() - 1);
    return CR_OK;
}

cr_status cr_batch_clear(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    if (!ctx->batch) return CR_OK;
    hipSetDevice(ctx->device);
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    Batch* b = ctx->batch;
    drop_graph(b);
    for (BatchJob& j : b->jobs) free_job(j);
    b->jobs.clear();
    b->tableDirty = true;
    return CR_OK;
}

cr_status cr_batch_size(cr_context* ctx, uint32_t* n)
{
    CR_CHECK_CTX(ctx);
    if (!n) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    *n = ctx->batch ? (uint32_t)ctx->batch->jobs.size() : 0u;
    return CR_OK;
}

cr_status cr_batch_set_frames(cr_context* ctx, uint32_t job, const cr_push_constants* frames, uint32_t n_frames)
{
    CR_CHECK_CTX(ctx);
    Batch* b = ctx->batch;
    if (!b || job >= b->jobs.size()) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "no such job");
    if (!frames || n_frames == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "a job needs at least one frame");
    BatchJob& j = b->jobs[job];
    std::vector<int32_t> fr(2 * (size_t)n_frames);
    for (uint32_t i = 0; i < n_frames; i++) {
        fr[2 * i] = frames[i].seed;
        fr[2 * i + 1] = (int32_t)frames[i].frameCount;
    }
    hipSetDevice(ctx->device);
    // the stream may still read the old list
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (fr.size() * 4 > j.framesCap) {
        int32_t* p = nullptr;
        CR_HIP(ctx, hipMalloc((void**)&p, fr.size() * 4), "hipMalloc frames");
        hipFree(j.dFrames);
        j.dFrames = p;
        j.framesCap = fr.size() * 4;
    }
    CR_HIP(ctx, hipMemcpyAsync(j.dFrames, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, ctx->stream), "copy frames");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync frames");
    // the graph reads the table from memory: rewriting it before the next render suffices
    j.nFrames = n_frames;
    b->tableDirty = true;
    return CR_OK;
}

cr_status cr_batch_clear_images(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    if (!ctx->batch) return CR_OK;
    for (BatchJob& j : ctx->batch->jobs)
        CR_HIP(ctx, zero_async(j.image, (size_t)j.W * j.H * 16, ctx->stream), "clear job image");
    return CR_OK;
}

cr_status cr_batch_set_tuning(cr_context* ctx, const cr_batch_tuning* t)
{
    CR_CHECK_CTX(ctx);
    if (!t || !valid_tuning(*t)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "invalid batch tuning");
    cr_status st = ensure_batch(ctx);
    if (st != CR_OK) return st;
    Batch* b = ctx->batch;
    b->tuning = *t;
    b->tuning._reserved = 0;
    b->tableDirty = true;
    drop_graph(b);
    return CR_OK;
}

cr_status cr_batch_get_tuning(cr_context* ctx, cr_batch_tuning* t)
{
    CR_CHECK_CTX(ctx);
    if (!t) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    if (ctx->batch) *t = ctx->batch->tuning;
    else cr_batch_tuning_default(t);
    return CR_OK;
}

cr_status cr_batch_render(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    Batch* b = ctx->batch;
    if (!b || b->jobs.empty()) return fail(ctx, CR_ERR_STATE, "empty batch");
    hipSetDevice(ctx->device);
    if (b->tableDirty) {
        // the graph may still be reading the table
        CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
        const std::vector<uint32_t> order = table_order(b);
        uint32_t total = 0;
        std::vector<PTArgs> table = build_table(b, order, b->tuning.tileWidth, nullptr, 0, &total);
        cr_status st = upload_table(ctx, b->dTable, table);
        if (st != CR_OK) return st;
        if (total != b->total) drop_graph(b);
        b->total = total;
        b->tableDirty = false;
    }
    if (b->graphDirty) {
        drop_graph(b);
        BatchArgs a;
        a.jobs = b->dTable;
        a.nJobs = (uint32_t)b->jobs.size();
        a.total = b->total;
        a.queue = b->dQueue;
        a.counters = b->dCounters;
        b->blocks = launch_blocks(ctx, b, b->tuning, a.nJobs, a.total);
        CR_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "begin capture");
        hipError_t e = zero_async(b->dQueue, 4, ctx->stream);
        if (e == hipSuccess)
            e = launch_pt_batch(a, b->dtype, b->layout, b->tracking, b->nee, b->blocks, (int)b->tuning.blockThreads,
                                ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e == hipSuccess) e = e2;
        if (e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return hip_fail(ctx, e, "capture batch launch");
        }
        b->graph = g;
        CR_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate batch graph");
        b->graphDirty = false;
    }
    if (b->eventsUsed >= 1024) {
        cr_status st = fold_events(ctx, b);
        if (st != CR_OK) return st;
    }
    if (b->eventsUsed == b->events.size()) {
        std::pair<hipEvent_t, hipEvent_t> ev;
        CR_HIP(ctx, hipEventCreate(&ev.first), "hipEventCreate");
        CR_HIP(ctx, hipEventCreate(&ev.second), "hipEventCreate");
        b->events.push_back(ev);
    }
    const auto& ev = b->events[b->eventsUsed];
    CR_HIP(ctx, hipEventRecord(ev.first, ctx->stream), "event");
    CR_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream), "launch batch graph");
    CR_HIP(ctx, hipEventRecord(ev.second, ctx->stream), "event");
    b->eventsUsed++;
    return CR_OK;
}

cr_status cr_batch_download(cr_context* ctx, uint32_t job, float* rgba, uint32_t* width, uint32_t* height)
{
    CR_CHECK_CTX(ctx);
    Batch* b = ctx->batch;
    if (!b || job >= b->jobs.size()) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "no such job");
    const BatchJob& j = b->jobs[job];
    if (width) *width = (uint32_t)j.W;
    if (height) *height = (uint32_t)j.H;
    if (rgba) {
        CR_HIP(ctx, hipMemcpyAsync(rgba, j.image, (size_t)j.W * j.H * 16, hipMemcpyDeviceToHost, ctx->stream),
               "copy job image");
        CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    }
    return CR_OK;
}

cr_status cr_batch_stats_get(cr_context* ctx, cr_pt_stats* out)
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    memset(out, 0, sizeof(*out));
    Batch* b = ctx->batch;
    if (!b) return CR_OK;
    unsigned long long c[8];
    CR_HIP(ctx, hipMemcpyAsync(c, b->dCounters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream), "copy counters");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    out->paths = c[0];
    out->vertices = c[1];
    out->lookups = c[2];
    out->svLookups = c[3];
    out->brickVisits = c[4];
    cr_status st = fold_events(ctx, b);
    if (st != CR_OK) return st;
    out->lastKernelMs = b->lastMs;
    out->totalKernelMs = b->totalMs;
    out->launches = b->launches;
    return CR_OK;
}

cr_status cr_batch_stats_reset(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    Batch* b = ctx->batch;
    if (!b) return CR_OK;
    cr_status st = fold_events(ctx, b);
    if (st != CR_OK) return st;
    CR_HIP(ctx, zero_async(b->dCounters, 8 * sizeof(unsigned long long), ctx->stream), "zero counters");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    b->totalMs = 0.0;
    b->lastMs = 0.0f;
    b->launches = 0;
    return CR_OK;
}

cr_status cr_batch_autotune(cr_context* ctx, int tune_bricks, cr_batch_tuning* out, cr_batch_tune_report* report)
{
    CR_CHECK_CTX(ctx);
    Batch* b = ctx->batch;
    if (!b || b->jobs.empty()) return fail(ctx, CR_ERR_STATE, "empty batch");
    hipSetDevice(ctx->device);
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    size_t px = 0;
    for (const BatchJob& j : b->jobs) px += (size_t)j.W * j.H;
    cr_status st = ensure_buffer(ctx, (void**)&b->dScratch, &b->scratchCap, px * 16, "autotune scratch images");
    if (st != CR_OK) return st;
    cr_batch_tune_report rep;
    memset(&rep, 0, sizeof(rep));
    cr_batch_tuning def;
    cr_batch_tuning_default(&def);

    // (1) majorant brick per job, each job alone with the default launch shape
    if (tune_bricks && b->tracking == CR_TRACK_BRICK_DDA) {
        const int sizes[3] = {8, 16, 32};
        for (uint32_t k = 0; k < b->jobs.size(); k++) {
            BatchJob& j = b->jobs[k];
            int best = majorant_brick(j.ext);
            float bestMs = 1e30f;
            for (int mb : sizes) {
                st = majorants_build(ctx, j.grid, mb);
                if (st != CR_OK) return st;
                j.ext.majorantBrick = (uint32_t)mb;
                float ms = 0.0f;
                st = trial(ctx, b, std::vector<uint32_t>{k}, def, 1, &ms);
                if (st != CR_OK) return st;
                if (ms < bestMs) {
                    bestMs = ms;
                    best = mb;
                }
            }
            st = majorants_build(ctx, j.grid, best);
            if (st != CR_OK) return st;
            j.ext.majorantBrick = (uint32_t)best;
        }
        b->tableDirty = true;
    }
    for (uint32_t k = 0; k < b->jobs.size(); k++) rep.jobBrick[k] = (uint32_t)majorant_brick(b->jobs[k].ext);

    // (2) launch shape over the whole batch
    const std::vector<uint32_t> order = table_order(b);
    st = trial(ctx, b, order, def, 1, &rep.defaultMs);
    if (st != CR_OK) return st;
    cr_batch_tuning best = def;
    float bestMs = rep.defaultMs;
    const uint32_t threads[3] = {64, 128, 256};
    const uint32_t tiles[3] = {8, 16, 32};
    for (uint32_t th : threads) {
        const int occ = pt_batch_blocks_per_cu(b->dtype, b->layout, b->tracking, b->nee, (int)th,
                                               (uint32_t)b->jobs.size());
        // below half the occupancy limit a persistent grid cannot hide its latency
        for (int per = std::max(1, occ / 2); per <= occ; per++) {
            for (uint32_t tw : tiles) {
                cr_batch_tuning t;
                t.blockThreads = th;
                t.blocksPerCU = (uint32_t)per;
                t.tileWidth = tw;
                t._reserved = 0;
                float ms = 0.0f;
                st = trial(ctx, b, order, t, 1, &ms);
                if (st != CR_OK) return st;
                rep.candidates++;
                if (ms < bestMs) {
                    bestMs = ms;
                    best = t;
                }
            }
        }
    }
    rep.bestMs = bestMs;
    b->tuning = best;
    b->tableDirty = true;
    drop_graph(b);
    if (out) *out = best;
    if (report) *report = rep;
    return CR_OK;
}

} // extern "C"
