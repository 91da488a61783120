// Allocation-free path-tracer entry point (cr_pt_render_static).
//
// cr_pt_render keeps its scratch in context buffers that it grows on demand:
// the work-queue counter, the statistics counters, the per-frame
// (seed, frameCount) table and the per-tile cost and claim-order buffers.
// This variant keeps the first three in one fixed device array per GPU,
// resolved once per device when a context is created, and renders into an
// image the caller owns; tiles are claimed in raster order, which changes the
// schedule only (every pixel's arithmetic is cr_pt_render's).  It calls no
// hipMalloc, hipFree or allocating library routine.
//
// The array is shared by every context on a GPU and keeps the last call's
// contents: the queue and counters are cleared by a fill kernel at the start
// of each call, the frame table is overwritten.  Calls that overlap on one
// GPU (two host threads, or two contexts whose streams run concurrently)
// race on it.
#include <mutex>

#include "cr_runtime.h"

using namespace cr;

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)

namespace {

// Layout: [0, 64) work-queue counter, [64, 192) 16 statistics counters,
// [256, 256 + 8 CR_PT_STATIC_MAX_FRAMES) the (seed, frameCount) pairs.
constexpr size_t QUEUE_OFF = 0;
constexpr size_t COUNTERS_OFF = 64;
constexpr size_t FRAMES_OFF = 256;
static_assert(FRAMES_OFF + 8ull * CR_PT_STATIC_MAX_FRAMES == CR_PT_STATIC_SCRATCH_BYTES, "scratch layout");

static __device__ __align__(256) unsigned char scratch[CR_PT_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
unsigned char* g_scratch[MAX_DEVICES] = {};   // this GPU's copy of `scratch`, by device ordinal

unsigned char* scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    return (device >= 0 && device < MAX_DEVICES) ? g_scratch[device] : nullptr;
}

} // namespace

namespace cr {

// Resolves the device address of this GPU's `scratch` (each GPU holds its own
// copy at its own address).  Called by cr_create with the context's device
// current; later contexts on the same device reuse the address.
cr_status static_scratch_init(cr_context* ctx)
{
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
        return fail(ctx, CR_ERR_NOT_SUPPORTED, "device ordinal beyond the static-scratch table");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[ctx->device]) return CR_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
    g_scratch[ctx->device] = (unsigned char*)p;
    return CR_OK;
}

} // namespace cr

extern "C" {

cr_status cr_pt_render_static(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, float* accum)
{
    CR_CHECK_CTX(ctx);
    if (!frames && n_frames) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames");
    if (!accum) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null accumulation image");
    if (n_frames > CR_PT_STATIC_MAX_FRAMES)
        return fail(ctx, CR_ERR_CAPACITY, "cr_pt_render_static renders at most CR_PT_STATIC_MAX_FRAMES (" +
                                              std::to_string(CR_PT_STATIC_MAX_FRAMES) + ") frames per call, got " +
                                              std::to_string(n_frames));
    if (n_frames == 0) return CR_OK;
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (ctx->W <= 0 || ctx->H <= 0) return fail(ctx, CR_ERR_STATE, "no camera set");
    if (ctx->ext.neeMode == CR_NEE_SHADOW_VOLUME && !ctx->svValid)
        return fail(ctx, CR_ERR_STATE, "shadow-volume NEE needs cr_sv_build after the last grid/shadow-volume change");
    unsigned char* base = scratch_of(ctx->device);
    if (!base) return fail(ctx, CR_ERR_STATE, "static scratch not resolved for this device");
    CR_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");

    int32_t fr[2 * CR_PT_STATIC_MAX_FRAMES];
    for (uint32_t i = 0; i < n_frames; i++) {
        fr[2 * i] = frames[i].seed;
        fr[2 * i + 1] = (int32_t)frames[i].frameCount;
    }
    int32_t* dFrames = (int32_t*)(base + FRAMES_OFF);
    // stream order: the copy waits for any earlier kernel that still reads the table
    CR_HIP(ctx, hipMemcpyAsync(dFrames, fr, (size_t)n_frames * 8, hipMemcpyHostToDevice, ctx->stream), "copy frames");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync frames");
    CR_HIP(ctx, zero_async(base + QUEUE_OFF, FRAMES_OFF - QUEUE_OFF, ctx->stream), "clear queue and counters");

    PTArgs a;
    fill_job_args(a, ctx->grid, ctx->sv, ctx->svDimsBuilt, ctx->cam, ctx->cloud, ctx->params, ctx->svp, ctx->ext,
                  dFrames, n_frames, accum, ctx->W, ctx->H, ctx->ptTileLog2W);
    a.queue = (uint32_t*)(base + QUEUE_OFF);
    a.counters = (unsigned long long*)(base + COUNTERS_OFF);
    a.tileOrder = nullptr;
    a.tileCost = nullptr;
    const DeviceGrid& g = ctx->grid;
    int blocks = ctx->numCUs * pt_blocks_per_cu(g.dtype, g.layout, ctx->ext.trackingMode, ctx->ext.neeMode, n_frames);
    const uint32_t needed = (a.totalIdx + 255u) / 256u;
    if ((uint32_t)blocks > needed) blocks = (int)needed;
    CR_HIP(ctx, launch_pt(a, g.dtype, g.layout, blocks, ctx->stream), "path tracer kernel");
    return CR_OK;
}

} // extern "C"
