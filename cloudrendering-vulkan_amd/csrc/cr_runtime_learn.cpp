// Host runtime of the learned renderer's support data (extension): the density
// mip pyramid and stencil features.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "cr_internal_learn.h"
#include "cr_runtime.h"

using namespace cr;

struct Learn {
    float* mips[LEARN_LEVELS] = {};
    size_t mipCap[LEARN_LEVELS] = {};
    int32_t dims[LEARN_LEVELS][3] = {};
    bool valid = false;
    uint64_t builtVersion = 0;
    float* points = nullptr;
    size_t pointsCap = 0;
    uint16_t* X = nullptr;
    size_t XCap = 0;
    // record pass, compaction, network outputs
    float* records = nullptr;
    size_t recordsCap = 0;
    uint32_t* slot = nullptr;
    size_t slotCap = 0;
    uint32_t* counts = nullptr;
    size_t countsCap = 0;
    uint32_t* mdev = nullptr;   // collision count of the last device-side compaction
    size_t mdevCap = 0;
    float* Y = nullptr;
    size_t YCap = 0;
    float* targets = nullptr;
    size_t targetsCap = 0;
    // training data set; step s of cr_learn_train draws its rows with step number trainSteps + s
    uint16_t* dsX = nullptr;
    float* dsT = nullptr;
    uint64_t dsN = 0, dsCap = 0;
    uint32_t trainSteps = 0;
    // whether the data set's targets leave out the first collision's
    // shadow-volume term (-1: empty data set), and the same for the data the
    // network was last trained on (-1: not trained from a data set)
    int dsFirstNee = -1;
    int netFirstNee = -1;
};

namespace cr {

// The first collision's next-event term that the path tracer adds with the
// current settings (on only for shadow-volume NEE and a bounce cap >= 1).
FirstNee first_nee_of(const cr_context* ctx)
{
    FirstNee e;
    memset(&e, 0, sizeof(e));
    e.on = (ctx->ext.neeMode == CR_NEE_SHADOW_VOLUME && ctx->ext.bounceCap >= 1u) ? 1u : 0u;
    e.sv = ctx->sv;
    e.svx = (int32_t)ctx->svDimsBuilt[0];
    e.svy = (int32_t)ctx->svDimsBuilt[1];
    e.svz = (int32_t)ctx->svDimsBuilt[2];
    e.svp = ctx->svp;
    e.params = ctx->params;
    e.intensity = ctx->params.lightIntensity;
    e.albedo = ctx->ext.albedo;
    return e;
}

void learn_network_reset(cr_context* ctx)
{
    if (ctx->learn) ctx->learn->netFirstNee = -1;
}

void learn_free(cr_context* ctx)
{
    Learn* L = ctx->learn;
    if (!L) return;
    for (int k = 0; k < LEARN_LEVELS; k++)
        if (L->mips[k]) hipFree(L->mips[k]);
    void* bufs[] = {L->points, L->X, L->records, L->slot, L->counts, L->mdev, L->Y, L->targets, L->dsX, L->dsT};
    for (void* b : bufs)
        if (b) hipFree(b);
    delete L;
    ctx->learn = nullptr;
}

static Learn* learn_get(cr_context* ctx)
{
    if (!ctx->learn) ctx->learn = new Learn();
    return ctx->learn;
}

cr_status learn_mips(cr_context* ctx)
{
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    Learn* L = learn_get(ctx);
    const DeviceGrid& g = ctx->grid;
    int32_t d[3] = {g.nx, g.ny, g.nz};
    for (int k = 0; k < LEARN_LEVELS; k++) {
        L->dims[k][0] = d[0]; L->dims[k][1] = d[1]; L->dims[k][2] = d[2];
        cr_status st = ensure_buffer(ctx, (void**)&L->mips[k], &L->mipCap[k], (size_t)d[0] * d[1] * d[2] * 4, "mip");
        if (st != CR_OK) return st;
        for (int a = 0; a < 3; a++) d[a] = (d[a] + 1) / 2;
    }
    hipError_t e = launch_grid_to_f32(g.data, g.dtype, g.layout, L->mips[0], g.nx, g.ny, g.nz, g.nbx, g.nby, ctx->stream);
    for (int k = 1; k < LEARN_LEVELS && e == hipSuccess; k++)
        e = launch_mip(L->mips[k - 1], L->dims[k - 1][0], L->dims[k - 1][1], L->dims[k - 1][2], L->mips[k], L->dims[k][0],
                       L->dims[k][1], L->dims[k][2], ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "mip pyramid");
    L->valid = true;
    L->builtVersion = ctx->gridVersion;
    return CR_OK;
}

// Features of device-resident points into device-resident X; ndev: optional
// device-resident point count (<= n, n sizes the launch).
cr_status learn_features_device(cr_context* ctx, const float* points, uint64_t n, uint16_t* X,
                                const uint32_t* ndev = nullptr)
{
    Learn* L = learn_get(ctx);
    if (!L->valid || L->builtVersion != ctx->gridVersion)
        return fail(ctx, CR_ERR_STATE, "mip pyramid not built for the current grid (cr_learn_build_mips)");
    FeatureArgs a;
    memset(&a, 0, sizeof(a));
    a.points = points;
    a.n = n;
    a.ndev = ndev;
    for (int k = 0; k < LEARN_LEVELS; k++) {
        a.mips[k] = L->mips[k];
        a.dims[k][0] = L->dims[k][0]; a.dims[k][1] = L->dims[k][1]; a.dims[k][2] = L->dims[k][2];
    }
    a.cloud = ctx->cloud;
    a.light[0] = ctx->svp.lightDirection[0];
    a.light[1] = ctx->svp.lightDirection[1];
    a.light[2] = ctx->svp.lightDirection[2];
    a.h0 = (ctx->cloud.bounds[1][0] - ctx->cloud.bounds[0][0]) / (float)ctx->grid.nx;
    a.X = X;
    hipError_t e = launch_features(a, ctx->stream);
    return e == hipSuccess ? CR_OK : hip_fail(ctx, e, "features");
}

} // namespace cr

extern "C" {

cr_status cr_learn_build_mips(cr_context* ctx)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    hipSetDevice(ctx->device);
    return learn_mips(ctx);
}

cr_status cr_learn_features(cr_context* ctx, const float* points, uint32_t n, uint16_t* X)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (!points || !X) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null points or output");
    if (n == 0) return CR_OK;
    hipSetDevice(ctx->device);
    Learn* L = learn_get(ctx);
    cr_status st = ensure_buffer(ctx, (void**)&L->points, &L->pointsCap, (size_t)n * 32, "points");
    if (st == CR_OK) st = ensure_buffer(ctx, (void**)&L->X, &L->XCap, (size_t)n * 448, "features");
    if (st != CR_OK) return st;
    hipError_t e = hipMemcpyAsync(L->points, points, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "copy points");
    st = learn_features_device(ctx, L->points, n, L->X);
    if (st != CR_OK) return st;
    e = hipMemcpyAsync(X, L->X, (size_t)n * 448, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? CR_OK : hip_fail(ctx, e, "copy features");
}

} // extern "C"

// ------------------------------------------------------------------------
// Learned renderer: record pass, continuation pass, compaction, training data,
// training loop and the learned render.

#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)
#define CR_TRY(call) \
    do { cr_status _s = (call); if (_s != CR_OK) return _s; } while (0)

namespace {

cr_status check_render_state(cr_context* ctx)
{
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    if (ctx->ext.neeMode == CR_NEE_SHADOW_VOLUME && !ctx->svValid)
        return fail(ctx, CR_ERR_STATE, "shadow-volume NEE needs cr_sv_build after the last grid/shadow-volume change");
    return CR_OK;
}

int pass_blocks(cr_context* ctx, int continuation, uint32_t work)
{
    const int per_cu = pt_learn_blocks_per_cu(ctx->grid.dtype, ctx->grid.layout, ctx->ext.trackingMode, ctx->ext.neeMode,
                                              continuation);
    int blocks = ctx->numCUs * per_cu;
    const uint32_t needed = (work + 255u) / 256u;
    if ((uint32_t)blocks > needed) blocks = (int)needed;
    return blocks > 0 ? blocks : 1;
}

// Record pass of n_frames frames into L->records; returns the record count.
cr_status record_pass(cr_context* ctx, Learn* L, const cr_push_constants* frames, uint32_t n_frames, uint64_t* n_out)
{
    CR_TRY(check_render_state(ctx));
    CR_TRY(upload_frames(ctx, frames, n_frames));
    const uint64_t n = (uint64_t)n_frames * (uint64_t)ctx->W * (uint64_t)ctx->H;
    CR_TRY(ensure_buffer(ctx, (void**)&L->records, &L->recordsCap, n * 32, "records"));
    PTArgs a;
    CR_TRY(fill_pt_args(ctx, a, n_frames));
    a.records = L->records;
    a.counters = ctx->dCounters + 8;
    CR_HIP(ctx, zero_async(ctx->dQueue, 4, ctx->stream), "zero queue");
    CR_HIP(ctx, launch_pt_record(a, ctx->grid.dtype, ctx->grid.layout, pass_blocks(ctx, 0, a.totalIdx), ctx->stream),
           "record pass");
    *n_out = n;
    return CR_OK;
}

// Continuation pass over m device records into out (float4 per record).
cr_status continue_pass(cr_context* ctx, const float* pts, uint32_t m, uint32_t spp, float* out)
{
    if (m == 0) return CR_OK;
    PTArgs a;
    CR_TRY(fill_pt_args(ctx, a, 0));
    a.points = pts;
    a.out = reinterpret_cast<float4*>(out);
    a.spp = spp;
    a.totalIdx = m;
    a.counters = ctx->dCounters + 8;
    CR_HIP(ctx, zero_async(ctx->dQueue, 4, ctx->stream), "zero queue");
    CR_HIP(ctx, launch_pt_continue(a, ctx->grid.dtype, ctx->grid.layout, pass_blocks(ctx, 1, m), ctx->stream),
           "continuation pass");
    return CR_OK;
}

// Collided records of L->records (n) into L->points in record order, with
// L->slot[i] = row or ~0; returns the count m (synchronises the stream).
cr_status compact(cr_context* ctx, Learn* L, uint64_t n, uint32_t* m_out)
{
    const uint64_t nb = (n + 255) / 256;
    CR_TRY(ensure_buffer(ctx, (void**)&L->counts, &L->countsCap, nb * 4, "counts"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->slot, &L->slotCap, n * 4, "slots"));
    CR_HIP(ctx, launch_count(L->records, n, L->counts, ctx->stream), "count collisions");
    std::vector<uint32_t> c(nb);
    CR_HIP(ctx, hipMemcpyAsync(c.data(), L->counts, nb * 4, hipMemcpyDeviceToHost, ctx->stream), "copy counts");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    uint64_t total = 0;
    for (uint64_t b = 0; b < nb; b++) {
        const uint32_t v = c[b];
        c[b] = (uint32_t)total;
        total += v;
    }
    if (total > 0xFFFFFFFEull) return fail(ctx, CR_ERR_CAPACITY, "more than 2^32-2 collisions in one pass");
    CR_HIP(ctx, hipMemcpyAsync(L->counts, c.data(), nb * 4, hipMemcpyHostToDevice, ctx->stream), "copy offsets");
    CR_TRY(ensure_buffer(ctx, (void**)&L->points, &L->pointsCap, (total ? total : 1) * 32, "points"));
    CR_HIP(ctx, launch_compact(L->records, n, L->counts, L->points, L->slot, ctx->stream), "compact collisions");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");  // c is read by the copy above
    *m_out = (uint32_t)total;
    return CR_OK;
}

// The same compaction with the scan on the device: no host round trip; the
// count stays in L->mdev and L->points / L->slot are sized for all n records.
cr_status compact_device(cr_context* ctx, Learn* L, uint64_t n)
{
    const uint64_t nb = (n + 255) / 256;
    if (nb > 0xFFFFFFFFull) return fail(ctx, CR_ERR_CAPACITY, "too many records in one pass");
    CR_TRY(ensure_buffer(ctx, (void**)&L->counts, &L->countsCap, nb * 4, "counts"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->slot, &L->slotCap, n * 4, "slots"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->mdev, &L->mdevCap, 4, "collision count"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->points, &L->pointsCap, n * 32, "points"));
    CR_HIP(ctx, launch_count(L->records, n, L->counts, ctx->stream), "count collisions");
    CR_HIP(ctx, launch_scan_counts(L->counts, (uint32_t)nb, L->mdev, ctx->stream), "scan counts");
    CR_HIP(ctx, launch_compact(L->records, n, L->counts, L->points, L->slot, ctx->stream), "compact collisions");
    return CR_OK;
}

cr_status ensure_mips(cr_context* ctx, Learn* L)
{
    if (L->valid && L->builtVersion == ctx->gridVersion) return CR_OK;
    return learn_mips(ctx);
}

cr_status ensure_dataset(cr_context* ctx, Learn* L, uint64_t need)
{
    if (need <= L->dsCap) return CR_OK;
    uint64_t cap = std::max<uint64_t>(need, L->dsCap * 2);
    cap = std::max<uint64_t>(cap, 1u << 16);
    uint16_t* X = nullptr;
    float* T = nullptr;
    CR_HIP(ctx, hipMalloc((void**)&X, cap * 448), "hipMalloc data set");
    hipError_t e = hipMalloc((void**)&T, cap * 12);
    if (e != hipSuccess) {
        hipFree(X);
        return hip_fail(ctx, e, "hipMalloc data set");
    }
    if (L->dsN) {
        e = hipMemcpyAsync(X, L->dsX, L->dsN * 448, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(T, L->dsT, L->dsN * 12, hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            hipFree(X);
            hipFree(T);
            return hip_fail(ctx, e, "grow data set");
        }
    }
    if (L->dsX) hipFree(L->dsX);
    if (L->dsT) hipFree(L->dsT);
    L->dsX = X;
    L->dsT = T;
    L->dsCap = cap;
    return CR_OK;
}

} // namespace

extern "C" {

cr_status cr_pt_record(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, float* records)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (!frames || n_frames == 0 || !records) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames or output");
    hipSetDevice(ctx->device);
    Learn* L = learn_get(ctx);
    uint64_t n = 0;
    CR_TRY(record_pass(ctx, L, frames, n_frames, &n));
    CR_HIP(ctx, hipMemcpyAsync(records, L->records, n * 32, hipMemcpyDeviceToHost, ctx->stream), "copy records");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_pt_continue(cr_context* ctx, const float* records, uint32_t n, uint32_t spp, float* out)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (!records || !out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null records or output");
    if (spp == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "spp must be >= 1");
    if (n == 0) return CR_OK;
    CR_TRY(check_render_state(ctx));
    hipSetDevice(ctx->device);
    Learn* L = learn_get(ctx);
    CR_TRY(ensure_buffer(ctx, (void**)&L->points, &L->pointsCap, (size_t)n * 32, "points"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->targets, &L->targetsCap, (size_t)n * 16, "targets"));
    CR_HIP(ctx, hipMemcpyAsync(L->points, records, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream), "copy records");
    CR_TRY(continue_pass(ctx, L->points, n, spp, L->targets));
    CR_HIP(ctx, hipMemcpyAsync(out, L->targets, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream), "copy output");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_learn_collect(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, uint32_t spp,
                           uint32_t max_points, uint32_t* added)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (!frames || n_frames == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames");
    if (spp == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "spp must be >= 1");
    if (added) *added = 0;
    hipSetDevice(ctx->device);
    Learn* L = learn_get(ctx);
    CR_TRY(check_render_state(ctx));
    CR_TRY(ensure_mips(ctx, L));
    const FirstNee nee = first_nee_of(ctx);
    if (L->dsN > 0 && L->dsFirstNee != (int)nee.on)
        return fail(ctx, CR_ERR_STATE, "the data set holds targets of the other NEE mode (shadow volume vs not); clear it first");
    uint64_t n = 0;
    uint32_t m = 0;
    CR_TRY(record_pass(ctx, L, frames, n_frames, &n));
    CR_TRY(compact(ctx, L, n, &m));
    if (m == 0) return CR_OK;
    const uint32_t stride = (max_points == 0 || m <= max_points) ? 1u : (m + max_points - 1u) / max_points;
    const uint32_t count = (m + stride - 1u) / stride;
    // the selected collisions, contiguous (the record buffer is free again)
    float* sel = L->points;
    if (stride > 1) {
        CR_HIP(ctx, launch_select(L->points, stride, count, L->records, ctx->stream), "select");
        sel = L->records;
    }
    CR_TRY(ensure_buffer(ctx, (void**)&L->targets, &L->targetsCap, (size_t)count * 16, "targets"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->X, &L->XCap, (size_t)count * 448, "features"));
    CR_TRY(continue_pass(ctx, sel, count, spp, L->targets));
    CR_TRY(learn_features_device(ctx, sel, count, L->X));
    CR_TRY(ensure_dataset(ctx, L, L->dsN + count));
    CR_HIP(ctx, launch_append(L->X, reinterpret_cast<const float4*>(L->targets), sel, nee, 1, count, L->dsX, L->dsT,
                              L->dsN, ctx->stream), "append");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    L->dsN += count;
    L->dsFirstNee = (int)nee.on;
    if (added) *added = count;
    return CR_OK;
}

cr_status cr_learn_dataset_size(cr_context* ctx, uint64_t* n)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (!n) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    *n = ctx->learn ? ctx->learn->dsN : 0;
    return CR_OK;
}

cr_status cr_learn_dataset_clear(cr_context* ctx)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (ctx->learn) {
        ctx->learn->dsN = 0;
        ctx->learn->dsFirstNee = -1;
    }
    return CR_OK;
}

cr_status cr_learn_dataset_download(cr_context* ctx, uint64_t first, uint32_t n, uint16_t* X, float* T)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    Learn* L = learn_get(ctx);
    if (first + n > L->dsN) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "rows beyond the data set");
    if (n == 0) return CR_OK;
    if (X) CR_HIP(ctx, hipMemcpyAsync(X, L->dsX + first * 224, (size_t)n * 448, hipMemcpyDeviceToHost, ctx->stream), "copy X");
    if (T) CR_HIP(ctx, hipMemcpyAsync(T, L->dsT + first * 3, (size_t)n * 12, hipMemcpyDeviceToHost, ctx->stream), "copy T");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_learn_train(cr_context* ctx, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* loss)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (batch == 0 || batch > (1u << 22)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "batch must be in [1, 2^22]");
    if (!(lr > 0.0f)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "lr must be > 0");
    hipSetDevice(ctx->device);
    Learn* L = learn_get(ctx);
    if (L->dsN == 0) return fail(ctx, CR_ERR_STATE, "empty training data set (cr_learn_collect)");
    if (!mlp_initialized(ctx)) CR_TRY(cr_mlp_init(ctx, seed));
    for (uint32_t s = 0; s < steps; s++) {
        const MlpDraw draw = {L->dsX, L->dsT, L->dsN, seed, L->trainSteps + s};
        CR_TRY(mlp_train_device(ctx, nullptr, nullptr, batch, lr, &draw));
    }
    L->trainSteps += steps;
    L->netFirstNee = L->dsFirstNee;
    if (loss) CR_TRY(mlp_last_loss(ctx, loss));
    return CR_OK;
}

cr_status cr_learn_render(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames)
{
    if (!ctx) return CR_ERR_INVALID_ARGUMENT;
    if (!frames && n_frames) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames");
    if (n_frames == 0) return CR_OK;
    if (!mlp_initialized(ctx)) return fail(ctx, CR_ERR_STATE, "no network (cr_mlp_init / cr_learn_train)");
    hipSetDevice(ctx->device);
    Learn* L = learn_get(ctx);
    CR_TRY(check_render_state(ctx));
    const FirstNee nee = first_nee_of(ctx);
    if (L->netFirstNee >= 0 && L->netFirstNee != (int)nee.on)
        return fail(ctx, CR_ERR_STATE, "the network was trained for the other NEE mode (shadow volume vs not)");
    CR_TRY(ensure_mips(ctx, L));
    // record -> compaction -> features -> network -> composition, all on the
    // stream: the collision count stays on the device and bounds the feature
    // and network launches, which are sized for every record.
    uint64_t n = 0;
    CR_TRY(record_pass(ctx, L, frames, n_frames, &n));
    CR_TRY(compact_device(ctx, L, n));
    CR_TRY(ensure_buffer(ctx, (void**)&L->X, &L->XCap, (size_t)n * 448, "features"));
    CR_TRY(ensure_buffer(ctx, (void**)&L->Y, &L->YCap, (size_t)n * 16, "network outputs"));
    CR_TRY(learn_features_device(ctx, L->points, n, L->X, L->mdev));
    CR_TRY(mlp_forward_device(ctx, L->X, n, L->Y, L->mdev));
    ComposeArgs a;
    a.image = reinterpret_cast<float4*>(ctx->image);
    a.pixels = (uint32_t)(ctx->W * ctx->H);
    a.nFrames = n_frames;
    a.frames = ctx->dFrames;
    a.records = L->records;
    a.slot = L->slot;
    a.Y = L->Y;
    a.nee = nee;
    CR_HIP(ctx, launch_compose(a, ctx->stream), "compose");
    return CR_OK;
}

} // extern "C"
