// Host runtime of the learned renderer's support data (extension): the density
// mip pyramid and stencil features.
#include <hip/hip_runtime.h>

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
    const void* builtFrom = nullptr;
    float* points = nullptr;
    size_t pointsCap = 0;
    uint16_t* X = nullptr;
    size_t XCap = 0;
};

namespace cr {

void learn_free(cr_context* ctx)
{
    Learn* L = ctx->learn;
    if (!L) return;
    for (int k = 0; k < LEARN_LEVELS; k++)
        if (L->mips[k]) hipFree(L->mips[k]);
    if (L->points) hipFree(L->points);
    if (L->X) hipFree(L->X);
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
    L->builtFrom = g.data;
    return CR_OK;
}

// Features of device-resident points into device-resident X.
cr_status learn_features_device(cr_context* ctx, const float* points, uint64_t n, uint16_t* X)
{
    Learn* L = learn_get(ctx);
    if (!L->valid) return fail(ctx, CR_ERR_STATE, "mip pyramid not built (cr_learn_build_mips)");
    FeatureArgs a;
    memset(&a, 0, sizeof(a));
    a.points = points;
    a.n = n;
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
