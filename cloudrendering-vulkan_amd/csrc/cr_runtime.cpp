// Host runtime behind the C-ABI: the context (device, stream, resident
// buffers), the density grid, the shadow-volume technique and the
// path-tracing technique.  Errors are returned as status codes with a message
// on the context; nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cr_internal.h"
#include "cr_runtime.h"

using namespace cr;

namespace cr {

cr_status fail(cr_context* ctx, cr_status code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

cr_status hip_fail(cr_context* ctx, hipError_t e, const char* what)
{
    return fail(ctx, CR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

cr_status ensure_buffer(cr_context* ctx, void** p, size_t* cap, size_t bytes, const char* what)
{
    if (*cap >= bytes && *p) return CR_OK;
    if (*p) {
        hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    if (bytes == 0) return CR_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, CR_ERR_OUT_OF_MEMORY, std::string("hipMalloc ") + what + " (" + std::to_string(bytes) + " B): " +
                                                   hipGetErrorString(e));
    }
    *cap = bytes;
    return CR_OK;
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Clears with a kernel rather than hipMemsetAsync so the same code can be
// captured into a HIP graph.
hipError_t zero_async(void* p, size_t bytes, hipStream_t s)
{
    uint32_t n = (uint32_t)((bytes + 3) / 4);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, (uint32_t*)p, n);
    return hipGetLastError();
}

} // namespace cr

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)

extern "C" {

const char* cr_status_string(cr_status s)
{
    switch (s) {
    case CR_OK: return "CR_OK";
    case CR_ERR_INVALID_ARGUMENT: return "CR_ERR_INVALID_ARGUMENT";
    case CR_ERR_OUT_OF_MEMORY: return "CR_ERR_OUT_OF_MEMORY";
    case CR_ERR_DEVICE: return "CR_ERR_DEVICE";
    case CR_ERR_CAPACITY: return "CR_ERR_CAPACITY";
    case CR_ERR_IO: return "CR_ERR_IO";
    case CR_ERR_STATE: return "CR_ERR_STATE";
    case CR_ERR_NOT_SUPPORTED: return "CR_ERR_NOT_SUPPORTED";
    default: return "CR_ERR_UNKNOWN";
    }
}

const char* cr_last_error(const cr_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

const char* cr_version(void) { return "cr-mi355x 0.1 gfx950"; }

cr_status cr_create(int device, cr_context** out)
{
    if (!out) return CR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return CR_ERR_DEVICE;
    if (device < 0 || device >= count) return CR_ERR_INVALID_ARGUMENT;
    cr_context* ctx = new cr_context();
    ctx->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess) { delete ctx; return CR_ERR_DEVICE; }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) { delete ctx; return CR_ERR_DEVICE; }
    ctx->numCUs = prop.multiProcessorCount;
    ctx->arch = prop.gcnArchName;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) { delete ctx; return CR_ERR_DEVICE; }
    hipEventCreate(&ctx->ev0);
    hipEventCreate(&ctx->ev1);
    hipEventCreate(&ctx->gev0);
    hipEventCreate(&ctx->gev1);
    cr_camera_default(&ctx->cam);
    cr_cloud_default(&ctx->cloud);
    cr_parameters_default(&ctx->params);
    cr_shadow_volume_default(&ctx->svp);
    cr_render_ext_default(&ctx->ext);
    cr_photon_map_default(&ctx->pm);
    if (ensure_buffer(ctx, (void**)&ctx->dQueue, &ctx->queueCap, 64, "queue") != CR_OK ||
        ensure_buffer(ctx, (void**)&ctx->dCounters, &ctx->countersCap, 16 * sizeof(unsigned long long), "counters") != CR_OK ||
        ensure_buffer(ctx, (void**)&ctx->dScalar, &ctx->scalarCap, 64, "scalar") != CR_OK) {
        cr_destroy(ctx);
        return CR_ERR_OUT_OF_MEMORY;
    }
    if (static_scratch_init(ctx) != CR_OK) {
        cr_destroy(ctx);
        return CR_ERR_DEVICE;
    }
    zero_async(ctx->dCounters, 16 * sizeof(unsigned long long), ctx->stream);
    hipStreamSynchronize(ctx->stream);
    *out = ctx;
    return CR_OK;
}

cr_status cr_destroy(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    void* bufs[] = {ctx->grid.data, ctx->grid.majorants, ctx->grid.skip, ctx->sv, ctx->image, ctx->dFrames, ctx->dQueue,
                    ctx->tileCost, ctx->tileSort,
                    ctx->dCounters, ctx->dScalar};
    for (void* b : bufs)
        if (b) hipFree(b);
    ppb_free(ctx);
    ppm_free(ctx);
    mlp_free(ctx);
    learn_free(ctx);
    batch_free(ctx);
    for (auto& ev : ctx->ptEvents) {
        hipEventDestroy(ev.first);
        hipEventDestroy(ev.second);
    }
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->gev0) hipEventDestroy(ctx->gev0);
    if (ctx->gev1) hipEventDestroy(ctx->gev1);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return CR_OK;
}

cr_status cr_synchronize(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return CR_OK;
}

void* cr_stream(cr_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ------------------------------------------------------------------ grid

cr_status cr_set_grid_layout(cr_context* ctx, int layout)
{
    CR_CHECK_CTX(ctx);
    if (layout != CR_GRID_DENSE && layout != CR_GRID_BRICK4)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "layout must be CR_GRID_DENSE or CR_GRID_BRICK4");
    ctx->gridLayout = layout;
    return CR_OK;
}

} // extern "C"

namespace cr {

// Allocate the device grid for (nx,ny,nz) in storage dtype and the context layout.
cr_status grid_alloc(cr_context* ctx, int32_t nx, int32_t ny, int32_t nz, int dtype)
{
    DeviceGrid& g = ctx->grid;
    int layout = ctx->gridLayout;
    int32_t nbx = (nx + 3) / 4, nby = (ny + 3) / 4, nbz = (nz + 3) / 4;
    size_t elems = layout == CR_GRID_BRICK4 ? (size_t)nbx * nby * nbz * 64 : (size_t)nx * ny * nz;
    if (elems >= (1ull << 32)) return fail(ctx, CR_ERR_CAPACITY, "grids hold fewer than 2^32 voxels (32-bit lookups)");
    size_t bytes = elems * (dtype == CR_DTYPE_F16 ? 2 : 4);
    if (g.data && g.bytes != bytes) {
        hipFree(g.data);
        g.data = nullptr;
        g.bytes = 0;
    }
    if (!g.data) {
        hipError_t e = hipMalloc(&g.data, bytes);
        if (e != hipSuccess) {
            g.data = nullptr;
            return fail(ctx, CR_ERR_OUT_OF_MEMORY, std::string("hipMalloc grid: ") + hipGetErrorString(e));
        }
        g.bytes = bytes;
    }
    if (layout == CR_GRID_BRICK4) {
        hipError_t e = zero_async(g.data, bytes, ctx->stream); // padding voxels of partial bricks
        if (e != hipSuccess) return hip_fail(ctx, e, "zero grid");
    }
    g.dtype = dtype;
    g.layout = layout;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nbx = nbx; g.nby = nby; g.nbz = nbz;
    return CR_OK;
}

int majorant_brick(const cr_render_ext& e) { return e.majorantBrick ? (int)e.majorantBrick : 16; }

// (Re)build the brick-majorant grid of g for mb^3 bricks.
cr_status majorants_build(cr_context* ctx, DeviceGrid& g, int mb)
{
    int32_t mx = (g.nx + mb - 1) / mb, my = (g.ny + mb - 1) / mb, mz = (g.nz + mb - 1) / mb;
    const size_t n = (size_t)mx * my * mz;
    // the brick walk addresses the two-float skip table with 32-bit byte offsets
    if (n >= ((size_t)1 << 29)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "more than 2^29 majorant bricks");
    if (g.majorants && (g.mx != mx || g.my != my || g.mz != mz)) {
        hipFree(g.majorants);
        hipFree(g.skip);
        g.majorants = nullptr;
        g.skip = nullptr;
    }
    if (!g.majorants) {
        hipError_t e = hipMalloc((void**)&g.majorants, n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void**)&g.skip, 2 * n * sizeof(float));
        if (e != hipSuccess) {
            if (g.majorants) hipFree(g.majorants);
            g.majorants = nullptr;
            g.skip = nullptr;
            return fail(ctx, CR_ERR_OUT_OF_MEMORY, "hipMalloc majorants");
        }
    }
    g.mx = mx; g.my = my; g.mz = mz; g.mb = mb;
    CR_HIP(ctx, launch_majorants(g.data, g.dtype, g.layout, g.nx, g.ny, g.nz, g.nbx, g.nby, mb, g.majorants, mx, my, mz,
                                 ctx->stream), "majorant kernel");
    uint8_t* tmp = nullptr;
    CR_HIP(ctx, hipMallocAsync((void**)&tmp, 2 * n, ctx->stream), "skip scratch");
    hipError_t e = launch_skip_grid(g.majorants, mx, my, mz, ctx->cloud.densityScaling, ctx->cloud.baseScaling, g.skip, tmp,
                                    ctx->stream);
    hipError_t e2 = hipFreeAsync(tmp, ctx->stream);
    CR_HIP(ctx, e, "skip grid");
    CR_HIP(ctx, e2, "free skip scratch");
    return CR_OK;
}

// Brick majorants, global maximum and SetCloudProperties (main.cpp:144-163).
cr_status grid_finish(cr_context* ctx, double vx, double vy, double vz)
{
    DeviceGrid& g = ctx->grid;
    cr_status st = majorants_build(ctx, g, majorant_brick(ctx->ext));
    if (st != CR_OK) return st;
    CR_HIP(ctx, zero_async(ctx->dScalar, 4, ctx->stream), "zero");
    CR_HIP(ctx, launch_grid_max(g.majorants, (uint32_t)((size_t)g.mx * g.my * g.mz), (unsigned int*)ctx->dScalar, ctx->stream),
           "max kernel");
    unsigned int bits = 0;
    CR_HIP(ctx, hipMemcpyAsync(&bits, ctx->dScalar, 4, hipMemcpyDeviceToHost, ctx->stream), "copy max");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    float gmax;
    memcpy(&gmax, &bits, 4);
    ctx->gridMax = gmax;
    ctx->voxelSize[0] = vx; ctx->voxelSize[1] = vy; ctx->voxelSize[2] = vz;
    cr_cloud_properties& c = ctx->cloud;
    // glm::vec3 cloudSize{ voxelSize * voxelCount * baseScaling } (double products, float storage)
    float sx = (float)(vx * (double)(uint32_t)g.nx * (double)c.baseScaling);
    float sy = (float)(vy * (double)(uint32_t)g.ny * (double)c.baseScaling);
    float sz = (float)(vz * (double)(uint32_t)g.nz * (double)c.baseScaling);
    c.maxExtinction = std::max(gmax, 0.01f);
    c.voxelCount[0] = (uint32_t)g.nx; c.voxelCount[1] = (uint32_t)g.ny; c.voxelCount[2] = (uint32_t)g.nz; c.voxelCount[3] = 0;
    c.bounds[0][0] = -sx / 2.0f; c.bounds[0][1] = -sy / 2.0f; c.bounds[0][2] = 0.0f; c.bounds[0][3] = 0.0f;
    c.bounds[1][0] = -c.bounds[0][0] + 0.0f;
    c.bounds[1][1] = -c.bounds[0][1] + 0.0f;
    c.bounds[1][2] = -c.bounds[0][2] + sz;
    c.bounds[1][3] = -c.bounds[0][3] + 0.0f;
    ctx->hasGrid = true;
    ctx->gridVersion++;
    ctx->sceneVersion++;
    ctx->svValid = false;
    return CR_OK;
}

} // namespace cr

extern "C" {

cr_status cr_grid_upload(cr_context* ctx, const cr_grid_header* hdr, const void* data, int src_dtype, int layout,
                         int storage_dtype)
{
    CR_CHECK_CTX(ctx);
    if (!hdr || !data) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null header or data");
    if (hdr->sizeX == 0 || hdr->sizeY == 0 || hdr->sizeZ == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "zero-sized grid");
    if (hdr->sizeX > 65535 * 256u || hdr->sizeY > 65535u || hdr->sizeZ > 65535u)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "grid too large");
    if ((src_dtype != CR_DTYPE_F32 && src_dtype != CR_DTYPE_F16) || (storage_dtype != CR_DTYPE_F32 && storage_dtype != CR_DTYPE_F16))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "dtype must be CR_DTYPE_F32 or CR_DTYPE_F16");
    if (layout != CR_LAYOUT_X_FASTEST && layout != CR_LAYOUT_Z_FASTEST) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "bad layout");
    if (!(hdr->voxelSizeX > 0.0) || !(hdr->voxelSizeY > 0.0) || !(hdr->voxelSizeZ > 0.0))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "voxel sizes must be positive");
    hipSetDevice(ctx->device);
    size_t n = (size_t)hdr->sizeX * hdr->sizeY * hdr->sizeZ;
    size_t srcBytes = n * (src_dtype == CR_DTYPE_F16 ? 2 : 4);
    void* staging = nullptr;
    CR_HIP(ctx, hipMalloc(&staging, srcBytes), "hipMalloc staging");
    hipError_t e = hipMemcpyAsync(staging, data, srcBytes, hipMemcpyHostToDevice, ctx->stream);
    cr_status st = CR_OK;
    if (e != hipSuccess) st = hip_fail(ctx, e, "upload");
    if (st == CR_OK) st = grid_alloc(ctx, (int32_t)hdr->sizeX, (int32_t)hdr->sizeY, (int32_t)hdr->sizeZ, storage_dtype);
    if (st == CR_OK) {
        DeviceGrid& g = ctx->grid;
        e = launch_convert_grid(staging, src_dtype, layout == CR_LAYOUT_Z_FASTEST, g.data, g.dtype, g.layout, g.nx, g.ny, g.nz,
                                g.nbx, g.nby, ctx->stream);
        if (e != hipSuccess) st = hip_fail(ctx, e, "convert kernel");
    }
    hipStreamSynchronize(ctx->stream);
    hipFree(staging);
    if (st != CR_OK) return st;
    return grid_finish(ctx, hdr->voxelSizeX, hdr->voxelSizeY, hdr->voxelSizeZ);
}

cr_status cr_grid_generate_fbm(cr_context* ctx, const cr_fbm_params* p, int storage_dtype)
{
    CR_CHECK_CTX(ctx);
    if (!p) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null params");
    if (p->size[0] == 0 || p->size[1] == 0 || p->size[2] == 0 || p->size[1] > 65535u || p->size[2] > 65535u)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "bad grid size");
    if (!(p->voxelSize > 0.0f) || p->octaves == 0 || p->octaves > 16)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "bad voxel size or octave count");
    if (storage_dtype != CR_DTYPE_F32 && storage_dtype != CR_DTYPE_F16) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "bad dtype");
    hipSetDevice(ctx->device);
    cr_status st = grid_alloc(ctx, (int32_t)p->size[0], (int32_t)p->size[1], (int32_t)p->size[2], storage_dtype);
    if (st != CR_OK) return st;
    DeviceGrid& g = ctx->grid;
    CR_HIP(ctx, launch_fbm(*p, g.data, g.dtype, g.layout, g.nbx, g.nby, ctx->stream), "fbm kernel");
    return grid_finish(ctx, (double)p->voxelSize, (double)p->voxelSize, (double)p->voxelSize);
}

cr_status cr_grid_download(cr_context* ctx, float* out)
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    hipSetDevice(ctx->device);
    DeviceGrid& g = ctx->grid;
    size_t n = (size_t)g.nx * g.ny * g.nz;
    float* tmp = nullptr;
    CR_HIP(ctx, hipMalloc((void**)&tmp, n * 4), "hipMalloc");
    hipError_t e = launch_grid_to_f32(g.data, g.dtype, g.layout, tmp, g.nx, g.ny, g.nz, g.nbx, g.nby, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipFree(tmp);
    if (e != hipSuccess) return hip_fail(ctx, e, "grid download");
    if (e2 != hipSuccess) return hip_fail(ctx, e2, "sync");
    return CR_OK;
}

cr_status cr_grid_majorants(cr_context* ctx, uint32_t dims[3], float* out)
{
    CR_CHECK_CTX(ctx);
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (dims) { dims[0] = (uint32_t)ctx->grid.mx; dims[1] = (uint32_t)ctx->grid.my; dims[2] = (uint32_t)ctx->grid.mz; }
    if (out) {
        size_t n = (size_t)ctx->grid.mx * ctx->grid.my * ctx->grid.mz;
        CR_HIP(ctx, hipMemcpyAsync(out, ctx->grid.majorants, n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy majorants");
        CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    }
    return CR_OK;
}

cr_status cr_grid_skip(cr_context* ctx, float* out)
{
    CR_CHECK_CTX(ctx);
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    size_t n = (size_t)ctx->grid.mx * ctx->grid.my * ctx->grid.mz;
    CR_HIP(ctx, hipMemcpy2DAsync(out, 4, ctx->grid.skip, 8, 4, n, hipMemcpyDeviceToHost, ctx->stream), "copy skip grid");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_math_probe(cr_context* ctx, int op, const float* in, uint32_t n, float* out)
{
    CR_CHECK_CTX(ctx);
    if (op < CR_MATH_SQRT || op > CR_MATH_BF16) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "unknown math op");
    if (n == 0) return CR_OK;
    if (!in || !out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null array");
    float* d = nullptr;
    hipError_t e = hipMalloc(&d, (size_t)n * 8);
    if (e != hipSuccess) return hip_fail(ctx, e, "probe buffer");
    e = hipMemcpyAsync(d, in, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_math_probe(op, d, n, d + n, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + n, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    return e == hipSuccess ? CR_OK : hip_fail(ctx, e, "math probe");
}

// Grid3D<float>::Load (Grid3D.h:44-90): 3 x u32 counts, 3 x f64 voxel sizes,
// then f32 voxels with z fastest.  A missing file returns CR_ERR_IO (the
// reference returns nullptr and keeps the old grid, main.cpp:171-176); a
// short file is rejected instead of reading garbage.
cr_status cr_grid_load_xyz(cr_context* ctx, const char* path, int storage_dtype)
{
    CR_CHECK_CTX(ctx);
    if (!path) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null path");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(ctx, CR_ERR_IO, std::string("cannot open ") + path);
    uint32_t sz[3];
    double vs[3];
    if (fread(sz, 4, 3, f) != 3 || fread(vs, 8, 3, f) != 3) {
        fclose(f);
        return fail(ctx, CR_ERR_IO, "short .xyz header");
    }
    size_t n = (size_t)sz[0] * sz[1] * sz[2];
    if (n == 0) {
        fclose(f);
        return fail(ctx, CR_ERR_IO, "zero-sized .xyz grid");
    }
    std::vector<float> buf;
    try {
        buf.resize(n);
    } catch (...) {
        fclose(f);
        return fail(ctx, CR_ERR_OUT_OF_MEMORY, "host buffer");
    }
    size_t got = fread(buf.data(), 4, n, f);
    fclose(f);
    if (got != n) return fail(ctx, CR_ERR_IO, "short .xyz data");
    cr_grid_header h;
    h.sizeX = sz[0]; h.sizeY = sz[1]; h.sizeZ = sz[2]; h._pad = 0;
    h.voxelSizeX = vs[0]; h.voxelSizeY = vs[1]; h.voxelSizeZ = vs[2];
    return cr_grid_upload(ctx, &h, buf.data(), CR_DTYPE_F32, CR_LAYOUT_Z_FASTEST, storage_dtype);
}

// Writes file order (z fastest) so that load(save(g)) == g; the reference's
// Grid3D::Save (Grid3D.h:168-178) writes memory order (x fastest), which its
// Load reads as z fastest, so there load(save(g)) != g for a non-cubic or
// non-symmetric grid.
cr_status cr_grid_save_xyz(cr_context* ctx, const char* path)
{
    CR_CHECK_CTX(ctx);
    if (!path) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null path");
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    DeviceGrid& g = ctx->grid;
    size_t n = (size_t)g.nx * g.ny * g.nz;
    std::vector<float> mem(n), file(n);
    cr_status st = cr_grid_download(ctx, mem.data());
    if (st != CR_OK) return st;
    for (size_t x = 0; x < (size_t)g.nx; x++)
        for (size_t y = 0; y < (size_t)g.ny; y++)
            for (size_t z = 0; z < (size_t)g.nz; z++)
                file[z + g.nz * (y + g.ny * x)] = mem[x + g.nx * (y + g.ny * z)];
    FILE* f = fopen(path, "wb");
    if (!f) return fail(ctx, CR_ERR_IO, std::string("cannot create ") + path);
    uint32_t sz[3] = {(uint32_t)g.nx, (uint32_t)g.ny, (uint32_t)g.nz};
    bool ok = fwrite(sz, 4, 3, f) == 3 && fwrite(ctx->voxelSize, 8, 3, f) == 3 && fwrite(file.data(), 4, n, f) == n;
    ok = (fclose(f) == 0) && ok;
    return ok ? CR_OK : fail(ctx, CR_ERR_IO, "write failed");
}

// --------------------------------------------------------------- uniforms

cr_status cr_set_camera(cr_context* ctx, const cr_camera_properties* cam)
{
    CR_CHECK_CTX(ctx);
    if (!cam) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null camera");
    if (cam->halfWidth <= 0 || cam->halfHeight <= 0 || cam->halfWidth > 8192 || cam->halfHeight > 8192)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "camera resolution out of range");
    for (int i = 0; i < 3; i++)
        if (!std::isfinite(cam->position[i]) || !std::isfinite(cam->forward[i]) || !std::isfinite(cam->right[i]) ||
            !std::isfinite(cam->up[i]))
            return fail(ctx, CR_ERR_INVALID_ARGUMENT, "non-finite camera vector");
    ctx->cam = *cam;
    ctx->sceneVersion++;
    int W = 2 * cam->halfWidth, H = 2 * cam->halfHeight;
    if (W != ctx->W || H != ctx->H || !ctx->image) {
        cr_status st = ensure_buffer(ctx, (void**)&ctx->image, &ctx->imageCap, (size_t)W * H * 16, "image");
        if (st != CR_OK) return st;
        ctx->W = W;
        ctx->H = H;
        CR_HIP(ctx, zero_async(ctx->image, (size_t)W * H * 16, ctx->stream), "clear image");
    }
    return CR_OK;
}

cr_status cr_set_cloud(cr_context* ctx, const cr_cloud_properties* c)
{
    CR_CHECK_CTX(ctx);
    if (!c) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null cloud properties");
    if (!(c->baseScaling > 0.0f) || !(c->maxExtinction > 0.0f) || !(c->densityScaling >= 0.0f) ||
        !(c->densityScaling < __builtin_inff()))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT,
                    "baseScaling and maxExtinction must be positive, densityScaling finite and non-negative");
    const bool rescale = ctx->hasGrid && (c->densityScaling != ctx->cloud.densityScaling ||
                                          c->baseScaling != ctx->cloud.baseScaling);
    ctx->cloud = *c;
    ctx->sceneVersion++;
    ctx->svValid = false;
    if (rescale) {
        const DeviceGrid& g = ctx->grid;
        CR_HIP(ctx, launch_skip_scale(g.majorants, (uint32_t)((size_t)g.mx * g.my * g.mz), c->densityScaling, c->baseScaling,
                                      g.skip, ctx->stream), "skip grid scale");
    }
    return CR_OK;
}

cr_status cr_get_cloud(cr_context* ctx, cr_cloud_properties* c)
{
    CR_CHECK_CTX(ctx);
    if (!c) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    *c = ctx->cloud;
    return CR_OK;
}

cr_status cr_set_parameters(cr_context* ctx, const cr_parameters* p)
{
    CR_CHECK_CTX(ctx);
    if (!p) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null parameters");
    ctx->params = *p;
    ctx->sceneVersion++;
    return CR_OK;
}

cr_status cr_set_shadow_volume(cr_context* ctx, const cr_shadow_volume_properties* s)
{
    CR_CHECK_CTX(ctx);
    if (!s) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null shadow volume properties");
    if (s->voxelAxisCount < 2 || s->voxelAxisCount > 2048)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "voxelAxisCount must be in [2, 2048]");
    ctx->svp = *s;
    ctx->sceneVersion++;
    ctx->svValid = false;
    return CR_OK;
}

cr_status cr_set_render_ext(cr_context* ctx, const cr_render_ext* e)
{
    CR_CHECK_CTX(ctx);
    if (!e) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null extension block");
    if (e->skyMode > CR_SKY_CONSTANT || e->neeMode > CR_NEE_NONE || e->trackingMode > CR_TRACK_BRICK_DDA)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "unknown sky, NEE or tracking mode");
    if (!(e->albedo >= 0.0f && e->albedo <= 1.0f)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "albedo must be in [0,1]");
    bool anyDim = e->svDims[0] || e->svDims[1] || e->svDims[2];
    bool allDim = e->svDims[0] && e->svDims[1] && e->svDims[2];
    if (anyDim && (!allDim || e->svDims[0] < 2 || e->svDims[1] < 2 || e->svDims[2] < 2 || e->svDims[0] > 4096 ||
                   e->svDims[1] > 4096 || e->svDims[2] > 4096))
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "svDims must be all zero or all in [2, 4096]");
    if (e->majorantBrick != 0 && e->majorantBrick != 4 && e->majorantBrick != 8 && e->majorantBrick != 16 &&
        e->majorantBrick != 32)
        return fail(ctx, CR_ERR_INVALID_ARGUMENT, "majorantBrick must be 0, 4, 8, 16 or 32");
    if (memcmp(e->svDims, ctx->ext.svDims, sizeof(e->svDims)) != 0) ctx->svValid = false;
    const int mb = majorant_brick(*e);
    ctx->ext = *e;
    ctx->sceneVersion++;
    if (ctx->hasGrid && ctx->grid.mb != mb) return majorants_build(ctx, ctx->grid, mb);
    return CR_OK;
}

// ----------------------------------------------------------- shadow volume

static void sv_dims_of(const cr_context* ctx, uint32_t d[3])
{
    if (ctx->ext.svDims[0] && ctx->ext.svDims[1] && ctx->ext.svDims[2]) {
        d[0] = ctx->ext.svDims[0]; d[1] = ctx->ext.svDims[1]; d[2] = ctx->ext.svDims[2];
    } else {
        d[0] = d[1] = d[2] = ctx->svp.voxelAxisCount;
    }
}

cr_status cr_sv_dims(cr_context* ctx, uint32_t out[3])
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    sv_dims_of(ctx, out);
    return CR_OK;
}

cr_status cr_sv_build(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    hipSetDevice(ctx->device);
    uint32_t d[3];
    sv_dims_of(ctx, d);
    if ((uint64_t)d[0] * d[1] * d[2] >= (1ull << 32))
        return fail(ctx, CR_ERR_CAPACITY, "shadow volumes hold fewer than 2^32 texels (32-bit lookups)");
    size_t bytes = (size_t)d[0] * d[1] * d[2] * 4;
    cr_status st = ensure_buffer(ctx, (void**)&ctx->sv, &ctx->svCap, bytes, "shadow volume");
    if (st != CR_OK) return st;
    SVArgs a;
    const DeviceGrid& g = ctx->grid;
    a.grid = g.data; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nbx = g.nbx; a.nby = g.nby;
    a.cloud = ctx->cloud;
    a.svp = ctx->svp;
    a.out = ctx->sv;
    a.dx = (int32_t)d[0]; a.dy = (int32_t)d[1]; a.dz = (int32_t)d[2];
    const uint32_t N = ctx->svp.voxelAxisCount;
    for (int k = 0; k < 3; k++) a.k[k] = (d[k] == N) ? 1.0f : (float)(N - 1u) / (float)(d[k] - 1u);
    CR_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream), "event");
    CR_HIP(ctx, launch_sv(a, g.dtype, g.layout, ctx->stream), "shadow volume kernel");
    CR_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream), "event");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->svMs = ms;
    ctx->svDimsBuilt[0] = d[0]; ctx->svDimsBuilt[1] = d[1]; ctx->svDimsBuilt[2] = d[2];
    ctx->svValid = true;
    ctx->sceneVersion++;
    return CR_OK;
}

cr_status cr_sv_download(cr_context* ctx, float* out)
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    if (!ctx->svValid) return fail(ctx, CR_ERR_STATE, "shadow volume not built");
    size_t bytes = (size_t)ctx->svDimsBuilt[0] * ctx->svDimsBuilt[1] * ctx->svDimsBuilt[2] * 4;
    CR_HIP(ctx, hipMemcpyAsync(out, ctx->sv, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

// ------------------------------------------------------------- path tracer

cr_status cr_pt_clear(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    CR_HIP(ctx, zero_async(ctx->image, (size_t)ctx->W * ctx->H * 16, ctx->stream), "clear");
    return CR_OK;
}

} // extern "C"

namespace cr {

// Wait for the recorded render launches and add their durations.
cr_status fold_pt_events(cr_context* ctx)
{
    if (ctx->ptEventsUsed == 0) return CR_OK;
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    for (size_t i = 0; i < ctx->ptEventsUsed; i++) {
        float ms = 0.0f;
        CR_HIP(ctx, hipEventElapsedTime(&ms, ctx->ptEvents[i].first, ctx->ptEvents[i].second), "hipEventElapsedTime");
        ctx->ptTotalMs += ms;
        ctx->ptLaunches++;
        ctx->ptLastMs = ms;
    }
    ctx->ptEventsUsed = 0;
    return CR_OK;
}

cr_status upload_frames(cr_context* ctx, const cr_push_constants* frames, uint32_t n)
{
    std::vector<int32_t> fr(2 * (size_t)n);
    for (uint32_t i = 0; i < n; i++) {
        fr[2 * i] = frames[i].seed;
        fr[2 * i + 1] = (int32_t)frames[i].frameCount;
    }
    cr_status st = ensure_buffer(ctx, (void**)&ctx->dFrames, &ctx->framesCap, fr.size() * 4, "frames");
    if (st != CR_OK) return st;
    CR_HIP(ctx, hipMemcpyAsync(ctx->dFrames, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, ctx->stream), "copy frames");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync frames");
    return CR_OK;
}

void fill_job_args(PTArgs& a, const DeviceGrid& g, const float* sv, const uint32_t svDims[3],
                   const cr_camera_properties& cam, const cr_cloud_properties& cloud, const cr_parameters& params,
                   const cr_shadow_volume_properties& svp, const cr_render_ext& ext, const int32_t* frames,
                   uint32_t n_frames, float* image, int W, int H, uint32_t tileLog2W)
{
    memset(&a, 0, sizeof(a));
    a.grid = g.data; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nbx = g.nbx; a.nby = g.nby;
    a.skip = g.skip; a.mx = g.mx; a.my = g.my; a.mz = g.mz;
    const int32_t dims[3] = {g.nx, g.ny, g.nz};
    for (int k = 0; k < 3; k++) {
        a.bw[k] = ((cloud.bounds[1][k] - cloud.bounds[0][k]) * (float)g.mb) / (float)dims[k];
        a.ibw[k] = 1.0f / a.bw[k];
    }
    a.sv = sv;
    a.svx = (int32_t)svDims[0]; a.svy = (int32_t)svDims[1]; a.svz = (int32_t)svDims[2];
    a.cam = cam;
    a.cloud = cloud;
    a.params = params;
    a.svp = svp;
    a.ext = ext;
    a.sigmaBar = (cloud.maxExtinction * cloud.densityScaling) / cloud.baseScaling;
    a.frames = frames;
    a.nFrames = n_frames;
    a.image = (float4*)image;
    a.W = W; a.H = H;
    a.padded = 32u * ((uint32_t)W / 32u + 1u);
    a.tileLog2W = tileLog2W;
    const uint32_t tw = 1u << tileLog2W, th = 64u >> tileLog2W;
    a.tilesX = ((uint32_t)W + tw - 1u) / tw;
    const uint32_t tilesY = ((uint32_t)H + th - 1u) / th;
    a.totalIdx = a.tilesX * tilesY * 64u;
    a.base = 0;
}

cr_status fill_pt_args(cr_context* ctx, PTArgs& a, uint32_t n_frames)
{
    fill_job_args(a, ctx->grid, ctx->sv, ctx->svDimsBuilt, ctx->cam, ctx->cloud, ctx->params, ctx->svp, ctx->ext,
                  ctx->dFrames, n_frames, ctx->image, ctx->W, ctx->H, ctx->ptTileLog2W);
    a.queue = ctx->dQueue;
    a.counters = ctx->dCounters;
    return CR_OK;
}

} // namespace cr

extern "C" {

cr_status cr_pt_render(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames)
{
    CR_CHECK_CTX(ctx);
    if (!frames && n_frames) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null frames");
    if (n_frames == 0) return CR_OK;
    if (!ctx->hasGrid) return fail(ctx, CR_ERR_STATE, "no grid");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    if (ctx->ext.neeMode == CR_NEE_SHADOW_VOLUME && !ctx->svValid)
        return fail(ctx, CR_ERR_STATE, "shadow-volume NEE needs cr_sv_build after the last grid/shadow-volume change");
    hipSetDevice(ctx->device);
    cr_status st = upload_frames(ctx, frames, n_frames);
    if (st != CR_OK) return st;
    PTArgs a;
    fill_pt_args(ctx, a, n_frames);
    const DeviceGrid& g = ctx->grid;
    int per_cu = pt_blocks_per_cu(g.dtype, g.layout, ctx->ext.trackingMode, ctx->ext.neeMode, n_frames);
    int blocks = ctx->numCUs * per_cu;
    uint32_t needed = (a.totalIdx + 255u) / 256u;
    if ((uint32_t)blocks > needed) blocks = (int)needed;
    if (ctx->ptBlocksOverride > 0) blocks = ctx->ptBlocksOverride;
    CR_HIP(ctx, zero_async(ctx->dQueue, 4, ctx->stream), "zero queue");
    // Claim tiles in decreasing order of the previous render's per-tile work
    // (same image shape), so the persistent grid drains on cheap tiles.  Only
    // the schedule changes; every pixel's arithmetic is the same.
    const uint32_t nTiles = a.totalIdx / 64u;
    if (ctx->tileN != nTiles || ctx->tileLog2W != a.tileLog2W || !ctx->tileCost) {
        if (ctx->tileCost) hipFree(ctx->tileCost);
        ctx->tileCost = nullptr;
        ctx->tileCostValid = false;
        CR_HIP(ctx, hipMalloc((void**)&ctx->tileCost, (size_t)nTiles * 4), "hipMalloc tile costs");
        ctx->tileN = nTiles;
        ctx->tileLog2W = a.tileLog2W;
    }
    if (ctx->tileCostValid) {
        st = ensure_buffer(ctx, (void**)&ctx->tileSort, &ctx->tileSortCap,
                           (size_t)nTiles * 16 + pt_order_counts(nTiles), "tile order");
        if (st != CR_OK) return st;
        uint32_t* k = ctx->tileSort;
        CR_HIP(ctx, launch_tile_order(ctx->tileCost, nTiles, k, k + nTiles, k + 2 * (size_t)nTiles, k + 3 * (size_t)nTiles,
                                      k + 4 * (size_t)nTiles, ctx->stream), "tile order");
        a.tileOrder = k + nTiles;
    }
    CR_HIP(ctx, zero_async(ctx->tileCost, (size_t)nTiles * 4, ctx->stream), "zero tile costs");
    a.tileCost = ctx->tileCost;
    ctx->tileCostValid = true;
    if (ctx->ptEventsUsed >= 1024) {
        st = fold_pt_events(ctx);
        if (st != CR_OK) return st;
    }
    if (ctx->ptEventsUsed == ctx->ptEvents.size()) {
        std::pair<hipEvent_t, hipEvent_t> ev;
        CR_HIP(ctx, hipEventCreate(&ev.first), "hipEventCreate");
        CR_HIP(ctx, hipEventCreate(&ev.second), "hipEventCreate");
        ctx->ptEvents.push_back(ev);
    }
    const auto& ev = ctx->ptEvents[ctx->ptEventsUsed];
    CR_HIP(ctx, hipEventRecord(ev.first, ctx->stream), "event");
    CR_HIP(ctx, launch_pt(a, g.dtype, g.layout, blocks, ctx->stream), "path tracer kernel");
    CR_HIP(ctx, hipEventRecord(ev.second, ctx->stream), "event");
    ctx->ptEventsUsed++;
    return CR_OK;
}

cr_status cr_pt_download(cr_context* ctx, float* rgba)
{
    CR_CHECK_CTX(ctx);
    if (!rgba) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    CR_HIP(ctx, hipMemcpyAsync(rgba, ctx->image, (size_t)ctx->W * ctx->H * 16, hipMemcpyDeviceToHost, ctx->stream), "copy image");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_pt_upload(cr_context* ctx, const float* rgba)
{
    CR_CHECK_CTX(ctx);
    if (!rgba) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null input");
    if (!ctx->image) return fail(ctx, CR_ERR_STATE, "no camera set");
    CR_HIP(ctx, hipMemcpyAsync(ctx->image, rgba, (size_t)ctx->W * ctx->H * 16, hipMemcpyHostToDevice, ctx->stream), "copy image");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_pt_stats_get(cr_context* ctx, cr_pt_stats* out)
{
    CR_CHECK_CTX(ctx);
    if (!out) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null output");
    unsigned long long c[8];
    CR_HIP(ctx, hipMemcpyAsync(c, ctx->dCounters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream), "copy counters");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    out->paths = c[0];
    out->vertices = c[1];
    out->lookups = c[2];
    out->svLookups = c[3];
    out->brickVisits = c[4];
    cr_status st = fold_pt_events(ctx);
    if (st != CR_OK) return st;
    out->lastKernelMs = ctx->ptLastMs;
    out->totalKernelMs = ctx->ptTotalMs;
    out->launches = ctx->ptLaunches;
    return CR_OK;
}

cr_status cr_pt_stats_reset(cr_context* ctx)
{
    CR_CHECK_CTX(ctx);
    cr_status st = fold_pt_events(ctx);
    if (st != CR_OK) return st;
    ctx->ptTotalMs = 0.0;
    ctx->ptLaunches = 0;
    CR_HIP(ctx, zero_async(ctx->dCounters, 8 * sizeof(unsigned long long), ctx->stream), "zero counters");
    return CR_OK;
}

} // extern "C"
