// Shadow-volume precompute (shaders/ShadowVolume.comp, dispatched by
// RenderTechniqueSV::RecordDrawCommands, RenderTechniqueSV.cpp:102-113).
//
// One lane per light-space column (x, y) marches the column along the light
// direction and stores T(x,y,z) = exp(-voxelSize * sum_{k<=z} sigma(p_k)),
// accumulated exactly as ShadowVolume.comp:84-93 states it
// (exp(-accumulatedDistance * accumulatedValue / (z + 1))).  Consecutive
// lanes own consecutive x, so every z-slice store is a coalesced row.
// With per-axis texel counts (extension) the lattice spacing along axis k is
// voxelSize * (N - 1) / (N_k - 1), N = voxelAxisCount; the cubic default is
// the reference lattice.
#include "cr_device.h"
#include "cr_internal.h"

using namespace crd;

namespace {

template <typename T, int LAYOUT>
__global__ void __launch_bounds__(256) sv_kernel(cr::SVArgs a)
{
    const int32_t x = (int32_t)(blockIdx.x * 64 + (threadIdx.x & 63));
    const int32_t y = (int32_t)(blockIdx.y * 4 + (threadIdx.x >> 6));
    if (x >= a.dx || y >= a.dy) return;
    GridView g;
    g.data = a.grid; g.nx = a.nx; g.ny = a.ny; g.nz = a.nz; g.nbx = a.nbx; g.nby = a.nby;
    const cr_shadow_volume_properties& P = a.svp;
    const float fx = (float)x * a.k[0], fy = (float)y * a.k[1];
    const float hz = P.voxelSize * a.k[2];
    float acc = 0.0f, accDist = 0.0f;
    const size_t plane = (size_t)a.dx * (size_t)a.dy;
    float* out = a.out + (size_t)x + (size_t)a.dx * (size_t)y;
    for (int32_t z = 0; z < a.dz; z++) {
        const float fz = (float)z * a.k[2];
        f3 p;
        p.x = P.bounds[0][0] + ((fx * P.right[0] + fy * P.up[0]) + fz * P.lightDirection[0]) * P.voxelSize;
        p.y = P.bounds[0][1] + ((fx * P.right[1] + fy * P.up[1]) + fz * P.lightDirection[1]) * P.voxelSize;
        p.z = P.bounds[0][2] + ((fx * P.right[2] + fy * P.up[2]) + fz * P.lightDirection[2]) * P.voxelSize;
        if (in_box(a.cloud.bounds[0], a.cloud.bounds[1], p)) {
            acc = acc + fdiv(trilinear<T, LAYOUT>(g, cloud_norm(a.cloud, p)) * a.cloud.densityScaling,
                             a.cloud.baseScaling);
        }
        accDist = accDist + hz;
        out[(size_t)z * plane] = crd_expf(fdiv((-accDist) * acc, (float)((uint32_t)z + 1u)));
    }
}

} // namespace

namespace cr {

hipError_t launch_sv(const SVArgs& a, int dtype, int layout, hipStream_t s)
{
    dim3 grid((unsigned)((a.dx + 63) / 64), (unsigned)((a.dy + 3) / 4));
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((sv_kernel<__half, LAYOUT_BRICK4>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((sv_kernel<__half, LAYOUT_DENSE>), grid, dim3(256), 0, s, a);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((sv_kernel<float, LAYOUT_BRICK4>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((sv_kernel<float, LAYOUT_DENSE>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

} // namespace cr
