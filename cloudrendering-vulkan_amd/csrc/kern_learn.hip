// Learned in-scattering support kernels (extension): the density mip pyramid,
// the hierarchical stencil features of a shading point, and the per-pixel
// composition of network outputs into the running-mean image.
//
// Features of a point x seen along omega (224 values, bf16):
//   for mip level k = 0..7 and (l, j, i) in {-1,0,1}^3 (i fastest):
//     p = x + h_k (i e_x + j e_y + l e_z),  h_k = h_0 2^k,
//     e_z = -L (towards the sun), (e_x, e_y) = createOrthonormalBasis(e_z)
//     f = density_k(p) * (densityScaling / baseScaling) * h_k   (0 outside the box)
//   then dot(omega, e_z), its square and six zeros.
#include "cr_device.h"
#include "cr_internal_learn.h"

using namespace crd;

namespace {

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Level k+1 = mean of the existing voxels of each 2x2x2 block of level k.
__global__ void __launch_bounds__(256) mip_kernel(const float* src, int32_t sx, int32_t sy, int32_t sz, float* dst,
                                                  int32_t dx, int32_t dy, int32_t dz)
{
    const int32_t x = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    const int32_t y = (int32_t)blockIdx.y, z = (int32_t)blockIdx.z;
    if (x >= dx) return;
    float acc = 0.0f, cnt = 0.0f;
    for (int c = 0; c < 8; c++) {
        const int32_t X = 2 * x + (c & 1), Y = 2 * y + ((c >> 1) & 1), Z = 2 * z + (c >> 2);
        if (X < sx && Y < sy && Z < sz) {
            acc += src[(size_t)X + (size_t)sx * ((size_t)Y + (size_t)sy * Z)];
            cnt += 1.0f;
        }
    }
    dst[(size_t)x + (size_t)dx * ((size_t)y + (size_t)dy * z)] = fdiv(acc, cnt);
}

__global__ void __launch_bounds__(256) features_kernel(cr::FeatureArgs a)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float* pt = a.points + i * 8;
    const f3 x = mk(pt[0], pt[1], pt[2]);
    const f3 w = mk(pt[3], pt[4], pt[5]);
    const f3 ez = mk(-a.light[0], -a.light[1], -a.light[2]);
    const float sz = ez.z >= 0.0f ? 1.0f : -1.0f;
    const float aa = fdiv(ez.y, 1.0f + fabsf(ez.z));
    const float bb = ez.y * aa;
    const float cc = -ez.x * aa;
    const f3 ex = mk(ez.z + sz * bb, sz * cc, -ez.x);
    const f3 ey = mk(cc, 1.0f - bb, -sz * ez.y);
    const float* b0 = a.cloud.bounds[0];
    const float* b1 = a.cloud.bounds[1];
    const float sscale = fdiv(a.cloud.densityScaling, a.cloud.baseScaling);
    uint16_t* out = a.X + i * 224;
    uint32_t packed[112];
    int idx = 0;
    float h = a.h0;
    for (int k = 0; k < cr::LEARN_LEVELS; k++) {
        GridView g;
        g.data = a.mips[k];
        g.nx = a.dims[k][0]; g.ny = a.dims[k][1]; g.nz = a.dims[k][2];
        g.nbx = 0; g.nby = 0;
        for (int l = -1; l <= 1; l++)
            for (int j = -1; j <= 1; j++)
                for (int ii = -1; ii <= 1; ii++) {
                    const float fi = (float)ii * h, fj = (float)j * h, fl = (float)l * h;
                    const f3 p = mk(x.x + ((fi * ex.x + fj * ey.x) + fl * ez.x), x.y + ((fi * ex.y + fj * ey.y) + fl * ez.y),
                                    x.z + ((fi * ex.z + fj * ey.z) + fl * ez.z));
                    float d = 0.0f;
                    if (in_box(b0, b1, p)) d = trilinear<float, LAYOUT_DENSE>(g, cloud_norm(a.cloud, p));
                    const uint16_t v = f2bf((d * sscale) * h);
                    if (idx & 1) packed[idx >> 1] |= (uint32_t)v << 16;
                    else packed[idx >> 1] = v;
                    idx++;
                }
        h = h * 2.0f;
    }
    const float c = dot(w, ez);
    packed[108] = (uint32_t)f2bf(c) | ((uint32_t)f2bf(c * c) << 16);
    packed[109] = packed[110] = packed[111] = 0u;
    uint4* o4 = reinterpret_cast<uint4*>(out);
#pragma unroll
    for (int q = 0; q < 28; q++) o4[q] = make_uint4(packed[4 * q], packed[4 * q + 1], packed[4 * q + 2], packed[4 * q + 3]);
}

// Per pixel, frames in order: L = escaped ? background : expm1(network) and
// the reference's running mean (PathTracer.comp:420-424).
__global__ void __launch_bounds__(256) compose_kernel(cr::ComposeArgs a)
{
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.pixels) return;
    float4 acc = a.image[p];
    for (uint32_t f = 0; f < a.nFrames; f++) {
        const size_t s = (size_t)f * a.pixels + p;
        const float4 e = a.sampleRes[s];  // w = 1 when the camera ray collided
        float4 L;
        if (e.w > 0.5f) {
            const float* y = a.Y + s * 4;
            L.x = fmaxf(expm1f(y[0]), 0.0f);
            L.y = fmaxf(expm1f(y[1]), 0.0f);
            L.z = fmaxf(expm1f(y[2]), 0.0f);
            L.w = 1.0f;
        } else {
            L = e;
            L.w = 1.0f;
        }
        const uint32_t fc = (uint32_t)a.frames[2 * f + 1];
        const float n = (float)fc, d = (float)(fc + 1u);
        acc.x = fdiv(L.x + acc.x * n, d);
        acc.y = fdiv(L.y + acc.y * n, d);
        acc.z = fdiv(L.z + acc.z * n, d);
        acc.w = fdiv(L.w + acc.w * n, d);
    }
    a.image[p] = acc;
}

} // namespace

namespace cr {

hipError_t launch_mip(const float* src, int32_t sx, int32_t sy, int32_t sz, float* dst, int32_t dx, int32_t dy, int32_t dz,
                      hipStream_t s)
{
    hipLaunchKernelGGL(mip_kernel, dim3((unsigned)((dx + 255) / 256), (unsigned)dy, (unsigned)dz), dim3(256), 0, s, src, sx,
                       sy, sz, dst, dx, dy, dz);
    return hipGetLastError();
}

hipError_t launch_features(const FeatureArgs& a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_compose(const ComposeArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(compose_kernel, dim3((a.pixels + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace cr
