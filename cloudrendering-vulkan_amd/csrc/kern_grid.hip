// Density-grid kernels: the seeded fBm cloud generator, format/layout
// conversion for uploads (Grid3D<float>::Load's z-fastest file order,
// Grid3D.h:73-85), FP32 readback, the 16^3 brick-majorant grid and the
// global maximum (Grid3D::GetMajorant, Grid3D.h:180-192).
//
// Generator definition (the reference ships no cloud; the build defines it):
//   hash    : PCG hash, Jarzynski & Olano, "Hash Functions for GPU Rendering",
//             JCGT 9(3), 2020.
//   lattice : h = H(seed ^ (octave * 0x9E3779B9)); h = H(h + ix); h = H(h + iy);
//             h = H(h + iz); value = (h >> 8) * 2^-24.
//   noise   : trilinear lattice interpolation with fade t^2 (3 - 2t),
//             lerp(a, b, t) = a + t (b - a).
//   fbm     : sum_o amp_o noise(p f_o), amp_0 = 1/2, amp *= gain, f *= lacunarity,
//             normalised by sum_o amp_o.
//   shape   : cumulus = radial falloff (0.20 to 0.48) about (0.5, 0.42, 0.5) of the normalised
//             box; anvil = height envelope times a radius widening from 0.22
//             to 0.48 between heights 0.50 and 0.80.
//   density : smoothstep(thresholdLo, thresholdHi, fbm * shape).
#include "cr_device.h"
#include "cr_internal.h"

using namespace crd;

namespace {

__device__ __forceinline__ uint32_t pcg_hash(uint32_t x)
{
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ float lattice(uint32_t hs, int32_t ix, int32_t iy, int32_t iz)
{
    uint32_t h = pcg_hash(hs + (uint32_t)ix);
    h = pcg_hash(h + (uint32_t)iy);
    h = pcg_hash(h + (uint32_t)iz);
    return (float)(h >> 8) * 5.9604644775390625e-8f;
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float fade(float t) { return (t * t) * (3.0f - 2.0f * t); }

__device__ __forceinline__ float value_noise(float x, float y, float z, uint32_t hs)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float sx = fade(x - fx), sy = fade(y - fy), sz = fade(z - fz);
    float c000 = lattice(hs, ix, iy, iz);
    float c100 = lattice(hs, ix + 1, iy, iz);
    float c010 = lattice(hs, ix, iy + 1, iz);
    float c110 = lattice(hs, ix + 1, iy + 1, iz);
    float c001 = lattice(hs, ix, iy, iz + 1);
    float c101 = lattice(hs, ix + 1, iy, iz + 1);
    float c011 = lattice(hs, ix, iy + 1, iz + 1);
    float c111 = lattice(hs, ix + 1, iy + 1, iz + 1);
    float n00 = lerpf(c000, c100, sx);
    float n10 = lerpf(c010, c110, sx);
    float n01 = lerpf(c001, c101, sx);
    float n11 = lerpf(c011, c111, sx);
    float n0 = lerpf(n00, n10, sy);
    float n1 = lerpf(n01, n11, sy);
    return lerpf(n0, n1, sz);
}

__device__ __forceinline__ float shape_fn(const cr_fbm_params& p, float ux, float uy, float uz)
{
    if (p.shape == CR_SHAPE_ANVIL) {
        float dx = ux - 0.5f, dz = uz - 0.5f;
        float rh = fsqrt(dx * dx + dz * dz);
        float R = 0.22f + 0.26f * smoothstep_f(0.50f, 0.80f, uy);
        float env = smoothstep_f(0.02f, 0.10f, uy) * (1.0f - smoothstep_f(0.90f, 0.98f, uy));
        return env * (1.0f - smoothstep_f(R * 0.55f, R, rh));
    }
    float dx = ux - 0.5f, dy = (uy - 0.42f) * 1.25f, dz = uz - 0.5f;
    float r = fsqrt((dx * dx + dy * dy) + dz * dz);
    return 1.0f - smoothstep_f(0.20f, 0.48f, r);
}

__device__ __forceinline__ float fbm_voxel(const cr_fbm_params& p, uint32_t ix, uint32_t iy, uint32_t iz)
{
    float cx = (float)ix + 0.5f, cy = (float)iy + 0.5f, cz = (float)iz + 0.5f;
    float ux = fdiv(cx, (float)p.size[0]), uy = fdiv(cy, (float)p.size[1]), uz = fdiv(cz, (float)p.size[2]);
    float sh = shape_fn(p, ux, uy, uz);
    if (sh <= 0.0f) return 0.0f;
    float x = cx * p.voxelSize, y = cy * p.voxelSize, z = cz * p.voxelSize;
    float sum = 0.0f, norm = 0.0f, amp = 0.5f, f = p.baseFrequency;
    for (uint32_t o = 0; o < p.octaves; o++) {
        uint32_t hs = pcg_hash(p.seed ^ (o * 0x9E3779B9u));
        sum = sum + amp * value_noise(x * f, y * f, z * f, hs);
        norm = norm + amp;
        amp = amp * p.gain;
        f = f * p.lacunarity;
    }
    float v = fdiv(sum, norm) * sh;
    return smoothstep_f(p.thresholdLo, p.thresholdHi, v);
}

template <int LAYOUT>
__device__ __forceinline__ uint64_t vidx(int32_t x, int32_t y, int32_t z, int32_t nx, int32_t ny, int32_t nbx, int32_t nby)
{
    GridView g;
    g.data = nullptr; g.nx = nx; g.ny = ny; g.nz = 0; g.nbx = nbx; g.nby = nby;
    return voxel_index<LAYOUT>(g, x, y, z);
}

template <typename T> __device__ __forceinline__ void store_voxel(T* p, uint64_t i, float v);
template <> __device__ __forceinline__ void store_voxel<float>(float* p, uint64_t i, float v) { p[i] = v; }
template <> __device__ __forceinline__ void store_voxel<__half>(__half* p, uint64_t i, float v)
{
    reinterpret_cast<uint16_t*>(p)[i] = f32_to_f16_rne(v);
}

// One thread per voxel of an x-fastest 3-D launch (x in blockIdx.x*256 + tid).
template <typename T, int LAYOUT>
__global__ void __launch_bounds__(256) fbm_kernel(cr_fbm_params p, T* out, int32_t nbx, int32_t nby)
{
    const int32_t x = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    const int32_t y = (int32_t)blockIdx.y, z = (int32_t)blockIdx.z;
    if (x >= (int32_t)p.size[0]) return;
    float v = fbm_voxel(p, (uint32_t)x, (uint32_t)y, (uint32_t)z);
    store_voxel<T>(out, vidx<LAYOUT>(x, y, z, (int32_t)p.size[0], (int32_t)p.size[1], nbx, nby), v);
}

template <typename S> __device__ __forceinline__ float load_src(const S* p, uint64_t i);
template <> __device__ __forceinline__ float load_src<float>(const float* p, uint64_t i) { return p[i]; }
template <> __device__ __forceinline__ float load_src<__half>(const __half* p, uint64_t i) { return __half2float(p[i]); }

// Source: dense x-fastest (Grid3D memory) or z-fastest (.xyz file order).
template <typename S, typename D, int LAYOUT>
__global__ void __launch_bounds__(256) convert_kernel(const S* src, int zfast, D* dst, int32_t nx, int32_t ny, int32_t nz,
                                                      int32_t nbx, int32_t nby)
{
    const int32_t x = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    const int32_t y = (int32_t)blockIdx.y, z = (int32_t)blockIdx.z;
    if (x >= nx) return;
    uint64_t si = zfast ? ((uint64_t)z + (uint64_t)nz * ((uint64_t)y + (uint64_t)ny * (uint64_t)x))
                        : ((uint64_t)x + (uint64_t)nx * ((uint64_t)y + (uint64_t)ny * (uint64_t)z));
    store_voxel<D>(dst, vidx<LAYOUT>(x, y, z, nx, ny, nbx, nby), load_src<S>(src, si));
}

template <typename T, int LAYOUT>
__global__ void __launch_bounds__(256) to_f32_kernel(const T* src, float* dst, int32_t nx, int32_t ny, int32_t nbx, int32_t nby)
{
    const int32_t x = (int32_t)(blockIdx.x * 256 + threadIdx.x);
    const int32_t y = (int32_t)blockIdx.y, z = (int32_t)blockIdx.z;
    if (x >= nx) return;
    dst[(uint64_t)x + (uint64_t)nx * ((uint64_t)y + (uint64_t)ny * (uint64_t)z)] =
        load_voxel<T>(src, vidx<LAYOUT>(x, y, z, nx, ny, nbx, nby));
}

// One block per mb^3 brick: max over the brick plus a one-voxel apron
// (the (mb+2)^3 voxels a trilinear lookup inside the brick can touch), start
// 0 as Grid3D::GetMajorant starts from T().
template <typename T, int LAYOUT>
__global__ void __launch_bounds__(256) majorant_kernel(const T* g, int32_t nx, int32_t ny, int32_t nz, int32_t nbx,
                                                       int32_t nby, int32_t mb, float* out, int32_t mx, int32_t my)
{
    const int32_t bx = (int32_t)blockIdx.x, by = (int32_t)blockIdx.y, bz = (int32_t)blockIdx.z;
    const int32_t x0 = max(bx * mb - 1, 0), x1 = min(bx * mb + mb, nx - 1);
    const int32_t y0 = max(by * mb - 1, 0), y1 = min(by * mb + mb, ny - 1);
    const int32_t z0 = max(bz * mb - 1, 0), z1 = min(bz * mb + mb, nz - 1);
    const int32_t ex = x1 - x0 + 1, ey = y1 - y0 + 1, ez = z1 - z0 + 1;
    const int32_t total = ex * ey * ez;
    float m = 0.0f;
    for (int32_t i = (int32_t)threadIdx.x; i < total; i += 256) {
        int32_t lx = i % ex, r = i / ex;
        int32_t ly = r % ey, lz = r / ey;
        float v = load_voxel<T>(g, vidx<LAYOUT>(x0 + lx, y0 + ly, z0 + lz, nx, ny, nbx, nby));
        if (v > m) m = v;
    }
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(uint64_t)bx + (uint64_t)mx * ((uint64_t)by + (uint64_t)my * (uint64_t)bz)] = red[0];
}

// Skip grid: occupancy, then 15 rounds of a 3x3x3 dilation (one axis per
// pass); a brick first covered in round d is d bricks (Chebyshev) from the
// nearest non-empty brick.  Bricks still uncovered keep -16.
// A non-empty brick stores its extinction majorant (m * densityScaling) /
// baseScaling, the value the brick walk needs.  Each brick holds two floats,
// that value and the majorant m itself, so the walk's one 8-byte fetch per
// brick also yields the m its density test divides by.
__global__ void __launch_bounds__(256) skip_init_kernel(const float* maj, uint32_t n, uint8_t* occ, float* skip, float ds,
                                                        float bs)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float m = maj[i];
    occ[i] = m > 0.0f ? 1 : 0;
    skip[2 * i] = m > 0.0f ? fdiv(m * ds, bs) : -16.0f;
    skip[2 * i + 1] = m;
}

__global__ void __launch_bounds__(256) skip_scale_kernel(const float* maj, uint32_t n, float* skip, float ds, float bs)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float m = maj[i];
    if (m > 0.0f) skip[2 * i] = fdiv(m * ds, bs);
}

__global__ void __launch_bounds__(256) dilate_kernel(const uint8_t* in, uint8_t* out, int32_t mx, int32_t my, int32_t mz,
                                                     int axis, float* skip, float d)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = (uint32_t)mx * (uint32_t)my * (uint32_t)mz;
    if (i >= n) return;
    const int32_t x = (int32_t)(i % (uint32_t)mx), y = (int32_t)((i / (uint32_t)mx) % (uint32_t)my);
    const int32_t z = (int32_t)(i / ((uint32_t)mx * (uint32_t)my));
    const int32_t c = axis == 0 ? x : (axis == 1 ? y : z);
    const int32_t N = axis == 0 ? mx : (axis == 1 ? my : mz);
    const uint32_t stride = axis == 0 ? 1u : (axis == 1 ? (uint32_t)mx : (uint32_t)mx * (uint32_t)my);
    uint8_t v = in[i];
    if (c > 0) v |= in[i - stride];
    if (c + 1 < N) v |= in[i + stride];
    out[i] = v;
    if (skip && v && skip[2 * i] == -16.0f) skip[2 * i] = -d;  // last pass of round d
}

// Majorants are >= 0, so their IEEE bit patterns order like unsigned integers.
__global__ void __launch_bounds__(256) max_kernel(const float* m, uint32_t n, unsigned int* out)
{
    unsigned int best = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) best = max(best, __float_as_uint(m[i]));
    for (int off = 32; off >= 1; off >>= 1) best = max(best, (unsigned int)__shfl_xor((int)best, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

dim3 vox_grid(int32_t nx, int32_t ny, int32_t nz) { return dim3((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)nz); }

} // namespace

namespace cr {

hipError_t launch_fbm(const cr_fbm_params& p, void* out, int dtype, int layout, int32_t nbx, int32_t nby, hipStream_t s)
{
    dim3 g = vox_grid((int32_t)p.size[0], (int32_t)p.size[1], (int32_t)p.size[2]);
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((fbm_kernel<__half, LAYOUT_BRICK4>), g, dim3(256), 0, s, p, (__half*)out, nbx, nby);
        else hipLaunchKernelGGL((fbm_kernel<__half, LAYOUT_DENSE>), g, dim3(256), 0, s, p, (__half*)out, nbx, nby);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((fbm_kernel<float, LAYOUT_BRICK4>), g, dim3(256), 0, s, p, (float*)out, nbx, nby);
        else hipLaunchKernelGGL((fbm_kernel<float, LAYOUT_DENSE>), g, dim3(256), 0, s, p, (float*)out, nbx, nby);
    }
    return hipGetLastError();
}

hipError_t launch_convert_grid(const void* src, int src_dtype, int zfast, void* dst, int dst_dtype, int dst_layout,
                               int32_t nx, int32_t ny, int32_t nz, int32_t nbx, int32_t nby, hipStream_t s)
{
    dim3 g = vox_grid(nx, ny, nz);
#define CR_CONV(S, D, L) hipLaunchKernelGGL((convert_kernel<S, D, L>), g, dim3(256), 0, s, (const S*)src, zfast, (D*)dst, nx, ny, nz, nbx, nby)
    if (src_dtype == CR_DTYPE_F32) {
        if (dst_dtype == CR_DTYPE_F32) { if (dst_layout == LAYOUT_BRICK4) CR_CONV(float, float, LAYOUT_BRICK4); else CR_CONV(float, float, LAYOUT_DENSE); }
        else { if (dst_layout == LAYOUT_BRICK4) CR_CONV(float, __half, LAYOUT_BRICK4); else CR_CONV(float, __half, LAYOUT_DENSE); }
    } else {
        if (dst_dtype == CR_DTYPE_F32) { if (dst_layout == LAYOUT_BRICK4) CR_CONV(__half, float, LAYOUT_BRICK4); else CR_CONV(__half, float, LAYOUT_DENSE); }
        else { if (dst_layout == LAYOUT_BRICK4) CR_CONV(__half, __half, LAYOUT_BRICK4); else CR_CONV(__half, __half, LAYOUT_DENSE); }
    }
#undef CR_CONV
    return hipGetLastError();
}

hipError_t launch_grid_to_f32(const void* src, int dtype, int layout, float* dst, int32_t nx, int32_t ny, int32_t nz,
                              int32_t nbx, int32_t nby, hipStream_t s)
{
    dim3 g = vox_grid(nx, ny, nz);
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((to_f32_kernel<__half, LAYOUT_BRICK4>), g, dim3(256), 0, s, (const __half*)src, dst, nx, ny, nbx, nby);
        else hipLaunchKernelGGL((to_f32_kernel<__half, LAYOUT_DENSE>), g, dim3(256), 0, s, (const __half*)src, dst, nx, ny, nbx, nby);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((to_f32_kernel<float, LAYOUT_BRICK4>), g, dim3(256), 0, s, (const float*)src, dst, nx, ny, nbx, nby);
        else hipLaunchKernelGGL((to_f32_kernel<float, LAYOUT_DENSE>), g, dim3(256), 0, s, (const float*)src, dst, nx, ny, nbx, nby);
    }
    return hipGetLastError();
}

hipError_t launch_majorants(const void* grid, int dtype, int layout, int32_t nx, int32_t ny, int32_t nz, int32_t nbx,
                            int32_t nby, int32_t mb, float* out, int32_t mx, int32_t my, int32_t mz, hipStream_t s)
{
    dim3 g((unsigned)mx, (unsigned)my, (unsigned)mz);
    if (dtype == CR_DTYPE_F16) {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((majorant_kernel<__half, LAYOUT_BRICK4>), g, dim3(256), 0, s, (const __half*)grid, nx, ny, nz, nbx, nby, mb, out, mx, my);
        else hipLaunchKernelGGL((majorant_kernel<__half, LAYOUT_DENSE>), g, dim3(256), 0, s, (const __half*)grid, nx, ny, nz, nbx, nby, mb, out, mx, my);
    } else {
        if (layout == LAYOUT_BRICK4) hipLaunchKernelGGL((majorant_kernel<float, LAYOUT_BRICK4>), g, dim3(256), 0, s, (const float*)grid, nx, ny, nz, nbx, nby, mb, out, mx, my);
        else hipLaunchKernelGGL((majorant_kernel<float, LAYOUT_DENSE>), g, dim3(256), 0, s, (const float*)grid, nx, ny, nz, nbx, nby, mb, out, mx, my);
    }
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) math_probe_kernel(int op, const float* in, uint32_t n, float* out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float x = in[i];
    float r = 0.0f, sn = 0.0f, cs = 0.0f;
    if (op == CR_MATH_SQRT) r = fsqrt(x);
    else if (op == CR_MATH_LOG) r = crd_logf(x);
    else if (op == CR_MATH_EXP) r = crd_expf(x);
    else if (op == CR_MATH_ACOS) r = crd_acosf(x);
    else if (op == CR_MATH_ATAN2) r = crd_atan2f(x, in[(i ^ 1u) < n ? (i ^ 1u) : i]);
    else if (op == CR_MATH_BF16) r = __uint_as_float(bf16x2_rne(x, 0.0f) << 16);
    else { crd_sincosf(x, &sn, &cs); r = op == CR_MATH_SIN ? sn : cs; }
    out[i] = r;
}

hipError_t launch_math_probe(int op, const float* in, uint32_t n, float* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(math_probe_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, op, in, n, out);
    return hipGetLastError();
}

hipError_t launch_skip_grid(const float* majorants, int32_t mx, int32_t my, int32_t mz, float ds, float bs, float* skip,
                            uint8_t* tmp, hipStream_t s)
{
    const uint32_t n = (uint32_t)mx * (uint32_t)my * (uint32_t)mz;
    const dim3 g((n + 255) / 256);
    uint8_t* a = tmp;
    uint8_t* b = tmp + n;
    hipLaunchKernelGGL(skip_init_kernel, g, dim3(256), 0, s, majorants, n, a, skip, ds, bs);
    for (int d = 1; d <= 15; d++) {
        hipLaunchKernelGGL(dilate_kernel, g, dim3(256), 0, s, a, b, mx, my, mz, 0, (float*)nullptr, 0.0f);
        hipLaunchKernelGGL(dilate_kernel, g, dim3(256), 0, s, b, a, mx, my, mz, 1, (float*)nullptr, 0.0f);
        hipLaunchKernelGGL(dilate_kernel, g, dim3(256), 0, s, a, b, mx, my, mz, 2, skip, (float)d);
        uint8_t* t = a;
        a = b;
        b = t;
    }
    return hipGetLastError();
}

hipError_t launch_skip_scale(const float* majorants, uint32_t n, float ds, float bs, float* skip, hipStream_t s)
{
    hipLaunchKernelGGL(skip_scale_kernel, dim3((n + 255) / 256), dim3(256), 0, s, majorants, n, skip, ds, bs);
    return hipGetLastError();
}

hipError_t launch_grid_max(const float* majorants, uint32_t n, unsigned int* out_bits, hipStream_t s)
{
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(max_kernel, dim3(blocks), dim3(256), 0, s, majorants, n, out_bits);
    return hipGetLastError();
}

} // namespace cr
