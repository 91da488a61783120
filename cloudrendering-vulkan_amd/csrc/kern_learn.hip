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

// The coarsest mip levels, as many as fit in this many floats (32 KB), are
// copied into LDS by each workgroup: for a 128^3 grid, levels 3-7 (16^3 down
// to 1^3, 18.3 KB), whose taps are 135 of a record's 216.
constexpr uint32_t FEATURE_LDS_FLOATS = 8192;

__device__ __forceinline__ uint32_t mip_floats(const cr::FeatureArgs& a, int k)
{
    return (uint32_t)a.dims[k][0] * (uint32_t)a.dims[k][1] * (uint32_t)a.dims[k][2];
}

__global__ void __launch_bounds__(256) features_kernel(cr::FeatureArgs a)
{
    __shared__ float mipLds[FEATURE_LDS_FLOATS];
    // The grid is sized for every record; the count of collisions among them
    // is known only on the device.  A workgroup wholly past it leaves before
    // staging (the test is the same for all its threads, so no thread is left
    // at the barrier).
    const uint64_t n = a.ndev ? (uint64_t)*a.ndev : a.n;
    if ((uint64_t)blockIdx.x * 256 >= n) return;
    // levels staged..LEARN_LEVELS-1 live in LDS, in that order
    int staged = cr::LEARN_LEVELS;
    uint32_t used = 0;
    while (staged > 0 && used + mip_floats(a, staged - 1) <= FEATURE_LDS_FLOATS) {
        staged--;
        used += mip_floats(a, staged);
    }
    {
        uint32_t off = 0;
        for (int k = staged; k < cr::LEARN_LEVELS; k++) {
            const uint32_t nk = mip_floats(a, k);
            for (uint32_t v = threadIdx.x; v < nk; v += 256u) mipLds[off + v] = a.mips[k][v];
            off += nk;
        }
    }
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
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
    // the row is written in four groups of two levels (54 bf16 = 27 words each,
    // the last with the two cosines and padding), so few values stay live
    uint32_t* out = reinterpret_cast<uint32_t*>(a.X + i * 224);
    float h = a.h0;
    uint32_t ldsOff = 0;  // the next staged level's place in mipLds
#pragma unroll 1
    for (int grp = 0; grp < 4; grp++) {
        uint32_t packed[28];
        int idx = 0;
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            const int k = 2 * grp + kk;
            GridView g;
            g.nx = a.dims[k][0]; g.ny = a.dims[k][1]; g.nz = a.dims[k][2];
            g.nbx = 0; g.nby = 0;
            // the 27 taps of level k; a staged level's copy reads LDS directly
            auto taps = [&](const GridView& gv) {
#pragma unroll
                for (int l = -1; l <= 1; l++)
#pragma unroll
                    for (int j = -1; j <= 1; j++)
#pragma unroll
                        for (int ii = -1; ii <= 1; ii++) {
                            const float fi = (float)ii * h, fj = (float)j * h, fl = (float)l * h;
                            const f3 p = mk(x.x + ((fi * ex.x + fj * ey.x) + fl * ez.x),
                                            x.y + ((fi * ex.y + fj * ey.y) + fl * ez.y),
                                            x.z + ((fi * ex.z + fj * ey.z) + fl * ez.z));
                            float d = 0.0f;
                            if (in_box(b0, b1, p)) d = trilinear<float, LAYOUT_DENSE>(gv, cloud_norm(a.cloud, p));
                            const uint16_t v = f2bf((d * sscale) * h);
                            if (idx & 1) packed[idx >> 1] |= (uint32_t)v << 16;
                            else packed[idx >> 1] = v;
                            idx++;
                        }
            };
            if (k >= staged) {
                g.data = mipLds + ldsOff;
                ldsOff += mip_floats(a, k);
                taps(g);
            } else {
                g.data = a.mips[k];
                taps(g);
            }
            h = h * 2.0f;
        }
        if (grp == 3) {
            const float c = dot(w, ez);
            packed[27] = (uint32_t)f2bf(c) | ((uint32_t)f2bf(c * c) << 16);
        }
        uint32_t* o = out + 27 * grp;
#pragma unroll
        for (int q = 0; q < 27; q++) o[q] = packed[q];
        if (grp == 3) {
            o[27] = packed[27];
            o[28] = 0u;
            o[29] = 0u;
            o[30] = 0u;
        }
    }
}

// The first collision's shadow-volume next-event term, in the path tracer's
// operation order (pt_kernel.h, EV_COLLIDE with thr = 1).
__device__ __forceinline__ float first_nee(const cr::FirstNee& e, f3 pos, f3 dir)
{
    const f3 ldir = ld3(e.svp.lightDirection);
    const float pdf = phase_pdf(e.params, neg(dir), ldir);
    const float w = 1.0f * e.albedo;
    GridView g;
    g.data = e.sv; g.nx = e.svx; g.ny = e.svy; g.nz = e.svz; g.nbx = 0; g.nby = 0;
    const float Tv = trilinear<float, LAYOUT_DENSE>(g, sv_norm(e.svp, pos));
    return (((1.0f * e.intensity) * pdf) * Tv) * w;
}

// Per pixel, frames in order: L = the escaped path's radiance, or
// expm1(network) at its first collision (plus the exact first next-event term
// when there is one), blended with the reference's running mean
// (PathTracer.comp:420-424).
__global__ void __launch_bounds__(256) compose_kernel(cr::ComposeArgs a)
{
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.pixels) return;
    float4 acc = a.image[p];
    for (uint32_t f = 0; f < a.nFrames; f++) {
        const size_t s = (size_t)f * a.pixels + p;
        const uint32_t slot = a.slot[s];
        float4 L;
        if (slot != 0xFFFFFFFFu) {
            const float* y = a.Y + (size_t)slot * 4;
            float e = 0.0f;
            if (a.nee.on) {
                const float4* r = reinterpret_cast<const float4*>(a.records) + 2 * s;
                const float4 r0 = r[0], r1 = r[1];
                e = first_nee(a.nee, mk(r0.x, r0.y, r0.z), mk(r0.w, r1.x, r1.y));
            }
            L.x = e + fmaxf(expm1f(y[0]), 0.0f);
            L.y = e + fmaxf(expm1f(y[1]), 0.0f);
            L.z = e + fmaxf(expm1f(y[2]), 0.0f);
            L.w = 1.0f;
        } else {
            L = reinterpret_cast<const float4*>(a.records)[2 * s];
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

// Stream compaction of the collided records, in record order (deterministic):
// per-block counts, a host-side exclusive scan, then the write pass.
__device__ __forceinline__ bool collided(const float* rec, uint64_t i) { return rec[i * 8 + 7] > 0.5f; }

__global__ void __launch_bounds__(256) count_kernel(const float* rec, uint64_t n, uint32_t* counts)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool f = i < n && collided(rec, i);
    __shared__ uint32_t wc[4];
    const unsigned long long bal = __ballot(f);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ void __launch_bounds__(256) compact_kernel(const float* rec, uint64_t n, const uint32_t* offsets, float* pts,
                                                      uint32_t* slot)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool f = i < n && collided(rec, i);
    __shared__ uint32_t wc[4];
    const unsigned long long bal = __ballot(f);
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) wc[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    if (i >= n) return;
    if (!f) {
        slot[i] = 0xFFFFFFFFu;
        return;
    }
    uint32_t s = offsets[blockIdx.x];
    for (uint32_t k = 0; k < w; k++) s += wc[k];
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    s += (uint32_t)__popcll(bal & lt);
    slot[i] = s;
    const float4* src = reinterpret_cast<const float4*>(rec + i * 8);
    float4* dst = reinterpret_cast<float4*>(pts + (size_t)s * 8);
    dst[0] = src[0];
    dst[1] = src[1];
}

// Exclusive scan of nb block counts in place, in one workgroup of 1024
// threads (each scans a contiguous run); *total = the sum.
__global__ void __launch_bounds__(1024) scan_counts_kernel(uint32_t* counts, uint32_t nb, uint32_t* total)
{
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023u) / 1024u;
    const uint32_t b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    uint32_t s = 0;
    for (uint32_t i = b0; i < b1; i++) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;  // exclusive prefix of this thread's run
    for (uint32_t i = b0; i < b1; i++) {
        const uint32_t c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (t == 1023u) *total = part[1023];
}

// Rows k*stride (k < count) of 8-float records copied to dst[k].
__global__ void __launch_bounds__(256) select_kernel(const float* src, uint32_t stride, uint32_t count, float* dst)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const float4* s4 = reinterpret_cast<const float4*>(src + (size_t)k * stride * 8);
    float4* d4 = reinterpret_cast<float4*>(dst + (size_t)k * 8);
    d4[0] = s4[0];
    d4[1] = s4[1];
}

// Rows k*stride (k < count) of X and T4 appended to the data set at `offset`.
__global__ void __launch_bounds__(256) append_kernel(const uint16_t* X, const float4* T4, const float* pts, cr::FirstNee nee,
                                                     uint32_t stride, uint32_t count, uint16_t* dsX, float* dsT,
                                                     uint64_t offset)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const uint64_t src = (uint64_t)k * stride, dst = offset + k;
    const uint4* xs = reinterpret_cast<const uint4*>(X + src * 224);
    uint4* xd = reinterpret_cast<uint4*>(dsX + dst * 224);
    for (int q = 0; q < 28; q++) xd[q] = xs[q];
    const float4 t = T4[src];
    float e = 0.0f;
    if (nee.on) {
        const float* p = pts + src * 8;
        e = first_nee(nee, mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]));
    }
    dsT[dst * 3 + 0] = t.x - e;
    dsT[dst * 3 + 1] = t.y - e;
    dsT[dst * 3 + 2] = t.z - e;
}

// A training batch: row r is data-set row hash(seed, step, r) mod n.
} // namespace

namespace cr {

hipError_t launch_count(const float* rec, uint64_t n, uint32_t* counts, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(count_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rec, n, counts);
    return hipGetLastError();
}

hipError_t launch_compact(const float* rec, uint64_t n, const uint32_t* offsets, float* pts, uint32_t* slot, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rec, n, offsets, pts, slot);
    return hipGetLastError();
}

hipError_t launch_scan_counts(uint32_t* counts, uint32_t nb, uint32_t* total, hipStream_t s)
{
    hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, counts, nb, total);
    return hipGetLastError();
}

hipError_t launch_select(const float* src, uint32_t stride, uint32_t count, float* dst, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(select_kernel, dim3((count + 255) / 256), dim3(256), 0, s, src, stride, count, dst);
    return hipGetLastError();
}

hipError_t launch_append(const uint16_t* X, const float4* T4, const float* pts, const FirstNee& nee, uint32_t stride,
                         uint32_t count, uint16_t* dsX, float* dsT, uint64_t offset, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(append_kernel, dim3((count + 255) / 256), dim3(256), 0, s, X, T4, pts, nee, stride, count, dsX, dsT,
                       offset);
    return hipGetLastError();
}

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
