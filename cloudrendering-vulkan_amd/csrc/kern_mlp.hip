// Learned in-scattering network on gfx950 MFMA (extension: the reference has
// no learned component).  224 -> 128 -> 128 -> 128 -> 128 -> 3, ReLU, bf16
// weights and activations, fp32 accumulation (v_mfma_f32_16x16x16_bf16).
//
// MFMA 16x16x16 operand layout (lane l of the wave, j = 0..3):
//   A (16 x 16, M x K): A[l % 16][4 (l / 16) + j]
//   B (16 x 16, K x N): B[4 (l / 16) + j][l % 16]
//   D (16 x 16, M x N): D[4 (l / 16) + j][l % 16]
// (checked on the device by mlp_probe_kernel against a host product).
//
// Forward: every wave owns 32 samples (two 16-row tiles) for the whole
// network.  Its activations stay in LDS (rows padded to 232 bf16 so the
// 8-byte fragment reads do not conflict); weight fragments are read from
// global memory in the transposed layout WT[n][k] so each lane's four k
// values are one 8-byte load.  Waves never exchange data, so the kernel has
// no barriers.  Inference (mlp_infer_kernel) instead copies each layer's WT
// into LDS once per workgroup and reads the B fragments there.
// Backward: per wave and layer, dW += A_in^T dZ (A^T fragments from LDS, the
// dZ accumulator tiles are already in B-operand layout), db += column sums of
// dZ, dA = dZ W^T (dZ staged through LDS into A layout, W read k-major), and
// dZ_prev = dA * (A_in > 0).  Every wave sums its parameter gradients into a
// slot of its own (the flat parameter layout; no atomics), looping over
// 32-row chunks w, w + nWaves, ...; the Adam kernel adds the slots in wave
// order, so a training step is bit-reproducible.
#include "cr_device.h"
#include "cr_internal_mlp.h"

using namespace crd;

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int ROW_STRIDE = 232;  // bf16 per LDS activation row (224 + 8)
constexpr int DZ_STRIDE = 136;   // bf16 per LDS dZ row (128 + 8)

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u = u + 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4s ld4(const uint16_t* p)
{
    v4s r;
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}

__global__ void __launch_bounds__(64) mlp_probe_kernel(const uint16_t* A, const uint16_t* B, float* D)
{
    // A: 16x16 row-major (M x K), B: 16x16 row-major (K x N), D = A B row-major
    const int l = threadIdx.x;
    v4s a, b;
    for (int j = 0; j < 4; j++) {
        a[j] = (short)A[(l % 16) * 16 + 4 * (l / 16) + j];
        b[j] = (short)B[(4 * (l / 16) + j) * 16 + (l % 16)];
    }
    v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = mfma(a, b, c);
    for (int j = 0; j < 4; j++) D[(4 * (l / 16) + j) * 16 + (l % 16)] = c[j];
}

// One dense layer for the wave's 32 rows: out = act[32 x K] * W + b.
template <int K, int N>
__device__ __forceinline__ void layer_fwd(const uint16_t* act, const uint16_t* WT, v4f (&acc)[2][8])
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int nt = 0; nt < N / 16; nt++) acc[rt][nt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for (int kb = 0; kb < K / 16; kb++) {
        const int kk = kb * 16 + 4 * (l / 16);
        const v4s a0 = ld4(act + (0 * 16 + l % 16) * ROW_STRIDE + kk);
        const v4s a1 = ld4(act + (1 * 16 + l % 16) * ROW_STRIDE + kk);
#pragma unroll
        for (int nt = 0; nt < N / 16; nt++) {
            const v4s b = ld4(WT + (size_t)(nt * 16 + l % 16) * K + kk);
            acc[0][nt] = mfma(a0, b, acc[0][nt]);
            acc[1][nt] = mfma(a1, b, acc[1][nt]);
        }
    }
}

template <bool TRAIN>
__global__ void __launch_bounds__(256) mlp_fwd_kernel(cr::MlpArgs a)
{
    __shared__ uint16_t lds[4 * 32 * ROW_STRIDE];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * 32;
    if (a.Bdev) a.B = (int64_t)*a.Bdev;
    if (row0 >= a.B) return;
    uint16_t* act = lds + wave * 32 * ROW_STRIDE;
    // X rows -> LDS (28 x 16-B chunks per 224-wide row)
    for (int idx = l; idx < 32 * 28; idx += 64) {
        const int r = idx / 28, c = idx % 28;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (row0 + r < a.B) v = *reinterpret_cast<const uint4*>(a.X + (size_t)(row0 + r) * 224 + c * 8);
        *reinterpret_cast<uint4*>(act + r * ROW_STRIDE + c * 8) = v;
    }
    v4f acc[2][8];
#pragma unroll
    for (int li = 0; li < 5; li++) {
        if (li == 0) layer_fwd<224, 128>(act, a.WT[0], acc);
        else if (li < 4) layer_fwd<128, 128>(act, a.WT[li], acc);
        else layer_fwd<128, 16>(act, a.WT[4], acc);
        const int N = li < 4 ? 128 : 16;
        const float* bias = a.bias[li];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int nt = 0; nt < 8; nt++) {
                if (nt * 16 >= N) break;
                const int n = nt * 16 + l % 16;
                const float bn = bias[n];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int r = rt * 16 + 4 * (l / 16) + j;
                    const float z = acc[rt][nt][j] + bn;
                    if (li < 4) {
                        const uint16_t h = f2bf(fmaxf(z, 0.0f));
                        act[r * ROW_STRIDE + n] = h;
                        if (TRAIN && row0 + r < a.B) a.acts[(size_t)li * a.actStride + (size_t)(row0 + r) * 128 + n] = h;
                    } else if (n < 3 && row0 + r < a.B) {
                        a.Y[(size_t)(row0 + r) * 4 + n] = z;
                    }
                }
            }
    }
}

// Inference: as mlp_fwd_kernel<false>, but each layer's weights are copied
// into LDS once per workgroup (rows padded by 8 bf16 against bank conflicts)
// and the four waves read their B fragments there, instead of every wave
// reading the whole network from L2.  The waves meet at two barriers per
// layer, so a wave whose rows lie past B still runs (on zero rows, storing
// nothing); only a workgroup with no row at all returns, all waves together.
__global__ void __launch_bounds__(256) mlp_infer_kernel(cr::MlpArgs a)
{
    __shared__ uint16_t lds_act[4 * 32 * ROW_STRIDE];
    __shared__ uint16_t lds_w[128 * ROW_STRIDE];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t B = a.Bdev ? (int64_t)*a.Bdev : a.B;
    if ((int64_t)blockIdx.x * 128 >= B) return;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * 32;
    uint16_t* act = lds_act + wave * 32 * ROW_STRIDE;
    for (int idx = l; idx < 32 * 28; idx += 64) {
        const int r = idx / 28, c = idx % 28;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (row0 + r < B) v = *reinterpret_cast<const uint4*>(a.X + (size_t)(row0 + r) * 224 + c * 8);
        *reinterpret_cast<uint4*>(act + r * ROW_STRIDE + c * 8) = v;
    }
    v4f acc[2][8];
#pragma unroll
    for (int li = 0; li < 5; li++) {
        const int K = li == 0 ? 224 : 128, N = li < 4 ? 128 : 16;
        const int WS = K + 8;
        const uint16_t* WT = a.WT[li];
        for (int idx = threadIdx.x; idx < N * (K / 8); idx += 256) {
            const int n = idx / (K / 8), c = idx % (K / 8);
            *reinterpret_cast<uint4*>(lds_w + n * WS + c * 8) = *reinterpret_cast<const uint4*>(WT + (size_t)n * K + c * 8);
        }
        __syncthreads();
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int nt = 0; nt < 8; nt++) acc[rt][nt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
        for (int kb = 0; kb < K / 16; kb++) {
            const int kk = kb * 16 + 4 * (l / 16);
            const v4s a0 = ld4(act + (0 * 16 + l % 16) * ROW_STRIDE + kk);
            const v4s a1 = ld4(act + (1 * 16 + l % 16) * ROW_STRIDE + kk);
#pragma unroll
            for (int nt = 0; nt < 8; nt++) {
                if (nt * 16 >= N) break;
                const v4s b = ld4(lds_w + (nt * 16 + l % 16) * WS + kk);
                acc[0][nt] = mfma(a0, b, acc[0][nt]);
                acc[1][nt] = mfma(a1, b, acc[1][nt]);
            }
        }
        const float* bias = a.bias[li];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int nt = 0; nt < 8; nt++) {
                if (nt * 16 >= N) break;
                const int n = nt * 16 + l % 16;
                const float bn = bias[n];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int r = rt * 16 + 4 * (l / 16) + j;
                    const float z = acc[rt][nt][j] + bn;
                    if (li < 4) act[r * ROW_STRIDE + n] = f2bf(fmaxf(z, 0.0f));
                    else if (n < 3 && row0 + r < B) a.Y[(size_t)(row0 + r) * 4 + n] = z;
                }
            }
        __syncthreads();  // every wave is done with this layer's weights
    }
}

// dY = 2 / (B * 3) (Y - log(1 + max(t, 0))); loss[block] = this block's share
// of the mean squared difference (the host adds the blocks in order).
__global__ void __launch_bounds__(256) mlp_loss_kernel(const float* Y, const float* T, int64_t B, float* dY, float* loss)
{
    __shared__ float waveSum[4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float s = 0.0f;
    if (i < B) {
        const float scale = 2.0f / ((float)B * 3.0f);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float d = Y[i * 4 + c] - log1pf(fmaxf(T[i * 3 + c], 0.0f));
            dY[i * 4 + c] = scale * d;
            s += d * d;
        }
        dY[i * 4 + 3] = 0.0f;
    }
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) waveSum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        loss[blockIdx.x] = (((waveSum[0] + waveSum[1]) + waveSum[2]) + waveSum[3]) / ((float)B * 3.0f);
}

// Backward through one layer for the wave's 32 rows.  dz holds dZ of this
// layer's output in accumulator layout; on return (li > 0) it holds dZ of the
// previous layer's output.  gW / gb: the wave's gradient slot of this layer;
// `first` (the wave's first chunk) stores instead of adding.
template <int K, int N>
__device__ __forceinline__ void layer_bwd(const cr::MlpArgs& a, int li, const uint16_t* ain, uint16_t* dzl, v4f (&dz)[2][8],
                                          float* gW, float* gb, bool first)
{
    const int l = threadIdx.x & 63;
    // (a) dW[k][n] += sum_rows ain[row][k] dz[row][n]
#pragma unroll 1
    for (int mt = 0; mt < K / 16; mt++) {
        v4s at[2];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int j = 0; j < 4; j++) at[rt][j] = (short)ain[(rt * 16 + 4 * (l / 16) + j) * ROW_STRIDE + mt * 16 + l % 16];
#pragma unroll
        for (int nt = 0; nt < N / 16; nt++) {
            v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                v4s b;
#pragma unroll
                for (int j = 0; j < 4; j++) b[j] = (short)f2bf(dz[rt][nt][j]);
                c = mfma(at[rt], b, c);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = mt * 16 + 4 * (l / 16) + j, n = nt * 16 + l % 16;
                float* g = gW + (size_t)k * N + n;
                *g = first ? c[j] : *g + c[j];
            }
        }
    }
    // (b) db[n] += sum_rows dz[row][n]
#pragma unroll
    for (int nt = 0; nt < N / 16; nt++) {
        float s = 0.0f;
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int j = 0; j < 4; j++) s += dz[rt][nt][j];
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (l < 16) {
            float* g = gb + nt * 16 + l;
            *g = first ? s : *g + s;
        }
    }
    if (li == 0) return;
    // (c) dA = dZ W^T: stage dZ (bf16) in LDS as [row][n]
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int nt = 0; nt < N / 16; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++) dzl[(rt * 16 + 4 * (l / 16) + j) * DZ_STRIDE + nt * 16 + l % 16] = f2bf(dz[rt][nt][j]);
    v4f da[2][8];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 8; ct++) da[rt][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* W = a.W[li];  // k-major [K][N]
#pragma unroll
    for (int kb = 0; kb < N / 16; kb++) {
        const int nn = kb * 16 + 4 * (l / 16);
        const v4s a0 = ld4(dzl + (0 * 16 + l % 16) * DZ_STRIDE + nn);
        const v4s a1 = ld4(dzl + (1 * 16 + l % 16) * DZ_STRIDE + nn);
#pragma unroll
        for (int ct = 0; ct < K / 16; ct++) {
            if (ct >= 8) break;
            const v4s b = ld4(W + (size_t)(ct * 16 + l % 16) * N + nn);
            da[0][ct] = mfma(a0, b, da[0][ct]);
            da[1][ct] = mfma(a1, b, da[1][ct]);
        }
    }
    // dZ of the previous layer's output: ReLU mask from its post-activation
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 8; ct++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = rt * 16 + 4 * (l / 16) + j, k = ct * 16 + l % 16;
                dz[rt][ct][j] = bf2f(ain[r * ROW_STRIDE + k]) > 0.0f ? da[rt][ct][j] : 0.0f;
            }
}

__global__ void __launch_bounds__(256) mlp_bwd_kernel(cr::MlpArgs a)
{
    __shared__ uint16_t lds_a[4 * 32 * ROW_STRIDE];
    __shared__ uint16_t lds_dz[4 * 32 * DZ_STRIDE];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t w = blockIdx.x * 4u + (uint32_t)wave;
    if (w >= a.nWaves) return;
    uint16_t* ain = lds_a + wave * 32 * ROW_STRIDE;
    uint16_t* dzl = lds_dz + wave * 32 * DZ_STRIDE;
    float* slot = a.partial + (size_t)w * a.total;
    for (int64_t chunk = w; chunk * 32 < a.B; chunk += a.nWaves) {
        const int64_t row0 = chunk * 32;
        const bool first = chunk == (int64_t)w;
        v4f dz[2][8];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t r = row0 + rt * 16 + 4 * (l / 16) + j;
                const int n = l % 16;
                dz[rt][0][j] = (r < a.B && n < 3) ? a.dY[r * 4 + n] : 0.0f;
            }
#pragma unroll
        for (int li = 4; li >= 0; li--) {
            // stage this layer's input rows (X or the previous layer's output) in LDS
            const int width = li == 0 ? 224 : 128;
            const int chunks = width / 8;
            for (int idx = l; idx < 32 * chunks; idx += 64) {
                const int r = idx / chunks, c = idx % chunks;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (row0 + r < a.B) {
                    const uint16_t* src = li == 0 ? a.X + (size_t)(row0 + r) * 224
                                                  : a.acts + (size_t)(li - 1) * a.actStride + (size_t)(row0 + r) * 128;
                    v = *reinterpret_cast<const uint4*>(src + c * 8);
                }
                *reinterpret_cast<uint4*>(ain + r * ROW_STRIDE + c * 8) = v;
            }
            float* gW = slot + a.wOff[li];
            float* gb = slot + a.bOff[li];
            if (li == 4) layer_bwd<128, 16>(a, li, ain, dzl, dz, gW, gb, first);
            else if (li > 0) layer_bwd<128, 128>(a, li, ain, dzl, dz, gW, gb, first);
            else layer_bwd<224, 128>(a, li, ain, dzl, dz, gW, gb, first);
        }
    }
}

// Gradient = the waves' slots added in order; Adam over the flat fp32
// parameters, the moving average when it is on, then refresh the bf16 copies
// (k-major W for backward, n-major WT for forward).
__global__ void __launch_bounds__(256) mlp_adam_kernel(cr::MlpArgs a, cr::AdamArgs o)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= o.n) return;
    // the waves' gradient slots, added in wave order
    float g = 0.0f;
    const float* part = a.partial + i;
#pragma unroll 8
    for (uint32_t w = 0; w < a.nWaves; w++) g += part[(size_t)w * a.total];
    a.grad[i] = g;
    const float m = o.beta1 * o.m[i] + (1.0f - o.beta1) * g;
    const float v = o.beta2 * o.v[i] + (1.0f - o.beta2) * g * g;
    o.m[i] = m;
    o.v[i] = v;
    const float p = o.params[i] - o.lr * fdiv(fdiv(m, o.c1), fsqrt(fdiv(v, o.c2)) + o.eps);
    o.params[i] = p;
    // moving average of the parameters, bias-corrected for inference
    float eh = 0.0f;
    if (o.ema) {
        const float e = o.emaDecay * o.ema[i] + (1.0f - o.emaDecay) * p;
        o.ema[i] = e;
        eh = e * o.emaCorr;
        o.emaHat[i] = eh;
    }
    // locate i in the padded layout; weights get their bf16 copies
    for (int li = 0; li < 5; li++) {
        const uint32_t K = li == 0 ? 224u : 128u, N = li < 4 ? 128u : 16u;
        const uint32_t w0 = a.wOff[li], b0 = a.bOff[li];
        if (i >= w0 && i < w0 + K * N) {
            const uint32_t k = (i - w0) / N, n = (i - w0) % N;
            const uint16_t h = f2bf(p);
            a.W[li][k * N + n] = h;
            a.WT[li][n * K + k] = h;
            if (o.ema) {
                const uint16_t he = f2bf(eh);
                o.EW[li][k * N + n] = he;
                o.EWT[li][n * K + k] = he;
            }
            return;
        }
        if (i >= b0 && i < b0 + N) return;
    }
}

} // namespace

namespace cr {

hipError_t launch_mlp_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpArgs& a, bool train, hipStream_t s)
{
    const unsigned blocks = (unsigned)((a.B + 127) / 128);
    if (blocks == 0) return hipSuccess;
    if (train) hipLaunchKernelGGL(mlp_fwd_kernel<true>, dim3(blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_loss(const float* Y, const float* T, int64_t B, float* dY, float* loss, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(mlp_loss_blocks(B)), dim3(256), 0, s, Y, T, B, dY, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const MlpArgs& a, hipStream_t s)
{
    const unsigned blocks = (a.nWaves + 3u) / 4u;
    if (blocks == 0 || a.B <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_bwd_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpArgs& a, const AdamArgs& o, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((o.n + 255) / 256), dim3(256), 0, s, a, o);
    return hipGetLastError();
}

} // namespace cr
