// Learned in-scattering network on gfx950 MFMA (extension: the reference has
// no learned component).  224 -> 128 -> 128 -> 128 -> 128 -> 3, ReLU, bf16
// weights and activations, fp32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// MFMA 16x16x32 operand layout (lane l of the wave, g = l / 16, m = l % 16):
//   A (16 x 32, M x K): A[m][8 g + j], j = 0..7
//   B (32 x 16, K x N): B[8 g + j][m], j = 0..7
//   D (16 x 16, M x N): D[4 g + j][m], j = 0..3
// (checked on the device by mlp_probe_kernel against a host product).
//
// Training step (8192 rows: about 3.4 GFLOP, a few tens of microseconds):
//  * mlp_train_kernel, one workgroup per 32-row tile: draws the tile's rows
//    from the data set (or reads the given batch), runs the forward pass with
//    the activations in LDS (each of the four waves owns 32 output columns of
//    every hidden layer), forms the loss share and dY, and runs the backward
//    chain dA = dZ W^T, dZ_prev = dA * (H > 0).  It writes every layer's input
//    and every layer's output gradient feature-major ([feature][row], bf16)
//    for the weight-gradient GEMM, and no activations are re-read.
//  * mlp_wgrad_kernel: dW_l = H_l^T dZ_l and db_l = 1^T dZ_l as MFMA GEMMs
//    reducing over the batch (32 x 32 output tiles; the batch split into a few
//    contiguous ranges, one gradient slot each; the four waves of a
//    workgroup add their partial tiles in wave order).
//  * mlp_adam_kernel adds the slots in slot order, so a step is
//    bit-reproducible, then Adam and the bf16 weight copies.
// Inference (mlp_infer_kernel): persistent, one workgroup per CU holding the
// whole network in LDS (159,744 B of bf16 weights + biases); each wave takes
// 64 rows at a time and keeps their activations in registers between layers:
// it computes Z^T = W^T H^T, weights as the A operand, so a layer's D tiles
// (4 consecutive features of one row per lane) are, packed in pairs, the next
// layer's B operand once that layer's reduction index is permuted to match.
// The permutation is applied to the weights as they are copied to LDS.
#include "cr_device.h"
#include "cr_internal_mlp.h"

using namespace crd;

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int XS = 232;  // bf16 per LDS row of the 224-wide input tile (224 + 8)
constexpr int HS = 136;  // bf16 per LDS row of 128-wide tiles (128 + 8)

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u = u + 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef short v2s __attribute__((ext_vector_type(2)));

// Two floats -> packed bf16 (v_cvt_pk_bf16_f32, round to nearest even).
__device__ __forceinline__ uint32_t bf2(float lo, float hi) { return bf16x2_rne(lo, hi); }
// max(., 0) of two packed bf16 (v_pk_max_i16: a negative bf16 is a negative
// int16, -0 included).  Rounding then clamping equals clamping then rounding.
__device__ __forceinline__ uint32_t relu2(uint32_t p)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(v2s, p), (v2s{0, 0})));
}

__device__ __forceinline__ v4f mfma(v8bf a, v8bf b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v8bf as_bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }
__device__ __forceinline__ v8bf ld16(const uint16_t* p) { return as_bf(*reinterpret_cast<const uint4*>(p)); }
__device__ __forceinline__ uint32_t pk(uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); }
__device__ __forceinline__ v8bf zero8() { return as_bf(make_uint4(0u, 0u, 0u, 0u)); }

__global__ void __launch_bounds__(64) mlp_probe_kernel(const uint16_t* A, const uint16_t* B, float* D)
{
    // A: 16 x 32 row-major (M x K), B: 32 x 16 row-major (K x N), D = A B (16 x 16 row-major)
    const int l = threadIdx.x, m = l % 16, g = l / 16;
    uint16_t a[8], b[8];
    for (int j = 0; j < 8; j++) {
        a[j] = A[m * 32 + 8 * g + j];
        b[j] = B[(8 * g + j) * 16 + m];
    }
    const uint4 ua = make_uint4(pk(a[0], a[1]), pk(a[2], a[3]), pk(a[4], a[5]), pk(a[6], a[7]));
    const uint4 ub = make_uint4(pk(b[0], b[1]), pk(b[2], b[3]), pk(b[4], b[5]), pk(b[6], b[7]));
    v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = mfma(as_bf(ua), as_bf(ub), c);
    for (int j = 0; j < 4; j++) D[(4 * g + j) * 16 + m] = c[j];
}

// ---------------------------------------------------------------------------
// Training

struct TrainLds {
    uint16_t x[32 * XS];       // the tile's input rows
    uint16_t h[4][32 * HS];    // H1..H4 (ReLU outputs of layers 0..3)
    uint16_t dz[2][32 * HS];   // output gradients, ping-pong between backward layers
    uint64_t src[32];          // the tile's source rows
    float lossPart[2];
};

// One hidden layer forward for the tile: wave w computes output columns
// [32 w, 32 w + 32) of all 32 rows, H_out = max(H_in WT^T + b, 0), into LDS and
// feature-major into `trOut` (the weight-gradient GEMM's copy).
template <int K, int IS>
__device__ __forceinline__ void fwd_layer(const uint16_t* in, const uint16_t* WT, const float* bias, uint16_t* out,
                                          uint16_t* trOut, uint64_t Bp, int64_t row0)
{
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
    const int n0 = wave * 32;
    v4f acc[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++) acc[rt][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < K / 32; q++) {
        const v8bf b0 = ld16(WT + (size_t)(n0 + m) * K + 32 * q + 8 * g);
        const v8bf b1 = ld16(WT + (size_t)(n0 + 16 + m) * K + 32 * q + 8 * g);
        const v8bf a0 = ld16(in + m * IS + 32 * q + 8 * g);
        const v8bf a1 = ld16(in + (16 + m) * IS + 32 * q + 8 * g);
        acc[0][0] = mfma(a0, b0, acc[0][0]);
        acc[0][1] = mfma(a0, b1, acc[0][1]);
        acc[1][0] = mfma(a1, b0, acc[1][0]);
        acc[1][1] = mfma(a1, b1, acc[1][1]);
    }
#pragma unroll
    for (int ct = 0; ct < 2; ct++) {
        const int n = n0 + ct * 16 + m;
        const float bn = bias[n];
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            const v4f z = acc[rt][ct];
            const uint2 h = make_uint2(relu2(bf2(z[0] + bn, z[1] + bn)), relu2(bf2(z[2] + bn, z[3] + bn)));
            uint16_t* o = out + (rt * 16 + 4 * g) * HS + n;
            o[0] = (uint16_t)h.x;
            o[HS] = (uint16_t)(h.x >> 16);
            o[2 * HS] = (uint16_t)h.y;
            o[3 * HS] = (uint16_t)(h.y >> 16);
            *reinterpret_cast<uint2*>(trOut + (size_t)n * Bp + row0 + rt * 16 + 4 * g) = h;
        }
    }
}

// One layer backward for the tile: dA = dZ W^T over the layer's N outputs
// (N = 16 for the last layer, zero-padded to one 32-wide step), then the
// input's gradient dZ_prev = dA * (H > 0).  Wave w owns input columns
// [32 w, 32 w + 32).
template <int N>
__device__ __forceinline__ void bwd_layer(const uint16_t* dzIn, const uint16_t* W, const uint16_t* h, uint16_t* dzOut,
                                          uint16_t* trOut, uint64_t Bp, int64_t row0)
{
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
    const int c0 = wave * 32;
    v4f acc[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++) acc[rt][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < (N + 31) / 32; q++) {
        v8bf b[2];
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
            b[ct] = (N == 16 && g >= 2) ? zero8() : ld16(W + (size_t)(c0 + ct * 16 + m) * N + 32 * q + 8 * g);
        const v8bf a0 = ld16(dzIn + m * HS + 32 * q + 8 * g);
        const v8bf a1 = ld16(dzIn + (16 + m) * HS + 32 * q + 8 * g);
        acc[0][0] = mfma(a0, b[0], acc[0][0]);
        acc[0][1] = mfma(a0, b[1], acc[0][1]);
        acc[1][0] = mfma(a1, b[0], acc[1][0]);
        acc[1][1] = mfma(a1, b[1], acc[1][1]);
    }
#pragma unroll
    for (int ct = 0; ct < 2; ct++) {
        const int c = c0 + ct * 16 + m;
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = rt * 16 + 4 * g + j;
                v[j] = bf2f(h[r * HS + c]) > 0.0f ? acc[rt][ct][j] : 0.0f;
            }
            const uint2 d = make_uint2(bf2(v[0], v[1]), bf2(v[2], v[3]));
            if (dzOut) {
                uint16_t* o = dzOut + (rt * 16 + 4 * g) * HS + c;
                o[0] = (uint16_t)d.x;
                o[HS] = (uint16_t)(d.x >> 16);
                o[2 * HS] = (uint16_t)d.y;
                o[3 * HS] = (uint16_t)(d.y >> 16);
            }
            *reinterpret_cast<uint2*>(trOut + (size_t)c * Bp + row0 + rt * 16 + 4 * g) = d;
        }
    }
}

// Forward, loss and the backward chain of one 32-row tile (see the file head).
// loss[tile] = the tile's share of the mean squared difference between the
// outputs and log(1 + max(target, 0)), dY = 2 / (3 B) (Y - log(1 + max(t, 0))).
__global__ void __launch_bounds__(256) mlp_train_kernel(cr::MlpArgs a)
{
    __shared__ TrainLds S;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, m = l & 15, g = l >> 4;
    const int64_t row0 = (int64_t)blockIdx.x * 32;
    const int64_t B = a.B;
    const uint64_t Bp = a.Bp;
    uint16_t* tr = a.tr;
    if (tid < 32)
        S.src[tid] = a.dsX ? cr::mlp_draw(a.seed, a.step, (uint32_t)(row0 + tid), a.dsN) : (uint64_t)(row0 + tid);
    __syncthreads();
    const uint16_t* Xs = a.dsX ? a.dsX : a.X;
    for (int idx = tid; idx < 32 * 28; idx += 256) {
        const int r = idx / 28, c = idx % 28;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (row0 + r < B) v = *reinterpret_cast<const uint4*>(Xs + S.src[r] * 224 + c * 8);
        *reinterpret_cast<uint4*>(S.x + r * XS + c * 8) = v;
    }
    __syncthreads();
    // X^T: (feature, 8 rows) per item
    for (int idx = tid; idx < 224 * 4; idx += 256) {
        const int k = idx >> 2, r8 = (idx & 3) * 8;
        uint16_t h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] = S.x[(r8 + j) * XS + k];
        *reinterpret_cast<uint4*>(tr + (size_t)(cr::MLP_TR_IN[0] + k) * Bp + row0 + r8) =
            make_uint4(pk(h[0], h[1]), pk(h[2], h[3]), pk(h[4], h[5]), pk(h[6], h[7]));
    }
    fwd_layer<224, XS>(S.x, a.WT[0], a.bias[0], S.h[0], tr + (size_t)cr::MLP_TR_IN[1] * Bp, Bp, row0);
    __syncthreads();
    fwd_layer<128, HS>(S.h[0], a.WT[1], a.bias[1], S.h[1], tr + (size_t)cr::MLP_TR_IN[2] * Bp, Bp, row0);
    __syncthreads();
    fwd_layer<128, HS>(S.h[1], a.WT[2], a.bias[2], S.h[2], tr + (size_t)cr::MLP_TR_IN[3] * Bp, Bp, row0);
    __syncthreads();
    fwd_layer<128, HS>(S.h[2], a.WT[3], a.bias[3], S.h[3], tr + (size_t)cr::MLP_TR_IN[4] * Bp, Bp, row0);
    __syncthreads();
    // output layer, loss and dY: waves 0 and 1, one 16-row tile each
    if (wave < 2) {
        const int rt = wave;
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 4; q++)
            acc = mfma(ld16(S.h[3] + (rt * 16 + m) * HS + 32 * q + 8 * g), ld16(a.WT[4] + m * 128 + 32 * q + 8 * g), acc);
        const float bn = a.bias[4][m];
        const float scale = 2.0f / ((float)B * 3.0f);
        float lsum = 0.0f;
        float dy[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = rt * 16 + 4 * g + j;
            const int64_t row = row0 + r;
            dy[j] = 0.0f;
            if (m < 3 && row < B) {
                const float t = a.dsX ? a.dsT[S.src[r] * 3 + m] : a.T[row * 3 + m];
                const float diff = (acc[j] + bn) - log1pf(fmaxf(t, 0.0f));
                dy[j] = scale * diff;
                lsum += diff * diff;
            }
        }
        const uint2 d = make_uint2(bf2(dy[0], dy[1]), bf2(dy[2], dy[3]));
        const uint16_t dh[4] = {(uint16_t)d.x, (uint16_t)(d.x >> 16), (uint16_t)d.y, (uint16_t)(d.y >> 16)};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = rt * 16 + 4 * g + j;
            S.dz[0][r * HS + m] = dh[j];
            S.dz[0][r * HS + 16 + m] = 0;  // the padded half of the 32-wide reduction step
        }
        *reinterpret_cast<uint2*>(tr + (size_t)(cr::MLP_TR_DZ[4] + m) * Bp + row0 + rt * 16 + 4 * g) = d;
        for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
        if (l == 0) S.lossPart[wave] = lsum;
    }
    __syncthreads();
    if (tid == 0) a.loss[blockIdx.x] = (S.lossPart[0] + S.lossPart[1]) / ((float)B * 3.0f);
    bwd_layer<16>(S.dz[0], a.W[4], S.h[3], S.dz[1], tr + (size_t)cr::MLP_TR_DZ[3] * Bp, Bp, row0);
    __syncthreads();
    bwd_layer<128>(S.dz[1], a.W[3], S.h[2], S.dz[0], tr + (size_t)cr::MLP_TR_DZ[2] * Bp, Bp, row0);
    __syncthreads();
    bwd_layer<128>(S.dz[0], a.W[2], S.h[1], S.dz[1], tr + (size_t)cr::MLP_TR_DZ[1] * Bp, Bp, row0);
    __syncthreads();
    bwd_layer<128>(S.dz[1], a.W[1], S.h[0], nullptr, tr + (size_t)cr::MLP_TR_DZ[0] * Bp, Bp, row0);
}

// One output tile of dW_l = H_l^T dZ_l (32 input features x NT * 16 outputs)
// over the slot's range of 32-row chunks, and with kt == 0 also db_l = 1^T dZ_l.
template <int NT>
__device__ __forceinline__ void wgrad_tile(const cr::MlpArgs& a, int layer, int kt, int nt, uint32_t s)
{
    __shared__ float red[3][2][2][4][64];
    __shared__ float redb[3][2][64];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, m = l & 15, g = l >> 4;
    const uint64_t Bp = a.Bp;
    const uint16_t* A = a.tr + (size_t)(cr::MLP_TR_IN[layer] + kt * 32) * Bp;
    const uint16_t* D = a.tr + (size_t)(cr::MLP_TR_DZ[layer] + nt * 32) * Bp;
    const uint64_t C = Bp / 32;
    const uint64_t c0 = s * C / a.slots, c1 = (s + 1) * C / a.slots;
    const v8bf ones = as_bf(make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u));
    v4f acc[2][2], accb[2];
#pragma unroll
    for (int ct = 0; ct < 2; ct++) {
        accb[ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int mt = 0; mt < 2; mt++) acc[mt][ct] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    for (uint64_t c = c0 + wave; c < c1; c += 4) {
        const uint64_t b0 = c * 32 + 8 * g;
        const v8bf a0 = ld16(A + (size_t)m * Bp + b0);
        const v8bf a1 = ld16(A + (size_t)(16 + m) * Bp + b0);
#pragma unroll
        for (int ct = 0; ct < NT; ct++) {
            const v8bf bb = ld16(D + (size_t)(ct * 16 + m) * Bp + b0);
            acc[0][ct] = mfma(a0, bb, acc[0][ct]);
            acc[1][ct] = mfma(a1, bb, acc[1][ct]);
            if (kt == 0) accb[ct] = mfma(ones, bb, accb[ct]);
        }
    }
    // add the waves' tiles in wave order
    if (wave > 0) {
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int ct = 0; ct < NT; ct++)
#pragma unroll
                for (int j = 0; j < 4; j++) red[wave - 1][mt][ct][j][l] = acc[mt][ct][j];
#pragma unroll
        for (int ct = 0; ct < NT; ct++) redb[wave - 1][ct][l] = accb[ct][0];
    }
    __syncthreads();
    if (wave != 0) return;
    const uint32_t N = cr::MLP_N[layer];
    float* slot = a.slot + (size_t)s * a.total;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int ct = 0; ct < NT; ct++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float v = ((acc[mt][ct][j] + red[0][mt][ct][j][l]) + red[1][mt][ct][j][l]) + red[2][mt][ct][j][l];
                const uint32_t k = kt * 32 + mt * 16 + 4 * g + j, n = nt * 32 + ct * 16 + m;
                slot[a.wOff[layer] + k * N + n] = v;
            }
    if (kt == 0 && g == 0) {
#pragma unroll
        for (int ct = 0; ct < NT; ct++)
            slot[a.bOff[layer] + nt * 32 + ct * 16 + m] = ((accb[ct][0] + redb[0][ct][l]) + redb[1][ct][l]) + redb[2][ct][l];
    }
}

// Grid: MLP_GRAD_TILES x slots workgroups; tile t of slot s is block s * tiles + t.
__global__ void __launch_bounds__(256) mlp_wgrad_kernel(cr::MlpArgs a)
{
    const uint32_t t = blockIdx.x % cr::MLP_GRAD_TILES, s = blockIdx.x / cr::MLP_GRAD_TILES;
    if (t < 28) {
        wgrad_tile<2>(a, 0, (int)(t / 4), (int)(t % 4), s);
    } else if (t < 76) {
        const uint32_t u = t - 28;
        wgrad_tile<2>(a, 1 + (int)(u / 16), (int)((u % 16) / 4), (int)(u % 4), s);
    } else {
        wgrad_tile<1>(a, 4, (int)(t - 76), 0, s);
    }
}

// Gradient = the slots added in order; Adam over the flat fp32 parameters,
// the moving average when it is on, then refresh the bf16 copies (k-major W
// for backward, n-major WT for forward).
__global__ void __launch_bounds__(256) mlp_adam_kernel(cr::MlpArgs a, cr::AdamArgs o)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= o.n) return;
    float g = 0.0f;
    const float* part = a.slot + i;
    for (uint32_t s = 0; s < a.slots; s++) g += part[(size_t)s * a.total];
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

// ---------------------------------------------------------------------------
// Inference

// LDS image of the network: per layer, per 32-wide reduction step q and
// output n, four 16-B chunks (the A fragments of lanes g = 0..3), chunk g
// stored at position g ^ swz(n % 16) so that the 16 lanes of one g read 16
// distinct 16-B bank groups.  Layer 0 keeps the natural reduction order
// (chunk g = WT[n][32 q + 8 g .. + 7]); layers 1-4 take their inputs from the
// previous layer's D tiles, so chunk g = WT[n][32 q + 4 g .. + 3] followed by
// WT[n][32 q + 16 + 4 g .. + 3].
constexpr uint32_t IW_OFF[cr::MLP_LAYERS] = {0, 57344, 90112, 122880, 155648};  // bytes
constexpr uint32_t IB_OFF = 159744;                                             // 4 x 128 + 16 fp32 biases
constexpr uint32_t I_LDS = IB_OFF + 528 * 4;
constexpr uint32_t I_CHUNKS = IB_OFF / 16;
constexpr int I_WAVES = 4;
constexpr int I_ROWS = 64;  // rows per wave task: four 16-row blocks

__device__ __forceinline__ int swz(int m) { return ((m >> 2) ^ ((m << 1) & 2)) & 3; }

__device__ __forceinline__ v8bf lds_w(const uint8_t* lds, uint32_t layerOff, int N, int q, int nt, int m, int g)
{
    return as_bf(*reinterpret_cast<const uint4*>(lds + layerOff + ((uint32_t)((q * N + nt * 16 + m) * 4 + (g ^ swz(m))) << 4)));
}

// ReLU of 8 output tiles (biases already in the accumulators) -> the next
// layer's B operands (4 steps x 4 row blocks).
__device__ __forceinline__ void to_act(const v4f (&acc)[8][4], v8bf (&act)[4][4])
{
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
            const v4f lo = acc[2 * q][bb], hi = acc[2 * q + 1][bb];
            act[q][bb] = as_bf(make_uint4(relu2(bf2(lo[0], lo[1])), relu2(bf2(lo[2], lo[3])), relu2(bf2(hi[0], hi[1])),
                                          relu2(bf2(hi[2], hi[3]))));
        }
}

// Accumulators of a layer's 8 output tiles, started at the biases of the
// outputs each lane holds (4 g + j of every 16-wide tile).
__device__ __forceinline__ void bias_init(const float* bias, int g, v4f (&acc)[8][4])
{
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
        const float4 b = *reinterpret_cast<const float4*>(bias + 16 * nt + 4 * g);
#pragma unroll
        for (int bb = 0; bb < 4; bb++) acc[nt][bb] = v4f{b.x, b.y, b.z, b.w};
    }
}

// Layer 0's B operands for rows [row0, row0 + 64): lane (g, m) takes 8
// features of row row0 + 16 bb + m per 32-wide step (zeros past the batch).
__device__ __forceinline__ void load_rows(const uint16_t* X, int64_t row0, int64_t B, int m, int g, v8bf (&x)[7][4])
{
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
        const int64_t row = row0 + bb * 16 + m;
        const bool in = row < B;
        const uint16_t* p = X + (size_t)(in ? row : 0) * 224 + 8 * g;
#pragma unroll
        for (int q = 0; q < 7; q++) x[q][bb] = in ? ld16(p + 32 * q) : zero8();
    }
}

__global__ void __launch_bounds__(I_WAVES * 64) mlp_infer_kernel(cr::MlpArgs a)
{
    __shared__ __align__(16) uint8_t lds[I_LDS];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, m = l & 15, g = l >> 4;
    // copy the network into LDS
    for (uint32_t ci = tid; ci < I_CHUNKS; ci += I_WAVES * 64) {
        int layer;
        uint32_t rem;
        if (ci < 3584) { layer = 0; rem = ci; }
        else if (ci < 3584 + 3 * 2048) { layer = 1 + (int)((ci - 3584) / 2048); rem = (ci - 3584) % 2048; }
        else { layer = 4; rem = ci - 3584 - 3 * 2048; }
        const uint32_t N = cr::MLP_N[layer], K = cr::MLP_K[layer];
        const uint32_t q = rem / (N * 4), n = (rem / 4) % N, gs = (rem % 4) ^ (uint32_t)swz((int)(n & 15));
        const uint16_t* row = a.WT[layer] + (size_t)n * K + 32 * q;
        uint4 v;
        if (layer == 0) {
            v = *reinterpret_cast<const uint4*>(row + 8 * gs);
        } else {
            const uint2 lo = *reinterpret_cast<const uint2*>(row + 4 * gs);
            const uint2 hi = *reinterpret_cast<const uint2*>(row + 16 + 4 * gs);
            v = make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
        *reinterpret_cast<uint4*>(lds + IW_OFF[layer] + (rem << 4)) = v;
    }
    float* bias = reinterpret_cast<float*>(lds + IB_OFF);
    for (int i = tid; i < 528; i += I_WAVES * 64) bias[i] = i < 512 ? a.bias[i >> 7][i & 127] : a.bias[4][i - 512];
    __syncthreads();

    const int64_t B = a.Bdev ? (int64_t)*a.Bdev : a.B;
    const int64_t stride = (int64_t)gridDim.x * I_WAVES;
    int64_t task = (int64_t)blockIdx.x * I_WAVES + wave;
    // The wave's next 64 feature rows, as layer 0's B operands: loaded one
    // task ahead, so their HBM latency overlaps the hidden layers.
    v8bf xn[7][4];
    load_rows(a.X, task * I_ROWS, B, m, g, xn);
    for (; task * I_ROWS < B; task += stride) {
        const int64_t row0 = task * I_ROWS;
        // An opaque zero offset: the weight fragments are loop-invariant LDS
        // reads, and hoisted out of the task loop they would need about 400
        // registers.
        uint32_t zo = 0;
        asm volatile("" : "+s"(zo));
        const uint8_t* W = lds + zo;
        v4f acc[8][4];
        bias_init(bias, g, acc);
#pragma unroll
        for (int q = 0; q < 7; q++) {
            // one step's fragments at a time (another opaque offset)
            uint32_t zq = zo;
            asm volatile("" : "+s"(zq));
#pragma unroll
            for (int nt = 0; nt < 8; nt++) {
                const v8bf w = lds_w(lds + zq, IW_OFF[0], 128, q, nt, m, g);
#pragma unroll
                for (int bb = 0; bb < 4; bb++) acc[nt][bb] = mfma(w, xn[q][bb], acc[nt][bb]);
            }
        }
        load_rows(a.X, (task + stride) * I_ROWS, B, m, g, xn);
        v8bf act[4][4];
        to_act(acc, act);
#pragma unroll
        for (int layer = 1; layer < 4; layer++) {
            bias_init(bias + 128 * layer, g, acc);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint32_t zq = zo;
                asm volatile("" : "+s"(zq));
#pragma unroll
                for (int nt = 0; nt < 8; nt++) {
                    const v8bf w = lds_w(lds + zq, IW_OFF[layer], 128, q, nt, m, g);
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) acc[nt][bb] = mfma(w, act[q][bb], acc[nt][bb]);
                }
            }
            to_act(acc, act);
        }
        v4f o[4];
        const float4 b4 = *reinterpret_cast<const float4*>(bias + 512 + 4 * g);
#pragma unroll
        for (int bb = 0; bb < 4; bb++) o[bb] = v4f{b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const v8bf w = lds_w(W, IW_OFF[4], 16, q, 0, m, g);
#pragma unroll
            for (int bb = 0; bb < 4; bb++) o[bb] = mfma(w, act[q][bb], o[bb]);
        }
        // D[output 4 g + j][row m]: the three outputs sit in lanes g == 0
        if (g == 0) {
#pragma unroll
            for (int bb = 0; bb < 4; bb++) {
                const int64_t row = row0 + bb * 16 + m;
                if (row < B) {
#pragma unroll
                    for (int j = 0; j < 3; j++) a.Y[(size_t)row * 4 + j] = o[bb][j];
                }
            }
        }
    }
}

} // namespace

namespace cr {

hipError_t launch_mlp_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_probe_kernel, dim3(1), dim3(64), 0, s, A, B, D);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpArgs& a, unsigned blocks, hipStream_t s)
{
    const uint64_t tasks = ((uint64_t)a.B + I_ROWS - 1) / I_ROWS;
    const uint64_t need = (tasks + I_WAVES - 1) / I_WAVES;
    if (need == 0) return hipSuccess;
    if (need < blocks) blocks = (unsigned)need;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(I_WAVES * 64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_train(const MlpArgs& a, hipStream_t s)
{
    const unsigned tiles = mlp_tiles((uint64_t)a.B);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_train_kernel, dim3(tiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_wgrad(const MlpArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(MLP_GRAD_TILES * a.slots), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpArgs& a, const AdamArgs& o, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((o.n + 255) / 256), dim3(256), 0, s, a, o);
    return hipGetLastError();
}

} // namespace cr
