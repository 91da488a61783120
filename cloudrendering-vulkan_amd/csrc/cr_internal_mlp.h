// Internal interfaces of the learned in-scattering kernels (kern_mlp.hip,
// kern_learn.hip) used by cr_runtime_mlp.cpp.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cr/abi.h"

namespace cr {

// Padded layer shapes: K = 224 | 128, N = 128 | 16 (the 3 outputs padded).
constexpr int MLP_LAYERS = 5;
constexpr uint32_t MLP_K[MLP_LAYERS] = {224, 128, 128, 128, 128};
constexpr uint32_t MLP_N[MLP_LAYERS] = {128, 128, 128, 128, 16};

// Rows per training tile (one workgroup of the fused training kernel).
constexpr uint32_t MLP_TILE = 32;
inline uint64_t mlp_padded_rows(uint64_t B) { return (B + MLP_TILE - 1) / MLP_TILE * MLP_TILE; }
inline unsigned mlp_tiles(uint64_t B) { return (unsigned)((B + MLP_TILE - 1) / MLP_TILE); }

// The training kernel keeps, for the weight-gradient GEMM, every layer's input
// and every layer's output gradient feature-major ([feature][Bp] bf16, Bp the
// padded row count): X^T (224 rows), H1..H4^T (128 each), dZ0..dZ3^T (128
// each), dZ4^T (16).  Offsets in elements, in units of Bp.
constexpr uint32_t MLP_TR_IN[MLP_LAYERS] = {0, 224, 352, 480, 608};
constexpr uint32_t MLP_TR_DZ[MLP_LAYERS] = {736, 864, 992, 1120, 1248};
constexpr uint32_t MLP_TR_ROWS = 1264;

// Gradient slots of the weight-gradient GEMM: the batch's 32-row chunks are
// split into `slots` contiguous ranges, each summed into a slot of its own and
// the slots added in order by the Adam kernel (no float atomics).
constexpr uint32_t MLP_MAX_SLOTS = 16;
inline uint32_t mlp_slots(uint64_t B)
{
    const uint64_t s = (mlp_tiles(B) + 31) / 32;
    return (uint32_t)(s < 1 ? 1 : (s > MLP_MAX_SLOTS ? MLP_MAX_SLOTS : s));
}
// Output tiles of the weight-gradient GEMM (32 x 32, 32 x 16 for the last layer).
constexpr uint32_t MLP_GRAD_TILES = 7 * 4 + 3 * 4 * 4 + 4;

struct MlpArgs {
    // batch source: X [B][224] bf16 and T [B][3], or, when dsX is set, rows
    // drawn from the data set (dsX [dsN][224], dsT [dsN][3]) by (seed, step)
    const uint16_t* X;
    const float* T;
    const uint16_t* dsX;
    const float* dsT;
    uint64_t dsN;
    uint32_t seed, step;
    int64_t B;
    uint64_t Bp;             // padded row count (a multiple of 32)
    const uint32_t* Bdev;    // inference only: when set, the row count is *Bdev (<= B, B sizes the launch)
    uint16_t* WT[MLP_LAYERS];  // bf16 [N][K]
    uint16_t* W[MLP_LAYERS];   // bf16 [K][N]
    const float* bias[MLP_LAYERS];
    float* Y;                // inference: [B][4] fp32 outputs (log(1 + radiance), 3 used)
    uint16_t* tr;            // training: feature-major inputs and output gradients (MLP_TR_ROWS x Bp)
    float* loss;             // training: per-tile loss shares, added in order by the host
    float* grad;             // flat fp32 gradients (same layout as the parameters), written by the Adam kernel
    float* slot;             // [slots][total] gradient sums of the weight-gradient GEMM
    uint32_t slots;
    uint32_t total;          // flat parameter count
    uint32_t wOff[MLP_LAYERS], bOff[MLP_LAYERS];
};

struct AdamArgs {
    float* params;
    float* m;
    float* v;
    uint32_t n;
    float lr, beta1, beta2, eps, c1, c2;
    // moving average (null ema: off): ema = d ema + (1 - d) params; emaHat = ema * emaCorr
    float* ema;
    float* emaHat;
    uint16_t* EW[MLP_LAYERS];   // bf16 copies of emaHat, [K][N] and [N][K]
    uint16_t* EWT[MLP_LAYERS];
    float emaDecay, emaCorr;
};

hipError_t launch_mlp_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s);
// Inference: persistent, `blocks` workgroups (one per CU).
hipError_t launch_mlp_infer(const MlpArgs& a, unsigned blocks, hipStream_t s);
// Training: the fused forward / loss / backward chain (mlp_tiles(B) loss
// shares), then the weight-gradient GEMM into a.slots slots, then Adam.
hipError_t launch_mlp_train(const MlpArgs& a, hipStream_t s);
hipError_t launch_mlp_wgrad(const MlpArgs& a, hipStream_t s);
hipError_t launch_mlp_adam(const MlpArgs& a, const AdamArgs& o, hipStream_t s);

// The data-set row a training step draws for batch row r (shared by the
// training kernel and the host-side description in cr_learn_train).
__host__ __device__ inline uint64_t mlp_draw(uint32_t seed, uint32_t step, uint32_t r, uint64_t n)
{
    uint32_t state = (seed * 0x9E3779B9u) ^ (step * 0x85EBCA6Bu) ^ (r * 0xC2B2AE35u);
    state = state * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    word = (word >> 22u) ^ word;
    return (uint64_t)word % n;
}

} // namespace cr
