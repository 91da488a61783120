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

struct MlpArgs {
    const uint16_t* X;       // [B][224] bf16 features
    int64_t B;
    const uint32_t* Bdev;    // forward only: when set, the row count is *Bdev (<= B, B sizes the launch)
    uint16_t* WT[MLP_LAYERS];  // bf16 [N][K] (forward B fragments)
    uint16_t* W[MLP_LAYERS];   // bf16 [K][N] (backward B fragments)
    const float* bias[MLP_LAYERS];
    float* Y;                // [B][4] fp32 outputs (log(1 + radiance), 3 used)
    uint16_t* acts;          // [4][B][128] bf16 hidden activations (training)
    size_t actStride;        // elements between layers in acts
    const float* dY;         // [B][4] fp32
    float* grad;             // flat fp32 gradients (same layout as the parameters), written by the Adam kernel
    float* partial;          // [nWaves][total] per-wave gradient sums of the backward pass
    uint32_t nWaves;         // backward waves: wave w takes the 32-row chunks w, w + nWaves, ...
    uint32_t total;          // flat parameter count
    uint32_t wOff[MLP_LAYERS], bOff[MLP_LAYERS];
};

// Backward waves for a batch of B rows: one per 32-row chunk, at most MLP_MAX_WAVES.
constexpr uint32_t MLP_MAX_WAVES = 1024;
inline uint32_t mlp_waves(uint64_t B)
{
    const uint64_t c = (B + 31) / 32;
    return (uint32_t)(c < MLP_MAX_WAVES ? c : MLP_MAX_WAVES);
}
// Loss kernel blocks (one partial loss each) for a batch of B rows.
inline unsigned mlp_loss_blocks(int64_t B) { return (unsigned)((B + 255) / 256); }

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
hipError_t launch_mlp_forward(const MlpArgs& a, bool train, hipStream_t s);
// loss: mlp_loss_blocks(B) partial sums, to be added in order
hipError_t launch_mlp_loss(const float* Y, const float* T, int64_t B, float* dY, float* loss, hipStream_t s);
hipError_t launch_mlp_backward(const MlpArgs& a, hipStream_t s);
hipError_t launch_mlp_adam(const MlpArgs& a, const AdamArgs& o, hipStream_t s);

} // namespace cr
