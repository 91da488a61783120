// Internal interfaces of the learned-renderer support kernels (kern_learn.hip).
// Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cr/abi.h"

namespace cr {

constexpr int LEARN_LEVELS = 8;

struct FeatureArgs {
    const float* points;    // [n][8]: x, y, z, omega_x, omega_y, omega_z, -, -
    uint64_t n;
    const float* mips[LEARN_LEVELS];
    int32_t dims[LEARN_LEVELS][3];
    cr_cloud_properties cloud;
    float light[3];         // propagation direction of sunlight
    float h0;               // world size of a level-0 voxel
    uint16_t* X;            // [n][224] bf16
};

struct ComposeArgs {
    float4* image;
    uint32_t pixels;
    uint32_t nFrames;
    const int32_t* frames;  // (seed, frameCount) pairs
    const float4* sampleRes;// [frames][pixels]: background radiance, w = 1 when the ray collided
    const float* Y;         // [frames][pixels][4] network outputs
};

hipError_t launch_mip(const float* src, int32_t sx, int32_t sy, int32_t sz, float* dst, int32_t dx, int32_t dy, int32_t dz,
                      hipStream_t s);
hipError_t launch_features(const FeatureArgs& a, hipStream_t s);
hipError_t launch_compose(const ComposeArgs& a, hipStream_t s);

} // namespace cr
