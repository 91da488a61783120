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
    const uint32_t* ndev;   // when set, the point count is *ndev (<= n, n sizes the launch)
    const float* mips[LEARN_LEVELS];
    int32_t dims[LEARN_LEVELS][3];
    cr_cloud_properties cloud;
    float light[3];         // propagation direction of sunlight
    float h0;               // world size of a level-0 voxel
    uint16_t* X;            // [n][224] bf16
};

// The next-event term the path tracer adds at a path's first collision with
// shadow-volume NEE: (((1 * I) * pdf) * Tv) * albedo, pdf the phase function
// toward the sun and Tv the shadow volume at the collision.  It is a function
// of the recorded position and direction alone, so the learned renderer
// evaluates it exactly (the data set's targets leave it out and the network
// predicts only the rest).  on = 0: ratio-tracked or no NEE, or bounce cap 0.
struct FirstNee {
    uint32_t on;
    const float* sv;
    int32_t svx, svy, svz;
    cr_shadow_volume_properties svp;
    cr_parameters params;
    float intensity;
    float albedo;
};

struct ComposeArgs {
    float4* image;
    uint32_t pixels;
    uint32_t nFrames;
    const int32_t* frames;  // (seed, frameCount) pairs
    const float* records;   // [frames][pixels] records of the record pass (8 floats)
    const uint32_t* slot;   // [frames][pixels]: row of Y, or ~0 for an escaped path
    const float* Y;         // [rows][4] network outputs
    FirstNee nee;           // added to every collided record's prediction
};

hipError_t launch_mip(const float* src, int32_t sx, int32_t sy, int32_t sz, float* dst, int32_t dx, int32_t dy, int32_t dz,
                      hipStream_t s);
hipError_t launch_features(const FeatureArgs& a, hipStream_t s);
hipError_t launch_compose(const ComposeArgs& a, hipStream_t s);
hipError_t launch_count(const float* rec, uint64_t n, uint32_t* counts, hipStream_t s);
hipError_t launch_compact(const float* rec, uint64_t n, const uint32_t* offsets, float* pts, uint32_t* slot, hipStream_t s);
hipError_t launch_scan_counts(uint32_t* counts, uint32_t nb, uint32_t* total, hipStream_t s);
hipError_t launch_select(const float* src, uint32_t stride, uint32_t count, float* dst, hipStream_t s);
// Append rows k * stride of X / T4 (and of the records pts) to the data set;
// the targets leave out nee's term of each record.
hipError_t launch_append(const uint16_t* X, const float4* T4, const float* pts, const FirstNee& nee, uint32_t stride,
                         uint32_t count, uint16_t* dsX, float* dsT, uint64_t offset, hipStream_t s);

} // namespace cr
