// Internal interfaces of the photon-mapping kernels (kern_ppm.hip) used by
// cr_runtime_ppm.cpp.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cr/abi.h"

namespace cr {

struct PpmGrid {
    const void* grid;
    int32_t nx, ny, nz, nbx, nby;
};

struct PpmTraceArgs {
    PpmGrid g;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_photon_map_properties pm;
    cr_render_ext ext;          // the sky the sky photons carry
    uint32_t numPaths;
    int32_t seed;
    float sunProbability;
    uint32_t maxCollisions;
    uint32_t* counts;           // count pass: deposits per path
    const uint32_t* offsets;    // write pass: first slot of each path
    cr_photon* photons;         // write pass: deposits in emission order
    uint32_t* cells;            //   their cells
    uint32_t* order;            //   their slots (the sort's values)
};

struct PpmEstimateArgs {
    PpmGrid g;
    const float* sv;
    int32_t svx, svy, svz;
    cr_camera_properties cam;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_shadow_volume_properties svp;
    cr_photon_map_properties pm;
    cr_render_ext ext;
    float ambientStep;
    const uint32_t* arrivals;   // per cell
    const uint32_t* start;      // per cell: first of its run in `photons`
    const cr_photon* photons;   // deposits sorted by cell, emission order within a cell
    const float4* dirs;         // their unit directions from (theta, phi), formed when they are sorted
    const int32_t* frames;      // (seed, frameCount, pmRadius bits) per frame
    uint32_t nFrames;
    float4* image;
    int32_t W, H;
    uint32_t padded;            // 32 * (W/32 + 1), the reference's pixel-id stride
};

hipError_t launch_ppm_trace(const PpmTraceArgs& a, bool write, int dtype, int layout, hipStream_t s);
hipError_t launch_ppm_runs(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t* start, uint32_t* arrivals,
                           const cr_photon* photons, cr_photon* sorted, float4* dirs, hipStream_t s);
hipError_t launch_ppm_estimate(const PpmEstimateArgs& a, int dtype, int layout, hipStream_t s);

} // namespace cr
