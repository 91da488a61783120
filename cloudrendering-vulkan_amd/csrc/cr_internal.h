// Internal interfaces between the host runtime (cr_context.cpp) and the
// kernel translation units.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cr/abi.h"

namespace cr {

// Device grid descriptor shared by the kernels.
struct DeviceGrid {
    void* data = nullptr;       // FP32 or FP16 voxels
    int dtype = CR_DTYPE_F32;
    int layout = 0;             // 0 dense x-fastest, 1 4^3 bricks
    int32_t nx = 0, ny = 0, nz = 0;
    int32_t nbx = 0, nby = 0, nbz = 0;
    size_t bytes = 0;
    float* majorants = nullptr; // ceil(N/mb)^3 brick maxima (with apron)
    float* skip = nullptr;      // per brick: (majorant * ds) / bs, or for an empty brick -(Chebyshev distance to a
                                // non-empty one, <= 16); then the majorant
    int32_t mx = 0, my = 0, mz = 0;
    int32_t mb = 16;            // voxels per side of a majorant brick
};

struct PTArgs {
    // density
    const void* grid;
    int32_t nx, ny, nz, nbx, nby;
    const float* skip;      // (scaled majorant or -(distance to a non-empty brick), majorant) per brick
    int32_t mx, my, mz;
    float bw[3];
    float ibw[3];           // 1 / bw
    // shadow volume
    const float* sv;
    int32_t svx, svy, svz;
    // uniforms
    cr_camera_properties cam;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_shadow_volume_properties svp;
    cr_render_ext ext;
    float sigmaBar;
    // frames
    const int32_t* frames;  // pairs (seed, frameCount)
    uint32_t nFrames;
    // image
    float4* image;
    int32_t W, H;
    uint32_t padded;        // 32 * (W/32 + 1), the reference's id stride
    uint32_t tilesX;
    uint32_t totalIdx;      // tiles * 64
    uint32_t* queue;        // pixel work counter (zeroed before launch)
    unsigned long long* counters; // paths, vertices, lookups, svLookups, brickVisits
    uint32_t tileLog2W;     // a wave's claims walk tiles of (1 << tileLog2W) x (64 >> tileLog2W) pixels
    uint32_t base;          // batched launch: first work index of this job
    // learned-renderer passes
    float* records;         // record pass: [nFrames][H][W] records of 8 floats
    const float* points;    // continuation pass: records to continue
    float4* out;            // continuation pass: mean radiance per record
    uint32_t spp;           // continuation pass: samples per record
    // scheduling (render pass): claim order of the tiles, and per-tile work of this launch
    const uint32_t* tileOrder;  // tile for each queue slot, or null (raster order)
    uint32_t* tileCost;         // lookups + brick visits per tile, accumulated, or null
};

// One persistent launch over several jobs: the job table (PTArgs per job,
// queue and counters unused) is copied to LDS by every workgroup; work index
// i belongs to the last job whose base is <= i.
struct BatchArgs {
    const PTArgs* jobs;
    uint32_t nJobs;
    uint32_t total;         // sum of the jobs' totalIdx
    uint32_t* queue;
    unsigned long long* counters;
};

// kern_pt.hip: a render of a.nFrames frames; launches of at most
// PT_FEW_FRAMES frames use the instantiation with gated pixel claims
// (kern_pt_few.hip), which is faster when each pixel carries little work.
constexpr uint32_t PT_FEW_FRAMES = 16;
hipError_t launch_pt(const PTArgs& a, int dtype, int layout, int grid_blocks, hipStream_t s);
int pt_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee, uint32_t n_frames);
// kern_pt_few.hip
hipError_t launch_pt_few(const PTArgs& a, int dtype, int layout, int grid_blocks, hipStream_t s);
int pt_few_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee);
hipError_t launch_pt_batch(const BatchArgs& b, int dtype, int layout, uint32_t tracking, uint32_t nee, int grid_blocks,
                           int block_threads, hipStream_t s);
int pt_batch_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee, int block_threads, uint32_t n_jobs);
// Tile claim order for the next render: tiles by decreasing cost (a stable
// radix sort of ~cost; ties keep raster order).  keys/vals/tmp: n words each,
// counts: the radix sort's histograms (pt_order_counts(n) bytes).
uint32_t pt_order_counts(uint32_t n);
hipError_t launch_tile_order(const uint32_t* cost, uint32_t n, uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp,
                             uint32_t* vals_tmp, uint32_t* counts, hipStream_t s);
// kern_pt_learn.hip
hipError_t launch_pt_record(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s);
hipError_t launch_pt_continue(const PTArgs& a, int dtype, int layout, int blocks, hipStream_t s);
int pt_learn_blocks_per_cu(int dtype, int layout, uint32_t tracking, uint32_t nee, int continuation);

struct SVArgs {
    const void* grid;
    int32_t nx, ny, nz, nbx, nby;
    cr_cloud_properties cloud;
    cr_shadow_volume_properties svp;
    float* out;
    int32_t dx, dy, dz;
    float k[3];
};
// kern_sv.hip
hipError_t launch_sv(const SVArgs& a, int dtype, int layout, hipStream_t s);

// kern_grid.hip
hipError_t launch_fbm(const cr_fbm_params& p, void* out, int dtype, int layout, int32_t nbx, int32_t nby,
                      hipStream_t s);
hipError_t launch_convert_grid(const void* src, int src_dtype, int src_layout_zfast, void* dst, int dst_dtype,
                               int dst_layout, int32_t nx, int32_t ny, int32_t nz, int32_t nbx, int32_t nby,
                               hipStream_t s);
hipError_t launch_grid_to_f32(const void* src, int dtype, int layout, float* dst, int32_t nx, int32_t ny, int32_t nz,
                              int32_t nbx, int32_t nby, hipStream_t s);
hipError_t launch_majorants(const void* grid, int dtype, int layout, int32_t nx, int32_t ny, int32_t nz, int32_t nbx,
                            int32_t nby, int32_t mb, float* out, int32_t mx, int32_t my, int32_t mz, hipStream_t s);
hipError_t launch_math_probe(int op, const float* in, uint32_t n, float* out, hipStream_t s);
// Skip grid from the majorants (tmp: 2 * n bytes of scratch); ds / bs are the
// cloud's densityScaling and baseScaling.  launch_skip_scale redoes only the
// non-empty bricks' values after those change.
hipError_t launch_skip_grid(const float* majorants, int32_t mx, int32_t my, int32_t mz, float ds, float bs, float* skip,
                            uint8_t* tmp, hipStream_t s);
hipError_t launch_skip_scale(const float* majorants, uint32_t n, float ds, float bs, float* skip, hipStream_t s);
hipError_t launch_grid_max(const float* majorants, uint32_t n, unsigned int* out_bits, hipStream_t s);

} // namespace cr
