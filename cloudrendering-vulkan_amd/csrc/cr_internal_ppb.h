// Internal interfaces of the photon-beam kernels (kern_lbvh.hip,
// kern_ppb.hip) used by cr_runtime_ppb.cpp.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cr/abi.h"

namespace cr {

struct BoxArg {
    float b0[3];
    float b1[3];
};

// One internal node in traversal form: both child boxes and indices.
// Center and half extent of each child's box, as intersectBox
// (PPB_PE.comp:282-300) forms them from the bounds, computed once at build.
struct TravNode {
    float cen[2][3];
    float rad[2][3];
    uint32_t child[2];
    uint32_t leafMask;  // bit c set: child c is a leaf
    uint32_t pad;
};
static_assert(sizeof(TravNode) == 64, "TravNode is one 64-B line");

// A beam in LBVH leaf order: the fields of cr_photon_beam the gather reads
// (its offset along the parent beam is kept in a separate array).
struct SortedBeam {
    float startPos[3];
    float radius;
    float endPos[3];
    uint32_t dataIdx;
};
static_assert(sizeof(SortedBeam) == 32, "SortedBeam is 32 B");

hipError_t launch_morton(const cr_photon_beam* beams, uint32_t n, const BoxArg& box, uint32_t* keys, uint32_t* vals,
                         hipStream_t s);
uint32_t sort_tiles(uint32_t n);
// words of the counts buffer launch_radix_sort needs for n keys (histograms and digit totals)
uint32_t sort_counts_words(uint32_t n);
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t* counts,
                             uint32_t n, hipStream_t s);
// meta[i] = (left, right, end, parent) of internal node i (end: the other end j
// of its leaf range [min(i, j), max(i, j)]); leafParent[l]: the parent of leaf l
hipError_t launch_hierarchy(const uint32_t* keys, uint32_t n, uint4* meta, uint32_t* leafParent, hipStream_t s);
hipError_t launch_leaves(const cr_photon_beam* beams, const uint32_t* perm, uint32_t n, SortedBeam* sorted, float* sortedOff,
                         float* lo, float* hi, hipStream_t s);
// exits: refit_exit_words(n) words, the last one and arrivals (n - 1
// counters) zeroed beforehand
uint64_t refit_exit_words(uint32_t n);
hipError_t launch_refit(uint32_t n, const uint4* meta, const uint32_t* leafParent, float* lo, float* hi, uint32_t* arrivals,
                        uint32_t* exits, hipStream_t s);
hipError_t launch_pack(uint32_t n, const uint4* meta, const float* lo, const float* hi, TravNode* out, hipStream_t s);
hipError_t launch_treenodes(uint32_t n, const uint4* meta, const uint32_t* leafParent, const float* lo, const float* hi,
                            cr_tree_node* out, hipStream_t s);

// Beam producer: the scene a photon path sees.
struct TraceArgs {
    const void* grid;
    int32_t nx, ny, nz, nbx, nby;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_beam_trace_params p;
    // write pass only
    const uint64_t* segOffsets;
    const uint64_t* dataOffsets;
    cr_photon_beam* beams;
    cr_photon_beam_data* data;
    uint64_t maxBeams;
    uint64_t maxData;
    // count pass only
    uint32_t* segCounts;
    uint32_t* dataCounts;
};
hipError_t launch_trace_count(const TraceArgs& a, int dtype, int layout, hipStream_t s);
hipError_t launch_trace_write(const TraceArgs& a, int dtype, int layout, hipStream_t s);

struct GatherArgs {
    const void* grid;
    int32_t nx, ny, nz, nbx, nby;
    const float* sv;
    int32_t svx, svy, svz;
    cr_camera_properties cam;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_shadow_volume_properties svp;
    const SortedBeam* beams;         // in leaf order
    const float* beamOff;            //   their offsets along their parent beams
    const cr_photon_beam_data* data;
    const TravNode* nodes;
    const float* leafLo;             // box of node 0 when n == 1
    const float* leafHi;
    uint32_t n;
    uint32_t refCompat, jitter;
    float albedo;
    const int32_t* frames;
    uint32_t nFrames;
    float4* image;
    int32_t W, H;
    uint32_t padded;
    unsigned long long* counters;    // rays, nodesVisited, beamsTested
    uint32_t* overflow;              // set when a traversal stack would overflow
};
hipError_t launch_gather(const GatherArgs& a, int dtype, int layout, hipStream_t s);

} // namespace cr
