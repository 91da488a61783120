// The context behind the opaque cr_context handle, shared by the runtime
// translation units.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "cr_internal.h"

// Progressive accumulation state (cr_runtime_output.cpp).
struct Progressive {
    uint32_t randState = 0;         // std::srand(0) (main.cpp:949)
    uint32_t firstFrameCount = 1;   // main.cpp:262
    bool clearOnRestart = false;
    bool started = false;
    uint64_t sceneVersion = 0;      // the context's sceneVersion at the last (re)start
    uint32_t frameCount = 1;
    uint32_t framesRendered = 0;
    uint32_t restarts = 0;
};

struct cr_context {
    int device = 0;
    int numCUs = 0;
    std::string arch;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // path tracer / shadow volume launches
    hipEvent_t gev0 = nullptr, gev1 = nullptr; // gather launches
    std::string err;

    // density grid
    cr::DeviceGrid grid;
    int gridLayout = CR_GRID_DENSE;
    bool hasGrid = false;
    uint64_t gridVersion = 0;   // bumped by every grid load or generation
    float gridMax = 0.0f;
    double voxelSize[3] = {0, 0, 0};

    // uniforms (RenderTechnique descriptor bindings)
    cr_camera_properties cam;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_shadow_volume_properties svp;
    cr_render_ext ext;

    // shadow volume
    float* sv = nullptr;
    size_t svCap = 0;
    bool svValid = false;
    uint32_t svDimsBuilt[3] = {0, 0, 0};
    float svMs = 0.0f;

    // accumulation image (RGBA32F)
    float* image = nullptr;
    size_t imageCap = 0;
    int W = 0, H = 0;

    // path tracer
    int32_t* dFrames = nullptr;
    size_t framesCap = 0;
    uint32_t* dQueue = nullptr;
    size_t queueCap = 0;
    unsigned long long* dCounters = nullptr;
    size_t countersCap = 0;
    void* dScalar = nullptr;
    size_t scalarCap = 0;
    bool ptPending = false;
    float ptLastMs = 0.0f;
    // per-launch event pairs, folded into ptTotalMs when stats are read
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ptEvents;
    size_t ptEventsUsed = 0;
    double ptTotalMs = 0.0;
    uint64_t ptLaunches = 0;
    int ptBlocksOverride = 0;
    // cost-ordered tile scheduling: per-tile work of the last render, and the sort buffers
    uint32_t* tileCost = nullptr;
    uint32_t* tileSort = nullptr;   // keys, vals, keys_tmp, vals_tmp (nTiles each) + histograms
    size_t tileSortCap = 0;
    uint32_t tileN = 0, tileLog2W = 0;
    bool tileCostValid = false;
    uint32_t ptTileLog2W = 3;                  // 8x8 pixel tiles

    // bumped by every change to what a render reads (grid, uniforms, shadow volume)
    uint64_t sceneVersion = 0;
    Progressive prog;

    // photon beams / LBVH (cr_runtime_ppb.cpp)
    struct Ppb* ppb = nullptr;
    // photon mapping (cr_runtime_ppm.cpp): grid properties and the photon map
    cr_photon_map_properties pm;
    bool hasPm = false;
    struct Ppm* ppm = nullptr;
    // learned in-scattering network (cr_runtime_mlp.cpp) and its support data (cr_runtime_learn.cpp)
    struct Mlp* mlp = nullptr;
    struct Learn* learn = nullptr;
    // batched render (cr_runtime_batch.cpp)
    struct Batch* batch = nullptr;
};

namespace cr {
cr_status fail(cr_context* ctx, cr_status code, const std::string& msg);
cr_status hip_fail(cr_context* ctx, hipError_t e, const char* what);
cr_status ensure_buffer(cr_context* ctx, void** p, size_t* cap, size_t bytes, const char* what);
hipError_t zero_async(void* p, size_t bytes, hipStream_t s);
cr_status upload_frames(cr_context* ctx, const cr_push_constants* frames, uint32_t n);
int majorant_brick(const cr_render_ext& e);
void fill_job_args(PTArgs& a, const DeviceGrid& g, const float* sv, const uint32_t svDims[3],
                   const cr_camera_properties& cam, const cr_cloud_properties& cloud, const cr_parameters& params,
                   const cr_shadow_volume_properties& svp, const cr_render_ext& ext, const int32_t* frames,
                   uint32_t n_frames, float* image, int W, int H, uint32_t tileLog2W);
cr_status fold_pt_events(cr_context* ctx);
cr_status majorants_build(cr_context* ctx, DeviceGrid& g, int mb);
void ppb_free(cr_context* ctx);
void ppm_free(cr_context* ctx);
void mlp_free(cr_context* ctx);
// Parameters have been set (cr_mlp_init / cr_mlp_set_params).
bool mlp_initialized(cr_context* ctx);
// New network parameters: forget which data set the network was trained on.
void learn_network_reset(cr_context* ctx);
void learn_free(cr_context* ctx);
void batch_free(cr_context* ctx);
// Rows drawn from a device-resident data set for one training step: batch row r
// takes data-set row mlp_draw(seed, step, r, n) (cr_internal_mlp.h).
struct MlpDraw {
    const uint16_t* X;  // [n][224] bf16 features
    const float* T;     // [n][3] targets
    uint64_t n;
    uint32_t seed, step;
};
cr_status mlp_train_device(cr_context* ctx, const uint16_t* X, const float* T, uint64_t B, float lr,
                           const MlpDraw* draw = nullptr);
// Bdev: optional device-resident row count (<= B; B sizes the launch).
cr_status mlp_forward_device(cr_context* ctx, const uint16_t* X, uint64_t B, float* Y, const uint32_t* Bdev = nullptr);
cr_status mlp_last_loss(cr_context* ctx, float* loss);
cr_status fill_pt_args(cr_context* ctx, PTArgs& a, uint32_t n_frames);
// Resolves this device's copy of the allocation-free entry point's scratch array (cr_runtime_static.cpp).
cr_status static_scratch_init(cr_context* ctx);
} // namespace cr
