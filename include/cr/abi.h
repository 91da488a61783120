/*
 * C-ABI of the MI355X (gfx950) volumetric hot path.
 *
 * The reference drives its GPU work through the RenderTechnique plugin API
 * (CloudRendering-Vulkan/RenderTechnique.h:15-66): QueueUpdate* binds the
 * uniform buffers and images, UpdateDescriptorSets commits them and
 * RecordDrawCommands records the dispatches.  This header is the same
 * contract without Vulkan: plain extern "C" functions, plain pointers and
 * sizes, status codes instead of abort()/exceptions, one opaque context per
 * host thread (a context is not thread-safe; contexts are independent).
 *
 * Ownership: the caller owns every host buffer it passes; the library owns
 * the device buffers (density grid, shadow volume, accumulation image, beams,
 * tree, network weights) and keeps them resident in HBM between calls.
 * Device work is issued on the context's own HIP stream; functions that hand
 * results back to host memory synchronise that stream first.
 *
 * Everything marked "extension" has no counterpart in the reference; with
 * the defaults (cr_render_ext_default) every path reduces to the reference's
 * behaviour.
 */
#ifndef CR_ABI_H
#define CR_ABI_H

#include "uniforms.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define CR_API __declspec(dllexport)
#else
#define CR_API __attribute__((visibility("default")))
#endif

/* ---------------------------------------------------------------- status */
typedef int32_t cr_status;
enum {
    CR_OK = 0,
    CR_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size, value out of range */
    CR_ERR_OUT_OF_MEMORY = 2,
    CR_ERR_DEVICE = 3,           /* HIP runtime error (message in cr_last_error) */
    CR_ERR_CAPACITY = 4,         /* output would exceed a stated capacity */
    CR_ERR_IO = 5,               /* file missing, short or malformed */
    CR_ERR_STATE = 6,            /* a prerequisite was not set (no grid, no tree, ...) */
    CR_ERR_NOT_SUPPORTED = 7
};

#define CR_UNBOUNDED 0xFFFFFFFFu

typedef struct cr_context cr_context;

/* Human-readable name of a status code (static storage). */
CR_API const char* cr_status_string(cr_status status);
/* Last error message recorded on this context (valid until the next call). */
CR_API const char* cr_last_error(const cr_context* ctx);
/* Library version string, e.g. "cr-mi355x 0.1 gfx950". */
CR_API const char* cr_version(void);

/* ---------------------------------------------------------------- context
 * Replaces the reference's device/technique construction
 * (main.cpp:765-910 InitializeVulkan, RenderTechniquePT.cpp:11-50). */
CR_API cr_status cr_create(int device, cr_context** out_ctx);
CR_API cr_status cr_destroy(cr_context* ctx);
/* Wait for all work issued on the context's stream. */
CR_API cr_status cr_synchronize(cr_context* ctx);
/* The context's hipStream_t, for callers that time or order work around it. */
CR_API void* cr_stream(cr_context* ctx);

/* ------------------------------------------------------- host-side setters
 * Restatements of the reference's struct setters, so a caller can build the
 * uniform structs without glm.  They touch host memory only. */
/* UniformBuffers.h:15-27 default members. */
CR_API void cr_camera_default(cr_camera_properties* cam);
/* UniformBuffers.h:46-50 CameraProperties::SetResolution. */
CR_API void cr_camera_set_resolution(cr_camera_properties* cam, int width, int height);
/* UniformBuffers.h:52-58 CameraProperties::SetFOV; *fov is clamped to [60,150] in place. */
CR_API void cr_camera_set_fov(cr_camera_properties* cam, float* fov_degrees);
/* UniformBuffers.h:29-35 CameraProperties::SetRotation; rotation[0] clamped to [-80,80]. */
CR_API void cr_camera_set_rotation(cr_camera_properties* cam, float rotation_degrees[2]);
/* UniformBuffers.h:80-91 default Parameters. */
CR_API void cr_parameters_default(cr_parameters* p);
/* UniformBuffers.h:93-106 Parameters::SetPhaseG (ignores |g|>1, clamps to +-0.95). */
CR_API void cr_parameters_set_phase_g(cr_parameters* p, float g);
/* UniformBuffers.h:71-78 default CloudProperties. */
CR_API void cr_cloud_default(cr_cloud_properties* c);
/* UniformBuffers.h:114-127 default ShadowVolumeProperties. */
CR_API void cr_shadow_volume_default(cr_shadow_volume_properties* sv);
/* UniformBuffers.h:131-143 SetLightDirection then UniformBuffers.h:145-151 SetOrigin,
 * in the order main.cpp:223-236 calls them (SetOrigin re-centres the
 * volume that SetLightDirection placed off centre). */
CR_API void cr_shadow_volume_set_light(cr_shadow_volume_properties* sv,
                                       const float light_dir[3],
                                       const float lower_bound[3],
                                       const float upper_bound[3]);
/* MSVC rand() after srand(seed): the per-frame seed sequence of main.cpp:722/:949. */
CR_API void cr_msvc_rand_sequence(uint32_t srand_seed, uint32_t count, int32_t* out);

/* ------------------------------------------------------------------- grid
 * Grid3D / .xyz (Grid3D.h:44-192, GridData.h:1-25, main.cpp:144-190). */
typedef struct cr_grid_header {
    uint32_t sizeX, sizeY, sizeZ;
    uint32_t _pad;
    double voxelSizeX, voxelSizeY, voxelSizeZ;
} cr_grid_header;

enum { CR_DTYPE_F32 = 0, CR_DTYPE_F16 = 1 };
enum { CR_LAYOUT_X_FASTEST = 0 /* Grid3D memory order */, CR_LAYOUT_Z_FASTEST = 1 /* .xyz file order */ };

/* Upload a density grid.  `data` holds sizeX*sizeY*sizeZ values of
 * `src_dtype` in `layout`; it is stored on the device as `storage_dtype`
 * (FP16 storage is an extension; values convert round-to-nearest-even).
 * Applies SetCloudProperties (main.cpp:144-163) with the context's current
 * baseScaling: bounds from voxel sizes, voxelCount, maxExtinction =
 * max(grid maximum, 0.01).  Also builds the brick-majorant grid (bricks of
 * cr_render_ext.majorantBrick voxels per side, 16 by default). */
CR_API cr_status cr_grid_upload(cr_context* ctx, const cr_grid_header* hdr, const void* data,
                                int src_dtype, int layout, int storage_dtype);
/* Device layout used by the next upload or generation (extension):
 * CR_GRID_DENSE = x-fastest as Grid3D memory; CR_GRID_BRICK4 = 4x4x4-voxel
 * bricks (128 B per FP16 brick) stored brick-major, so a trilinear footprint
 * usually falls in one or two cache lines.  Results do not depend on it. */
enum { CR_GRID_DENSE = 0, CR_GRID_BRICK4 = 1 };
CR_API cr_status cr_set_grid_layout(cr_context* ctx, int layout);
/* Read an .xyz file (Grid3D<float>::Load semantics: z-fastest file, x-fastest memory). */
CR_API cr_status cr_grid_load_xyz(cr_context* ctx, const char* path, int storage_dtype);
/* Write the current grid as .xyz in file order (z fastest), so load(save(g)) == g. */
CR_API cr_status cr_grid_save_xyz(cr_context* ctx, const char* path);
/* Copy the device grid back as FP32, x fastest (sizeX*sizeY*sizeZ floats). */
CR_API cr_status cr_grid_download(cr_context* ctx, float* out);
/* Brick-majorant grid (extension): ceil(N/b) per axis for b-voxel bricks, x
 * fastest, max density over each b^3 brick plus a one-voxel apron. */
CR_API cr_status cr_grid_majorants(cr_context* ctx, uint32_t dims_out[3], float* out /* may be NULL */);
/* The empty-space skip grid the brick walk reads (extension): for a non-empty
 * brick its extinction majorant (majorant * densityScaling) / baseScaling, for
 * an empty brick -d, d being the max-norm distance in bricks to the nearest
 * non-empty brick (d <= 16; 16 = at least 16).  The walk itself uses only the
 * sign (an empty brick is crossed without a lookup). */
CR_API cr_status cr_grid_skip(cr_context* ctx, float* out);

/* Seeded value-noise fBm cloud generated on the device (extension; the
 * reference ships no clouds, models/PUT_YOUR_MODELS_HERE.txt). */
typedef struct cr_fbm_params {
    uint32_t size[3];
    float voxelSize;       /* world units per voxel (isotropic) */
    uint32_t seed;
    uint32_t octaves;
    float lacunarity;
    float gain;
    float baseFrequency;   /* lattice cells per world unit at octave 0 */
    uint32_t shape;        /* CR_SHAPE_* */
    float thresholdLo;     /* smoothstep edges applied to fbm * shape */
    float thresholdHi;
} cr_fbm_params;
enum { CR_SHAPE_CUMULUS = 0, CR_SHAPE_ANVIL = 1 };
CR_API void cr_fbm_params_default(cr_fbm_params* p);
CR_API cr_status cr_grid_generate_fbm(cr_context* ctx, const cr_fbm_params* p, int storage_dtype);

/* --------------------------------------------------------------- uniforms
 * RenderTechnique::QueueUpdate{CameraProperties,CloudData,Parameters,
 * ShadowVolume} + UpdateDescriptorSets (RenderTechnique.h:41-46). */
CR_API cr_status cr_set_camera(cr_context* ctx, const cr_camera_properties* cam);
/* baseScaling and maxExtinction must be > 0, densityScaling finite and >= 0. */
CR_API cr_status cr_set_cloud(cr_context* ctx, const cr_cloud_properties* cloud);
CR_API cr_status cr_get_cloud(cr_context* ctx, cr_cloud_properties* cloud);
CR_API cr_status cr_set_parameters(cr_context* ctx, const cr_parameters* params);
CR_API cr_status cr_set_shadow_volume(cr_context* ctx, const cr_shadow_volume_properties* sv);

/* Render extensions (none exist in the reference). */
typedef struct cr_render_ext {
    float albedo;          /* scattering albedo; 1 = the reference (which has none) */
    uint32_t bounceCap;    /* real collisions that scatter; CR_UNBOUNDED = reference */
    uint32_t skyMode;      /* CR_SKY_* */
    float skyColor[4];     /* radiance of the constant sky */
    uint32_t neeMode;      /* CR_NEE_* */
    uint32_t trackingMode; /* CR_TRACK_* */
    uint32_t svDims[3];    /* shadow-volume texels per axis; 0,0,0 = voxelAxisCount^3 */
    uint32_t majorantBrick; /* voxels per side of a majorant brick: 4, 8, 16 or 32; 0 = 16 */
    uint32_t _reserved[3];
} cr_render_ext;
enum { CR_SKY_GRADIENT = 0 /* PathTracer.comp:104-113 */, CR_SKY_CONSTANT = 1 };
enum { CR_NEE_SHADOW_VOLUME = 0 /* PathTracer.comp:405-411 */, CR_NEE_RATIO_TRACKING = 1, CR_NEE_NONE = 2 };
enum { CR_TRACK_GLOBAL_MAJORANT = 0 /* PathTracer.comp:335-355 */, CR_TRACK_BRICK_DDA = 1 };
CR_API void cr_render_ext_default(cr_render_ext* ext);
CR_API cr_status cr_set_render_ext(cr_context* ctx, const cr_render_ext* ext);

/* ----------------------------------------------------------- shadow volume
 * RenderTechniqueSV::RecordDrawCommands (RenderTechniqueSV.cpp:102-113) running
 * shaders/ShadowVolume.comp: light-space transmittance volume. */
CR_API cr_status cr_sv_build(cr_context* ctx);
CR_API cr_status cr_sv_dims(cr_context* ctx, uint32_t dims_out[3]);
CR_API cr_status cr_sv_download(cr_context* ctx, float* out /* dims x*y*z, x fastest */);

/* ------------------------------------------------------------ path tracer
 * RenderTechniquePT::RecordDrawCommands (RenderTechniquePT.cpp:129-168) running
 * shaders/PathTracer.comp, with the blit/present dropped.  The accumulation
 * image is RGBA32F, width x height = 2*halfWidth x 2*halfHeight of the camera,
 * row-major from the top-left pixel (the reference's image coordinates).
 * One call renders n_frames frames in one persistent launch; frame f uses
 * frames[f].seed and blends with the running mean
 *     acc = (L + acc * frameCount) / (frameCount + 1)       (PathTracer.comp:420-424)
 * exactly as n_frames consecutive reference frames would. */
CR_API cr_status cr_pt_clear(cr_context* ctx);
CR_API cr_status cr_pt_render(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames);
CR_API cr_status cr_pt_download(cr_context* ctx, float* rgba);
CR_API cr_status cr_pt_upload(cr_context* ctx, const float* rgba);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they render (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): cr_pt_render's frames rendered
 * into a caller-owned device image `accum` (W x H RGBA32F of the current
 * camera, read and written as the running mean) with no allocation.  The
 * render's scratch -- work-queue counter, statistics counters and the
 * per-frame seed table -- is one fixed CR_PT_STATIC_SCRATCH_BYTES array per
 * GPU, resolved at cr_create; at most CR_PT_STATIC_MAX_FRAMES frames per call
 * (CR_ERR_CAPACITY beyond).  Tiles are claimed in raster order instead of by
 * the previous render's cost; every pixel equals cr_pt_render's bit for bit.
 * cr_pt_stats does not see these launches.  The array is shared by every
 * context on a GPU: calls that overlap on one GPU race.  Blocks until the
 * frame table is copied, then returns with the render queued on the
 * context's stream (cr_synchronize waits for it). */
#define CR_PT_STATIC_MAX_FRAMES 4096u
#define CR_PT_STATIC_SCRATCH_BYTES (256u + 8u * CR_PT_STATIC_MAX_FRAMES)
CR_API cr_status cr_pt_render_static(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames,
                                     float* accum);

/* Progressive accumulation: the reference's render loop, headless.  Per frame
 * the loop draws seed = std::rand() (after std::srand(0), main.cpp:722, :949)
 * and increments frameCount (main.cpp:728); every parameter, camera or cloud
 * change goes through UpdateShadowVolume, which restarts frameCount at 1
 * without clearing the image (main.cpp:262, :560-590).  cr_pt_progressive
 * renders the loop's next n_frames frames; when anything a render reads has
 * changed since its previous call (a cr_set_* call, a grid load or
 * generation, cr_sv_build, cr_batch_add) it first restarts frameCount at
 * first_frame_count, clearing the image if clear_on_restart.  The defaults
 * (srand seed 0, first_frame_count 1, no clear) reproduce the reference;
 * first_frame_count 0 with clear_on_restart gives the exact mean of the
 * frames since the last change.  cr_pt_progressive_reset sets them and makes
 * the next call start afresh. */
typedef struct cr_progressive_state {
    uint32_t frameCount;      /* frameCount of the next frame */
    uint32_t framesRendered;  /* frames since the last (re)start */
    uint32_t restarts;        /* restarts caused by changes since cr_pt_progressive_reset */
    uint32_t randState;       /* the std::rand state after the last frame's seed */
} cr_progressive_state;
CR_API cr_status cr_pt_progressive_reset(cr_context* ctx, uint32_t srand_seed, uint32_t first_frame_count,
                                         int clear_on_restart);
CR_API cr_status cr_pt_progressive(cr_context* ctx, uint32_t n_frames, cr_progressive_state* state_out);
/* The accumulation image as a portable float map (RGB float32, little-endian,
 * bottom row first): the headless stand-in for the swapchain blit of
 * RenderTechniquePT::RecordDrawCommands (RenderTechniquePT.cpp:143-167). */
CR_API cr_status cr_pt_write_pfm(cr_context* ctx, const char* path);

typedef struct cr_pt_stats {
    uint64_t paths;      /* camera paths traced (one per pixel per frame) */
    uint64_t vertices;   /* camera-ray origins + accepted (real) collisions */
    uint64_t lookups;    /* density lookups (primary tracking + shadow rays) */
    uint64_t svLookups;  /* shadow-volume lookups */
    uint64_t brickVisits;/* brick-majorant cells entered (DDA tracking) */
    double lastKernelMs; /* duration of the last render launch (HIP events) */
    double totalKernelMs;/* sum of render-launch durations since the last reset */
    uint64_t launches;   /* render launches since the last reset */
} cr_pt_stats;
CR_API cr_status cr_pt_stats_get(cr_context* ctx, cr_pt_stats* out);
CR_API cr_status cr_pt_stats_reset(cr_context* ctx);

/* ---------------------------------------------------- batched render (extension)
 * Many independent path-tracing jobs, each a complete RenderTechniquePT set-up
 * (grid, uniforms, shadow volume, image, frame list), rendered by ONE
 * persistent launch per call: one work queue runs over all jobs' pixels, so
 * small jobs do not leave the GPU idle.  The launch is recorded once into a
 * HIP graph and replayed on later calls; frame lists may change between
 * calls.  Every pixel of a job equals cr_pt_render of the same set-up bit for
 * bit.  No reference counterpart: the reference renders one technique per
 * frame (main.cpp:603-735, RenderLoop/DrawFrame). */
#define CR_BATCH_MAX_JOBS 64
typedef struct cr_batch_tuning {
    uint32_t blockThreads; /* workgroup size of the persistent grid: 64, 128 or 256 */
    uint32_t blocksPerCU;  /* workgroups per CU; 0 = as many as fit */
    uint32_t tileWidth;    /* width of the 64-pixel tiles a wave's claims walk: 8, 16 or 32 */
    uint32_t _reserved;
} cr_batch_tuning;
typedef struct cr_batch_tune_report {
    float defaultMs;       /* the trial render (1/32 of every job's frames) with the default tuning */
    float bestMs;          /* the same with the chosen tuning */
    uint32_t candidates;   /* launch shapes timed */
    uint32_t jobBrick[CR_BATCH_MAX_JOBS]; /* chosen majorant brick per job (insertion order) */
} cr_batch_tune_report;
CR_API void cr_batch_tuning_default(cr_batch_tuning* t);
/* Moves the context's grid (with its majorants) and shadow volume into a new
 * job, with copies of the current camera, cloud, parameters, shadow-volume and
 * extension uniforms and of the n_frames frames; the job gets its own cleared
 * RGBA32F image.  The context is left without grid and shadow volume.  All
 * jobs share the grid storage type and layout and the tracking and NEE modes;
 * at most CR_BATCH_MAX_JOBS jobs (CR_ERR_CAPACITY beyond). */
CR_API cr_status cr_batch_add(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, uint32_t* job);
CR_API cr_status cr_batch_clear(cr_context* ctx); /* frees every job */
CR_API cr_status cr_batch_size(cr_context* ctx, uint32_t* n_jobs);
CR_API cr_status cr_batch_set_frames(cr_context* ctx, uint32_t job, const cr_push_constants* frames, uint32_t n_frames);
CR_API cr_status cr_batch_clear_images(cr_context* ctx);
CR_API cr_status cr_batch_set_tuning(cr_context* ctx, const cr_batch_tuning* t);
CR_API cr_status cr_batch_get_tuning(cr_context* ctx, cr_batch_tuning* t);
/* Renders every job's frames (graph launch; asynchronous like cr_pt_render). */
CR_API cr_status cr_batch_render(cr_context* ctx);
/* Copies job's image (W x H RGBA32F) to rgba; width/height may be NULL. */
CR_API cr_status cr_batch_download(cr_context* ctx, uint32_t job, float* rgba, uint32_t* width, uint32_t* height);
CR_API cr_status cr_batch_stats_get(cr_context* ctx, cr_pt_stats* out);
CR_API cr_status cr_batch_stats_reset(cr_context* ctx);
/* Autotuner: times a trial render of 1/32 of every job's frames (at least
 * one; into scratch images, the jobs' images are untouched) for each launch
 * shape (blockThreads x blocksPerCU from half the occupancy limit up x
 * tileWidth) and keeps the fastest.  With tune_bricks, first times each job
 * alone with majorant bricks of 8, 16 and 32 voxels and keeps the fastest
 * (this changes the job's extension block: the job then equals cr_pt_render
 * with that majorantBrick).  Installs and returns the chosen tuning. */
CR_API cr_status cr_batch_autotune(cr_context* ctx, int tune_bricks, cr_batch_tuning* out, cr_batch_tune_report* report);


/* ------------------------------------------------------------ photon beams
 * RenderTechniquePPB (RenderTechniquePPB.h:7-129): beam production
 * (shaders/PPB_PT.comp), the LBVH build (Morton codes, radix sort,
 * PPB_GenerateHierarchy.comp, PPB_CalculateAABB.comp) and the gather
 * (PPB_PE.comp).  The reference builds these pipelines but dispatches only the
 * tracer and one local-sort pass (RenderTechniquePPB.cpp:436-519); here every
 * stage runs. */
typedef struct cr_beam_trace_params {
    uint32_t numPaths;        /* photon paths (one per thread; id = path index) */
    int32_t seed;             /* initializeRandom(seed * id) as PPB_PT.comp:513 */
    float radius;             /* beam radius (pmRadius, fixed) */
    float albedo;             /* continuation probability at a real collision (1 - absorption) */
    float lightDirection[4];  /* propagation direction of sunlight (normalised) */
    float lightIntensity;
    uint32_t maxInteractions; /* cap on scattering events per path */
    uint32_t _reserved[2];
} cr_beam_trace_params;
CR_API void cr_beam_trace_params_default(cr_beam_trace_params* p);

/* RenderTechniquePPB::UpdateRadius (RenderTechniquePPB.cpp:530-546), called
 * with PushConstants.frameCount before each frame's beams are traced
 * (:439): frame_number <= 1 gives initial_radius; otherwise the previous
 * frame's radius times sum_{j=1..beams_per_pass} (c + j + alpha) / (c + j + 1)
 * with c = (frame_number - 2) * beams_per_pass, in the reference's integer
 * (unsigned, then int) and float arithmetic.  The reference's constants are
 * CR_PPB_ALPHA and CR_PPB_BEAMS_PER_PASS (RenderTechniquePPB.h:115-120).
 * As written the radius is multiplied by a sum of ratios near 1, so it grows
 * about beams_per_pass-fold per frame; the function keeps that behaviour (a
 * caller wanting a fixed radius passes its own to cr_ppb_trace_beams). */
#define CR_PPB_ALPHA 0.8f
#define CR_PPB_BEAMS_PER_PASS 56u
CR_API float cr_ppb_update_radius(uint32_t frame_number, uint32_t beams_per_pass, float alpha, float initial_radius,
                                  float previous_radius);

/* Trace sun photon paths through the current grid and deposit beams, in path
 * order (deterministic: a counting pass, a scan and a writing pass replace
 * the reference's atomic appends, PPB_PT.comp:340, :381).  Keeps at most
 * max_beams segments; *out_beams and *out_data receive the counts kept. */
CR_API cr_status cr_ppb_trace_beams(cr_context* ctx, const cr_beam_trace_params* p, uint32_t max_beams,
                                    uint32_t* out_beams, uint32_t* out_data);
/* Replace the device beams with caller data (n beams, n_data PhotonBeamData). */
CR_API cr_status cr_beams_upload(cr_context* ctx, const cr_photon_beam* beams, uint32_t n,
                                 const cr_photon_beam_data* data, uint32_t n_data);
CR_API cr_status cr_beams_count(cr_context* ctx, uint32_t* n, uint32_t* n_data);
CR_API cr_status cr_beams_download(cr_context* ctx, cr_photon_beam* beams, cr_photon_beam_data* data);

/* RenderTechniquePPB::AllocateResources / FreeResources (RenderTechniquePPB.h:13-16,
 * RenderTechniquePPB.cpp:282-355): reserve every beam, sort and tree buffer for
 * up to max_beams beams (CR_ERR_CAPACITY outside [2, 2^28]), so later traces and
 * builds within that size allocate nothing; release them all (the beams too). */
CR_API cr_status cr_ppb_allocate(cr_context* ctx, uint32_t max_beams);
CR_API cr_status cr_ppb_free(cr_context* ctx);

/* LBVH build over the device beams: 30-bit Morton codes of centroids in the
 * cloud box, 4 x 8-bit stable LSD radix sort, Karras hierarchy, AABB refit.
 * Beams are reordered on the device into sorted order (as the reference's
 * sort moves whole PhotonBeam records). */
CR_API cr_status cr_lbvh_build(cr_context* ctx);
/* Host copies of the last build: sorted codes (n), permutation (n: sorted
 * position -> input index), nodes (2n-1, reference TreeNode layout).  Any
 * pointer may be NULL. */
CR_API cr_status cr_lbvh_download(cr_context* ctx, uint32_t* morton_sorted, uint32_t* perm, cr_tree_node* nodes);
/* Per-stage device times of the last build in ms: morton, sort, hierarchy, refit, pack. */
CR_API cr_status cr_lbvh_timings(cr_context* ctx, double out_ms[5]);

/* Gather options (extension). */
typedef struct cr_gather_ext {
    uint32_t refCompat; /* 1: PPB_PE.comp:362-404 term as written (sin^2, no 1/r, unscaled density, segment-relative T_beam) */
    uint32_t jitter;    /* 1: sub-pixel jitter drawn from the pixel's RNG stream */
    float albedo;       /* scattering albedo used in the corrected term */
    uint32_t _reserved;
} cr_gather_ext;
CR_API void cr_gather_ext_default(cr_gather_ext* g);
/* PPB_PE.comp main (:471-487) per frame with running-mean accumulation into
 * the same RGBA32F image as cr_pt_render.  Uses the current shadow volume
 * as camera transmittance (main.cpp:226-229 builds it along the camera). */
CR_API cr_status cr_ppb_gather(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames,
                               const cr_gather_ext* ext);
typedef struct cr_gather_stats {
    uint64_t rays, nodesVisited, beamsTested;
    double lastKernelMs;
} cr_gather_stats;
CR_API cr_status cr_gather_stats_get(cr_context* ctx, cr_gather_stats* out);

/* ---------------------------------------------------------- photon mapping
 * RenderTechniquePPM (RenderTechniquePPM.cpp:229-312): per frame, photon
 * tracing into a uniform grid of voxelCount cells, at most
 * CR_PPM_PHOTONS_PER_CELL photons each (shaders/PPM_PT.comp), then a
 * ray-marched radiance estimate per pixel: shadow-volume direct light, one
 * sky direction, and the photons within pmRadius (shaders/PPM_PE.comp).
 * Corrections (the estimate is otherwise the reference's):
 *   - a cell keeps its first CR_PPM_PHOTONS_PER_CELL photons in emission
 *     order (path id, then collision), not in atomic arrival order, so the
 *     map is a pure function of the inputs; the estimate reads only those
 *     (the reference loops over every arrival, past the cell's 32 slots);
 *   - a photon stores the azimuth atan2(y, x) in [0, 2 pi) (the reference
 *     doubles it, PPM_PT.comp:187-190, so the estimate rebuilt another
 *     direction), and a deposit at the grid's far faces goes to the last cell;
 *   - sun photons start on the entry faces uniformly over the projected area
 *     (as cr_ppb_trace_beams does; the reference projects uniform points of
 *     the volume back along the light, which its pdf 1/area does not
 *     describe), sky photons take a uniform direction and the same area
 *     sampling for it (the reference swaps the polar angles and offsets the
 *     start by 0.1); the powers keep the reference's factors (1/2 per branch
 *     probability, the two "manual adjustment" halvings, PPM_PT.comp:460-476);
 *   - the estimate's sky ray marches with ppm.ambientStep (100 world units in
 *     the reference, PPM_PE.comp:515) and skips the six density taps whose
 *     mean the reference computes but never uses (PPM_PE.comp:355-366). */
#define CR_PPM_PHOTONS_PER_CELL 32u
#define CR_PPM_ALPHA 0.8f
typedef struct cr_ppm_params {
    uint32_t numPaths;        /* photon paths per frame: 200 x 64 (RenderTechniquePPM.cpp:253, PPM_PT.comp:3) */
    float sunProbability;     /* 0.5 (PPM_PT.comp:462) */
    float ambientStep;        /* the estimate's sky-ray step (world units) */
    uint32_t maxCollisions;   /* ceiling on real collisions per photon path (the reference has none) */
} cr_ppm_params;
CR_API void cr_ppm_params_default(cr_ppm_params* p);
/* PhotonMapProperties defaults (UniformBuffers.h:214-222) and SetBounds (:225-243). */
CR_API void cr_photon_map_default(cr_photon_map_properties* p);
CR_API void cr_photon_map_set_bounds(cr_photon_map_properties* p, const float b0[4], const float b1[4]);
CR_API cr_status cr_set_photon_map(cr_context* ctx, const cr_photon_map_properties* p);
/* RenderTechniquePPM::UpdateRadius (RenderTechniquePPM.cpp:302-312):
 * initial_radius for frames 0 and 1, else r * ((n - 1 + alpha) / n)^(1/3). */
CR_API float cr_ppm_update_radius(uint32_t frame_number, float alpha, float initial_radius, float previous_radius);
/* Photon tracing with the frame seed (PPM_PT.comp) into the photon grid;
 * deposits_out (optional): photons deposited, kept or not. */
CR_API cr_status cr_ppm_trace(cr_context* ctx, int32_t seed, const cr_ppm_params* p, uint32_t* deposits_out);
/* The radiance estimate (PPM_PE.comp) of n_frames frames over the current
 * photon grid into the accumulation image (cr_pt_render's running mean);
 * frames[f].pmRadius is the query radius.  Needs cr_sv_build. */
CR_API cr_status cr_ppm_estimate(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames,
                                 const cr_ppm_params* p);
/* The photon grid: arrivals per cell (as the reference's collisions[cell].x;
 * n_cells = voxelCount product) and every deposited photon sorted by cell in
 * emission order (n_deposits); a cell's kept photons are the first
 * min(arrivals, CR_PPM_PHOTONS_PER_CELL) of its run.  NULL pointers return only
 * the sizes. */
CR_API cr_status cr_ppm_download(cr_context* ctx, uint32_t* n_cells, uint32_t* n_deposits, uint32_t* arrivals,
                                 cr_photon* photons);

/* --------------------------------------------- learned in-scattering (extension)
 * A 224 -> 128 -> 128 -> 128 -> 128 -> 3 ReLU network on bf16 MFMA with fp32
 * accumulation, predicting log(1 + Y) per RGB channel, where Y is the radiance
 * the plain path tracer adds after a real collision (no reference
 * counterpart).  Parameters live on the device as fp32 masters in a padded
 * flat layout (layer l: weights [K_l][N_l] row-major at wOffset[l], bias
 * [N_l] at bOffset[l]; K = 224,128,128,128,128; N = 128,128,128,128,16 of
 * which the last layer uses 3) plus bf16 copies used by the kernels. */
#define CR_MLP_INPUTS 224
CR_API void cr_mlp_param_layout(uint32_t K[5], uint32_t N[5], uint32_t wOffset[5], uint32_t bOffset[5],
                                uint32_t* total);
/* He-normal weights, zero biases, zero Adam state (deterministic in seed). */
CR_API cr_status cr_mlp_init(cr_context* ctx, uint32_t seed);
CR_API cr_status cr_mlp_set_params(cr_context* ctx, const float* params, uint32_t n);
CR_API cr_status cr_mlp_get_params(cr_context* ctx, float* params, uint32_t n);
/* Forward pass of B samples: X is [B][224] bf16 (host), Y is [B][3] (host). */
CR_API cr_status cr_mlp_forward(cr_context* ctx, const uint16_t* X, uint32_t B, float* Y);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared
 * error of the output against log(1 + targets); returns the loss and, when
 * grads is not NULL, the parameter gradients (padded layout) of this step. */
CR_API cr_status cr_mlp_train_step(cr_context* ctx, const uint16_t* X, const float* targets, uint32_t B, float lr,
                                   float* loss, float* grads);
/* Exponential moving average of the parameters over training steps: after
 * each Adam step, ema = decay * ema + (1 - decay) * params, and inference
 * (cr_mlp_forward, cr_learn_render) uses ema / (1 - decay^steps) in place of
 * the parameters once a step has been averaged.  decay in [0, 1); 0 (the
 * default) turns it off.  cr_mlp_set_ema, cr_mlp_init and cr_mlp_set_params
 * restart the average; training itself always updates the parameters. */
CR_API cr_status cr_mlp_set_ema(cr_context* ctx, float decay);
/* The bias-corrected average (padded layout); CR_ERR_STATE before any averaged step. */
CR_API cr_status cr_mlp_get_ema(cr_context* ctx, float* params, uint32_t n);
/* One 16x16x32 bf16 MFMA on row-major A (16x32) and B (32x16): D = A B (16x16). */
CR_API cr_status cr_mlp_probe(cr_context* ctx, const uint16_t* A, const uint16_t* B, float* D);

/* The kernels' elementary functions applied to n floats on the device
 * (in and out are host arrays).  The path tracer's log/exp/sin/cos/sqrt are
 * these exact functions; the probe lets a test compare them bit for bit with a
 * CPU implementation.  No reference counterpart: the reference calls GLSL's
 * built-ins, whose precision the Vulkan driver defines. */
/* CR_MATH_ATAN2 pairs the inputs: out[i] = atan2(in[i], in[i ^ 1]) (in[i] itself
 * for the last of an odd count).  CR_MATH_BF16: in[i] rounded to bf16 by the
 * learned network's conversion, returned as the float it represents. */
enum { CR_MATH_SQRT = 0, CR_MATH_LOG = 1, CR_MATH_EXP = 2, CR_MATH_SIN = 3, CR_MATH_COS = 4, CR_MATH_ACOS = 5,
       CR_MATH_ATAN2 = 6, CR_MATH_BF16 = 7 };
CR_API cr_status cr_math_probe(cr_context* ctx, int op, const float* in, uint32_t n, float* out);

/* Density mip pyramid (8 levels, 2x2x2 means) of the current grid, used by the
 * network's stencil features; rebuilt after every grid change. */
CR_API cr_status cr_learn_build_mips(cr_context* ctx);
/* Stencil features of n shading points: points is [n][8] floats (x, y, z,
 * omega_x, omega_y, omega_z, unused, unused); X receives [n][224] bf16.  Uses
 * the current cloud properties and shadow-volume light direction. */
CR_API cr_status cr_learn_features(cr_context* ctx, const float* points, uint32_t n, uint16_t* X);

/* Record pass: each frame's camera path (cr_pt_render's sample) up to its
 * first real collision.  records receives [n_frames][H][W] records of 8
 * floats: a collision is (x, y, z, dx, dy, dz, RNG state bits, 1); a path
 * that escapes before any collision is (r, g, b, a, 0, 0, 0, 0) with its
 * radiance.  The accumulation image is not touched. */
CR_API cr_status cr_pt_record(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, float* records);
/* Continuation pass: for each of n records, the mean radiance (out: [n][4])
 * of spp continuations of the path from its collision (NEE, scattering,
 * further bounces, exactly as cr_pt_render continues it).  The first
 * continuation resumes the recorded RNG state, so record + continue with
 * spp = 1 reproduces cr_pt_render's sample bit for bit; escaped records give 0. */
CR_API cr_status cr_pt_continue(cr_context* ctx, const float* records, uint32_t n, uint32_t spp, float* out);
/* Training data from the current scene: record pass over `frames`, the
 * collisions compacted in record order, every k-th of them (k chosen so that
 * at most max_points remain; 0 = all) continued with spp samples (the target)
 * and featurised; appended to the library's data set.  *added = rows added.
 * With shadow-volume NEE (and a bounce cap >= 1) the target leaves out the
 * next-event term of the collision itself, (((1 * I) * pdf) * Tv) * albedo,
 * which is a function of the recorded position and direction; the learned
 * render adds it exactly.  A data set holds one kind of target: collecting
 * under the other NEE mode fails with CR_ERR_STATE until it is cleared. */
CR_API cr_status cr_learn_collect(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames, uint32_t spp,
                                  uint32_t max_points, uint32_t* added);
CR_API cr_status cr_learn_dataset_size(cr_context* ctx, uint64_t* n);
CR_API cr_status cr_learn_dataset_clear(cr_context* ctx);
/* Rows [first, first+n) of the data set: X [n][224] bf16, T [n][3]; either may be NULL. */
CR_API cr_status cr_learn_dataset_download(cr_context* ctx, uint64_t first, uint32_t n, uint16_t* X, float* T);
/* steps Adam steps on minibatches of `batch` rows drawn from the data set by a
 * seeded hash; initialises the network (cr_mlp_init(seed)) if there is none.
 * *loss = the last step's loss. */
CR_API cr_status cr_learn_train(cr_context* ctx, uint32_t steps, uint32_t batch, float lr, uint32_t seed, float* loss);
/* Learned render: record pass, features and network at every collision, and
 * the running mean into the accumulation image (cr_pt_render's contract, with
 * the network standing for the rest of each path; with shadow-volume NEE the
 * collision's own next-event term is evaluated exactly and the network stands
 * for the rest).  CR_ERR_STATE if the network was trained (cr_learn_train) on
 * targets of the other NEE mode. */
CR_API cr_status cr_learn_render(cr_context* ctx, const cr_push_constants* frames, uint32_t n_frames);

#ifdef __cplusplus
}
#endif
#endif /* CR_ABI_H */
