/*
 * ORACLE — test infrastructure only.
 *
 * A clean-room CPU restatement of the reference's hot path
 * (jeanfilho/CloudRendering-Vulkan) used to check the gfx950 kernels.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker or the CPU baseline; the product never links it.
 *
 * Each function cites the reference file:line it restates.  Where the
 * reference is defective (an unguarded sort histogram, a racy prefix sum,
 * sun photons started on the exit face, and the like) the corrected method
 * is implemented and, where it matters, a flag reproduces the reference.  All floating point is IEEE binary32 evaluated in a fixed order
 * with -ffp-contract=off, so the HIP kernels, which follow the same order,
 * can be compared bit for bit.
 */
#ifndef CR_ORACLE_H
#define CR_ORACLE_H

#include <stdint.h>
#include "../include/cr/uniforms.h"
#include "../include/cr/abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Everything the path tracer reads: the reference's seven descriptor
 * bindings (RenderTechniquePT.cpp:20-35) plus the extension state. */
typedef struct or_scene {
    const float* grid_f32;      /* x-fastest density, or NULL when grid_f16 is set */
    const uint16_t* grid_f16;   /* IEEE half bits, x fastest */
    uint32_t dims[3];
    const float* majorants;     /* brick maxima (16^3 + apron), x fastest */
    uint32_t mdims[3];
    const float* sv;            /* shadow volume, x fastest */
    uint32_t svdims[3];
    cr_camera_properties cam;
    cr_cloud_properties cloud;
    cr_parameters params;
    cr_shadow_volume_properties svp;
    cr_render_ext ext;
} or_scene;

typedef struct or_counters {
    uint64_t paths, vertices, lookups, svLookups, brickVisits;
} or_counters;

/* ------------------------------------------------------------------ math */
float or_logf(float x);
float or_expf(float x);
void or_sincosf(float x, float* s, float* c);
/* op is a CR_MATH_* code; sqrt is the C library's (correctly rounded). */
void or_math_array(int op, const float* in, uint32_t n, float* out);

/* ------------------------------------------------------------------- RNG
 * shaders/PathTracer.comp:220-241 (identical copies PPB_PT.comp:210-231). */
uint32_t or_xorshift32(uint32_t state);
/* initializeRandom(seed * id) then `count` raw draws (PathTracer.comp:368). */
void or_rng_stream(int32_t seed, uint32_t id, uint32_t count, uint32_t* out);
/* MSVC rand() after srand(s): per-frame seeds (main.cpp:722, :949). */
void or_msvc_rand(uint32_t srand_seed, uint32_t count, int32_t* out);

/* ---------------------------------------------------------------- sampling */
/* texture() with a linear, clamp-to-edge sampler (Initializers.cpp:408-418) at
 * normalized coordinates, exact-float trilinear weights. */
float or_trilinear_f32(const float* grid, const uint32_t dims[3], const float uvw[3]);
/* Henyey-Greenstein pdf (PathTracer.comp:247-250, :259-274). */
float or_phase_pdf(const cr_parameters* p, const float in_dir[3], const float out_dir[3]);
/* scatterRay (PathTracer.comp:277-311) with the two given uniforms drawn in order. */
void or_scatter_dir(const cr_parameters* p, const float in_dir[3], uint32_t rng_state,
                    float out_dir[3], uint32_t* rng_state_out);
/* sampleBackground (PathTracer.comp:104-113) or the constant sky. */
void or_background(const cr_render_ext* ext, const float dir[3], float out[4]);
/* getCameraRay (PathTracer.comp:144-151). */
void or_camera_ray(const cr_camera_properties* cam, int32_t x, int32_t y, float pos[3], float dir[3]);
/* intersectCloud (PathTracer.comp:154-184); returns 1 on hit. */
int or_intersect_box(const float b0[3], const float b1[3], const float pos[3], const float dir[3],
                     float* tmin, float* tmax);

/* ------------------------------------------------------------ shadow volume
 * shaders/ShadowVolume.comp:77-93 over a (dims) light-space lattice. */
void or_sv_dims(const or_scene* s, uint32_t out[3]);
void or_sv_build(const or_scene* s, float* out);

/* ------------------------------------------------------------ path tracer
 * shaders/PathTracer.comp:366-425 for one pixel and one frame seed. */
void or_pt_sample(const or_scene* s, int32_t x, int32_t y, int32_t seed, float out[4], or_counters* c);
/* n_frames consecutive frames with running-mean accumulation into accum
 * (RGBA32F, W x H); OpenMP over rows when threads > 1. */
void or_pt_render(const or_scene* s, const cr_push_constants* frames, uint32_t n_frames,
                  float* accum, int32_t threads, or_counters* c);
/* The same for the pixel rectangle [x0, x0+w) x [y0, y0+h) only. */
void or_pt_render_region(const or_scene* s, const cr_push_constants* frames, uint32_t n_frames, float* accum,
                         int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t threads, or_counters* c);

/* ------------------------------------------------------------------ grids */
uint32_t or_hash32(uint32_t x);
float or_fbm_voxel(const cr_fbm_params* p, uint32_t ix, uint32_t iy, uint32_t iz);
void or_fbm_grid(const cr_fbm_params* p, float* out, int32_t threads);
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);
/* Grid3D::GetMajorant (Grid3D.h:180-192). */
float or_grid_max(const float* grid, uint64_t n);
/* Brick majorants: max over each brick^3 brick plus a one-voxel apron. */
void or_brick_majorants(const float* grid, const uint32_t dims[3], uint32_t brick, float* out, uint32_t mdims_out[3]);

/* rounding correction of the kernels' square root, checked over [lo_bits, hi_bits] */
uint64_t or_sqrt_correction_errors(uint32_t lo_bits, uint32_t hi_bits);

/* progressive beam radius (RenderTechniquePPB.cpp:530-546) */
float or_ppb_update_radius(uint32_t frameNumber, uint32_t beamsPerPass, float alpha, float initialRadius,
                           float pmRadius);

#ifdef __cplusplus
}
#endif
#endif
