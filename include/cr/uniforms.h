/*
 * Shader-visible data contracts, byte-identical to the reference's
 * CloudRendering-Vulkan/UniformBuffers.h (glm 0.9.9.7 layouts, vec3 = 12 B,
 * vec4 = 16 B with 4-B alignment, mat4 column-major).
 *
 * Every struct here is what a caller of the reference fills on the host and
 * hands to a technique; this library takes the same bytes through its C-ABI.
 * The only representational change is Parameters.isotropic: the reference
 * declares a C++ bool followed by 3 padding bytes and the GLSL side reads a
 * 4-byte bool (UniformBuffers.h:90 vs shaders/PathTracer.comp:76).  Here the
 * four bytes are one uint32_t and the library tests only the low byte, so a
 * struct written by the reference's own C++ (bool + garbage padding) and one
 * written through this header read the same.
 *
 * Plain C (C99 + C11 static asserts) so that C, C++, HIP and ctypes callers
 * share it.
 */
#ifndef CR_UNIFORMS_H
#define CR_UNIFORMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define CR_STATIC_ASSERT(c, m) static_assert(c, m)
extern "C" {
#else
#define CR_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

/* UniformBuffers.h:6-13 — per-frame push constants (24 B, align 8). */
typedef struct cr_push_constants {
    double time;
    int32_t seed;          /* per-frame seed: std::rand() after std::srand(0) (main.cpp:722, :949) */
    uint32_t frameCount;   /* running-mean weight (PathTracer.comp:420-424) */
    float pmRadius;        /* photon-beam radius (RenderTechniquePPB.cpp:530-546) */
    uint32_t currentBuffer;
} cr_push_constants;

/* UniformBuffers.h:15-69 — pinhole camera (68 B). */
typedef struct cr_camera_properties {
    float position[3];
    int32_t halfWidth;
    float forward[3];
    int32_t halfHeight;
    float right[3];
    float nearPlane;
    float up[3];
    float pixelSizeY;
    float pixelSizeX;
} cr_camera_properties;

/* UniformBuffers.h:71-78 — cloud box and density scaling (60 B). */
typedef struct cr_cloud_properties {
    float bounds[2][4];
    uint32_t voxelCount[4];
    float maxExtinction;
    float baseScaling;
    float densityScaling;
} cr_cloud_properties;

/* UniformBuffers.h:80-112 — phase function and light (28 B). */
typedef struct cr_parameters {
    uint32_t maxRayBounces;
    float lightIntensity;
    float phaseG;
    float phaseOnePlusG2;
    float phaseOneMinusG2;
    float phaseOneOver2G;
    uint32_t isotropic; /* low byte = reference bool; upper bytes are padding */
} cr_parameters;

/* UniformBuffers.h:114-167 — light-space shadow volume frame (152 B). */
typedef struct cr_shadow_volume_properties {
    float bounds[2][4];
    float lightDirection[4];
    float right[4];
    float up[4];
    float basisChange[16]; /* column-major 4x4: basisChange[col*4 + row] */
    uint32_t voxelAxisCount;
    float voxelSize;
} cr_shadow_volume_properties;

/* UniformBuffers.h:169-177 — one photon-beam cylinder segment (48 B). */
typedef struct cr_photon_beam {
    float startPos[3];
    uint32_t mortonCode;
    float endPos[3];
    float radius;
    uint32_t dataIdx;
    uint32_t _padding_beam[3];
} cr_photon_beam;

/* UniformBuffers.h:179-188 — LBVH node (64 B). */
typedef struct cr_tree_node {
    uint32_t parent;
    uint32_t left;     /* internal: left child; leaf: index into the sorted beams */
    uint32_t right;
    uint32_t isLeaf;   /* low byte = reference bool */
    float bounds[2][4];
    uint32_t processed;
    uint32_t _padding_node[3];
} cr_tree_node;

/* UniformBuffers.h:190-194 — per-path beam data (80 B). */
typedef struct cr_photon_beam_data {
    float power[4];
    float trDistances[16];
} cr_photon_beam_data;

/* UniformBuffers.h:202-210 — one photon of the photon map (48 B; the
 * reference's comment says 36).  theta = acos(dir.z), phi = atan2(dir.y, dir.x)
 * in [0, 2 pi) of the photon's propagation direction. */
typedef struct cr_photon {
    float position[4];
    float power[4];
    float phi;
    float theta;
    float _padding[2];
} cr_photon;

/* UniformBuffers.h:212-249 — photon-map grid (76 B): a cube from the cloud's
 * lower corner with the cloud's largest extent as side (SetBounds), split in
 * voxelCount cells of voxelSize; stepSize is the estimate's ray-march step. */
typedef struct cr_photon_map_properties {
    float lightDirection[4];
    float bounds[2][4];
    int32_t voxelCount[3];
    float voxelSize;
    uint32_t photonSize;
    float stepSize;
    float absorption;
} cr_photon_map_properties;

/* UniformBuffers.h:196-200 — radix-sort element (8 B). */
typedef struct cr_sort_element {
    uint32_t idx;
    uint32_t mortonCode;
} cr_sort_element;

/* RenderTechniquePPB.h:41-47 — LBVH pass constants (12 B). */
typedef struct cr_lbvh_push_constants {
    uint32_t baseShift;
    uint32_t currentBuffer;
    uint32_t workGroupCount;
} cr_lbvh_push_constants;

CR_STATIC_ASSERT(sizeof(cr_push_constants) == 24, "PushConstants is 24 B");
CR_STATIC_ASSERT(sizeof(cr_camera_properties) == 68, "CameraProperties is 68 B");
CR_STATIC_ASSERT(sizeof(cr_cloud_properties) == 60, "CloudProperties is 60 B");
CR_STATIC_ASSERT(offsetof(cr_cloud_properties, maxExtinction) == 48, "maxExtinction@48");
CR_STATIC_ASSERT(offsetof(cr_cloud_properties, baseScaling) == 52, "baseScaling@52");
CR_STATIC_ASSERT(offsetof(cr_cloud_properties, densityScaling) == 56, "densityScaling@56");
CR_STATIC_ASSERT(sizeof(cr_parameters) == 28, "Parameters is 28 B");
CR_STATIC_ASSERT(offsetof(cr_parameters, isotropic) == 24, "isotropic@24");
CR_STATIC_ASSERT(sizeof(cr_shadow_volume_properties) == 152, "ShadowVolumeProperties is 152 B");
CR_STATIC_ASSERT(offsetof(cr_shadow_volume_properties, basisChange) == 80, "basisChange@80");
CR_STATIC_ASSERT(offsetof(cr_shadow_volume_properties, voxelAxisCount) == 144, "voxelAxisCount@144");
CR_STATIC_ASSERT(offsetof(cr_shadow_volume_properties, voxelSize) == 148, "voxelSize@148");
CR_STATIC_ASSERT(sizeof(cr_photon_beam) == 48, "PhotonBeam is 48 B");
CR_STATIC_ASSERT(offsetof(cr_photon_beam, mortonCode) == 12, "mortonCode@12");
CR_STATIC_ASSERT(offsetof(cr_photon_beam, endPos) == 16, "endPos@16");
CR_STATIC_ASSERT(offsetof(cr_photon_beam, radius) == 28, "radius@28");
CR_STATIC_ASSERT(offsetof(cr_photon_beam, dataIdx) == 32, "dataIdx@32");
CR_STATIC_ASSERT(sizeof(cr_tree_node) == 64, "TreeNode is 64 B");
CR_STATIC_ASSERT(offsetof(cr_tree_node, isLeaf) == 12, "isLeaf@12");
CR_STATIC_ASSERT(offsetof(cr_tree_node, bounds) == 16, "bounds@16");
CR_STATIC_ASSERT(offsetof(cr_tree_node, processed) == 48, "processed@48");
CR_STATIC_ASSERT(sizeof(cr_photon_beam_data) == 80, "PhotonBeamData is 80 B");
CR_STATIC_ASSERT(sizeof(cr_sort_element) == 8, "SortElement is 8 B");
CR_STATIC_ASSERT(sizeof(cr_photon) == 48, "Photon is 48 B");
CR_STATIC_ASSERT(offsetof(cr_photon, phi) == 32, "phi@32");
CR_STATIC_ASSERT(sizeof(cr_photon_map_properties) == 76, "PhotonMapProperties is 76 B");
CR_STATIC_ASSERT(offsetof(cr_photon_map_properties, voxelCount) == 48, "voxelCount@48");
CR_STATIC_ASSERT(offsetof(cr_photon_map_properties, stepSize) == 68, "stepSize@68");
CR_STATIC_ASSERT(sizeof(cr_lbvh_push_constants) == 12, "LBVHPushConstants is 12 B");

#ifdef __cplusplus
}
#endif
#endif /* CR_UNIFORMS_H */
