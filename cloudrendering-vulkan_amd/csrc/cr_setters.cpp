// Host-side restatements of the reference's uniform-struct setters
// (CloudRendering-Vulkan/UniformBuffers.h) and of the MSVC rand() sequence
// that seeds each frame (main.cpp:722, :949).  Host memory only; these let a
// caller build the exact bytes the reference would hand to its shaders
// without glm.
#include <math.h>
#include <string.h>

#include "../../include/cr/abi.h"

namespace {

const float DEG2RAD = 0.01745329251994329576923690768489f; // glm::radians

// utilities::GetOrthonormalBasis (Utilities.cpp:7-16).
void ortho_basis(const float f[3], float r[3], float u[3])
{
    float sz = f[2] >= 0.0f ? 1.0f : -1.0f;
    float a = f[1] / (1.0f + fabsf(f[2]));
    float b = f[1] * a;
    float c = -f[0] * a;
    r[0] = f[2] + sz * b; r[1] = sz * c; r[2] = -f[0];
    u[0] = c; u[1] = 1.0f - b; u[2] = -sz * f[1];
}

float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

} // namespace

extern "C" {

void cr_camera_default(cr_camera_properties* c)
{
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->position[0] = 0.0f; c->position[1] = 0.0f; c->position[2] = -800.0f;
    c->halfWidth = 800 / 2;
    c->forward[2] = 1.0f;
    c->halfHeight = 600 / 2;
    c->right[0] = 1.0f;
    c->nearPlane = 50.0f;
    c->up[1] = 1.0f;
    c->pixelSizeY = (2.0f * tanf((120.0f / 2.0f) * DEG2RAD) * c->nearPlane) / (float)(c->halfHeight * 2);
    c->pixelSizeX = c->pixelSizeY * (float)c->halfHeight / (float)c->halfWidth;
}

void cr_camera_set_resolution(cr_camera_properties* c, int width, int height)
{
    if (!c) return;
    c->halfWidth = width / 2;
    c->halfHeight = height / 2;
}

void cr_camera_set_fov(cr_camera_properties* c, float* fov)
{
    if (!c || !fov) return;
    *fov = clampf(*fov, 60.0f, 150.0f);
    float height = 2.0f * tanf((*fov / 2.0f) * DEG2RAD) * c->nearPlane;
    c->pixelSizeY = height / (float)(c->halfHeight * 2);
    c->pixelSizeX = c->pixelSizeY * (float)c->halfHeight / (float)c->halfWidth;
}

// forward = Rx(rot.x) * Ry(rot.y) * (0,0,1) = (sin y, -sin x cos y, cos x cos y).
void cr_camera_set_rotation(cr_camera_properties* c, float rot[2])
{
    if (!c || !rot) return;
    rot[0] = clampf(rot[0], -80.0f, 80.0f);
    float ax = rot[0] * DEG2RAD, ay = rot[1] * DEG2RAD;
    c->forward[0] = sinf(ay);
    c->forward[1] = -sinf(ax) * cosf(ay);
    c->forward[2] = cosf(ax) * cosf(ay);
    ortho_basis(c->forward, c->right, c->up);
}

void cr_parameters_default(cr_parameters* p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->maxRayBounces = 5;
    p->lightIntensity = 5.0f;
    p->phaseG = 0.00000001f;
    p->phaseOnePlusG2 = 1.0f + p->phaseG * p->phaseG;
    p->phaseOneMinusG2 = 1.0f - p->phaseG * p->phaseG;
    p->phaseOneOver2G = 0.5f / p->phaseG;
    p->isotropic = fabsf(p->phaseG) < 0.0001f ? 1u : 0u;
}

void cr_parameters_set_phase_g(cr_parameters* p, float value)
{
    if (!p) return;
    if (value > 1.0f || value < -1.0f) return;
    p->phaseG = clampf(value, -0.95f, 0.95f);
    p->isotropic = fabsf(p->phaseG) < 0.0001f ? 1u : 0u;
    if (!p->isotropic) {
        p->phaseOnePlusG2 = 1.0f + p->phaseG * p->phaseG;
        p->phaseOneMinusG2 = 1.0f - p->phaseG * p->phaseG;
        p->phaseOneOver2G = 0.5f / p->phaseG;
    }
}

void cr_cloud_default(cr_cloud_properties* c)
{
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->maxExtinction = 0.2f;
    c->baseScaling = 1000.0f;
    c->densityScaling = 200.0f;
}

void cr_shadow_volume_default(cr_shadow_volume_properties* s)
{
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->lightDirection[0] = 1.0f; s->lightDirection[1] = -1.0f;
    s->right[0] = 1.0f;
    s->up[1] = 1.0f;
    // glm::mat4{lightDirection, right, up, vec4(0)} (columns)
    for (int i = 0; i < 4; i++) {
        s->basisChange[0 * 4 + i] = s->lightDirection[i];
        s->basisChange[1 * 4 + i] = s->right[i];
        s->basisChange[2 * 4 + i] = s->up[i];
        s->basisChange[3 * 4 + i] = 0.0f;
    }
    s->voxelAxisCount = 500;
    s->voxelSize = 1.0f;
}

// SetLightDirection (UniformBuffers.h:131-143) followed by SetOrigin
// (:145-151); UpdateOrigin (:159-166).  The inverse of the orthonormal basis
// matrix {right, up, light, e_w} is taken as its transpose.
void cr_shadow_volume_set_light(cr_shadow_volume_properties* s, const float d_in[3], const float lo[3],
                                const float hi[3])
{
    if (!s || !d_in || !lo || !hi) return;
    float len = sqrtf(d_in[0] * d_in[0] + d_in[1] * d_in[1] + d_in[2] * d_in[2]);
    float d[3] = {d_in[0] / len, d_in[1] / len, d_in[2] / len};
    float r[3], u[3];
    ortho_basis(d, r, u);
    for (int i = 0; i < 3; i++) {
        s->lightDirection[i] = d[i];
        s->right[i] = r[i];
        s->up[i] = u[i];
    }
    s->lightDirection[3] = s->right[3] = s->up[3] = 0.0f;
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    float radius = sqrtf(dx * dx + dy * dy + dz * dz) / 2.0f;
    s->voxelSize = (2.0f * radius) / (float)(s->voxelAxisCount - 1u);
    float center[3] = {(lo[0] + hi[0]) / 2.0f, (lo[1] + hi[1]) / 2.0f, (lo[2] + hi[2]) / 2.0f};
    for (int i = 0; i < 3; i++) {
        float corner = radius * ((-d[i] - r[i]) - u[i]);
        s->bounds[0][i] = center[i] + corner;
        s->bounds[1][i] = center[i] - corner;
    }
    s->bounds[0][3] = 1.0f;
    s->bounds[1][3] = 1.0f;
    // inverse(mat4{right, up, light, (0,0,0,1)}) = transpose: row i of the
    // basis becomes column i.
    const float* basis[3] = {r, u, d};
    for (int col = 0; col < 3; col++) {
        for (int row = 0; row < 3; row++) s->basisChange[col * 4 + row] = basis[row][col];
        s->basisChange[col * 4 + 3] = 0.0f;
    }
    s->basisChange[12] = 0.0f; s->basisChange[13] = 0.0f; s->basisChange[14] = 0.0f; s->basisChange[15] = 1.0f;
}

// MSVC CRT rand(): state = state * 214013 + 2531011; return (state >> 16) & 0x7fff.
void cr_msvc_rand_sequence(uint32_t srand_seed, uint32_t count, int32_t* out)
{
    if (!out) return;
    uint32_t st = srand_seed;
    for (uint32_t i = 0; i < count; i++) {
        st = st * 214013u + 2531011u;
        out[i] = (int32_t)((st >> 16) & 0x7fffu);
    }
}

void cr_render_ext_default(cr_render_ext* e)
{
    if (!e) return;
    memset(e, 0, sizeof(*e));
    e->albedo = 1.0f;
    e->bounceCap = CR_UNBOUNDED;
    e->skyMode = CR_SKY_GRADIENT;
    e->skyColor[0] = e->skyColor[1] = e->skyColor[2] = e->skyColor[3] = 1.0f;
    e->neeMode = CR_NEE_SHADOW_VOLUME;
    e->trackingMode = CR_TRACK_GLOBAL_MAJORANT;
}

void cr_fbm_params_default(cr_fbm_params* p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->size[0] = p->size[1] = p->size[2] = 128;
    p->voxelSize = 1.0f / 128.0f;
    p->seed = 1;
    p->octaves = 6;
    p->lacunarity = 2.0f;
    p->gain = 0.5f;
    p->baseFrequency = 4.0f;
    p->shape = CR_SHAPE_CUMULUS;
    p->thresholdLo = 0.22f;
    p->thresholdHi = 0.42f;
}

void cr_beam_trace_params_default(cr_beam_trace_params* p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->numPaths = 56; // RenderTechniquePPB.h:118-120: 28 workgroups x 2
    p->seed = 38;
    p->radius = 0.002f;
    p->albedo = 1.0f;
    const float inv = 1.0f / sqrtf(2.0f);
    p->lightDirection[0] = inv; p->lightDirection[1] = -inv; p->lightDirection[2] = 0.0f; p->lightDirection[3] = 0.0f;
    p->lightIntensity = 5.0f;
    p->maxInteractions = 64;
}

float cr_ppb_update_radius(uint32_t frame_number, uint32_t beams_per_pass, float alpha, float initial_radius,
                           float previous_radius)
{
    if (frame_number <= 1u) return initial_radius;
    // (frameNumber - 2) * m_beamsPerPass is formed in unsigned arithmetic and
    // stored in an int; constant + j converts it back to unsigned.
    const int32_t constant = (int32_t)((frame_number - 2u) * beams_per_pass);
    float sum = 0.0f;
    for (uint32_t j = 1; j <= beams_per_pass; j++) {
        const float cj = (float)((uint32_t)constant + j);
        sum = sum + (cj + alpha) / (cj + 1.0f);
    }
    return previous_radius * sum;
}

void cr_gather_ext_default(cr_gather_ext* g)
{
    if (!g) return;
    memset(g, 0, sizeof(*g));
    g->refCompat = 0;
    g->jitter = 0;
    g->albedo = 1.0f;
}

} // extern "C"
