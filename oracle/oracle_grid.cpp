/*
 * ORACLE — test infrastructure only.
 *
 * Density-grid helpers: Grid3D::GetMajorant (Grid3D.h:180-192), the
 * brick-majorant grid (extension), IEEE half conversion, and the seeded
 * value-noise fBm cloud generator (extension: the reference ships no cloud,
 * models/PUT_YOUR_MODELS_HERE.txt).
 *
 * Generator definition (the build defines it; the device kernel evaluates the
 * same float sequence):
 *   hash      : PCG hash of Jarzynski & Olano, "Hash Functions for GPU
 *               Rendering", JCGT 9(3), 2020.
 *   lattice   : h = H(seed ^ (octave * 0x9E3779B9)); h = H(h + ix); h = H(h + iy);
 *               h = H(h + iz); value = (h >> 8) * 2^-24.
 *   noise     : trilinear interpolation of the lattice with the cubic fade
 *               s(t) = t^2 (3 - 2t), lerp(a,b,t) = a + t (b - a).
 *   fbm       : sum_o amp_o noise(p * f_o, o), amp_0 = 1/2, amp *= gain,
 *               f_0 = baseFrequency, f *= lacunarity, divided by sum_o amp_o.
 *   shape     : cumulus = radial falloff (0.20 to 0.48) around (0.5, 0.42, 0.5) of the
 *               normalized box; anvil = height envelope times a radius that
 *               widens from 0.22 to 0.48 between heights 0.50 and 0.80.
 *   density   : smoothstep(thresholdLo, thresholdHi, fbm * shape).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_core.h"

using namespace orc;

namespace {

inline uint32_t pcg_hash(uint32_t x)
{
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

inline float lattice(int32_t ix, int32_t iy, int32_t iz, uint32_t octave, uint32_t seed)
{
    uint32_t h = pcg_hash(seed ^ (octave * 0x9E3779B9u));
    h = pcg_hash(h + (uint32_t)ix);
    h = pcg_hash(h + (uint32_t)iy);
    h = pcg_hash(h + (uint32_t)iz);
    return (float)(h >> 8) * 5.9604644775390625e-8f;
}

inline float lerpf(float a, float b, float t) { return a + t * (b - a); }
inline float fade(float t) { return (t * t) * (3.0f - 2.0f * t); }

inline float value_noise(float x, float y, float z, uint32_t octave, uint32_t seed)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float sx = fade(x - fx), sy = fade(y - fy), sz = fade(z - fz);
    float c000 = lattice(ix, iy, iz, octave, seed);
    float c100 = lattice(ix + 1, iy, iz, octave, seed);
    float c010 = lattice(ix, iy + 1, iz, octave, seed);
    float c110 = lattice(ix + 1, iy + 1, iz, octave, seed);
    float c001 = lattice(ix, iy, iz + 1, octave, seed);
    float c101 = lattice(ix + 1, iy, iz + 1, octave, seed);
    float c011 = lattice(ix, iy + 1, iz + 1, octave, seed);
    float c111 = lattice(ix + 1, iy + 1, iz + 1, octave, seed);
    float n00 = lerpf(c000, c100, sx);
    float n10 = lerpf(c010, c110, sx);
    float n01 = lerpf(c001, c101, sx);
    float n11 = lerpf(c011, c111, sx);
    float n0 = lerpf(n00, n10, sy);
    float n1 = lerpf(n01, n11, sy);
    return lerpf(n0, n1, sz);
}

inline float fbm(const cr_fbm_params& p, float x, float y, float z)
{
    float sum = 0.0f, norm = 0.0f, amp = 0.5f, f = p.baseFrequency;
    for (uint32_t o = 0; o < p.octaves; o++) {
        sum = sum + amp * value_noise(x * f, y * f, z * f, o, p.seed);
        norm = norm + amp;
        amp = amp * p.gain;
        f = f * p.lacunarity;
    }
    return sum / norm;
}

inline float shape_fn(const cr_fbm_params& p, float ux, float uy, float uz)
{
    if (p.shape == CR_SHAPE_ANVIL) {
        float dx = ux - 0.5f, dz = uz - 0.5f;
        float rh = sqrtf(dx * dx + dz * dz);
        float R = 0.22f + 0.26f * smoothstep_f(0.50f, 0.80f, uy);
        float env = smoothstep_f(0.02f, 0.10f, uy) * (1.0f - smoothstep_f(0.90f, 0.98f, uy));
        return env * (1.0f - smoothstep_f(R * 0.55f, R, rh));
    }
    float dx = ux - 0.5f, dy = (uy - 0.42f) * 1.25f, dz = uz - 0.5f;
    float r = sqrtf((dx * dx + dy * dy) + dz * dz);
    return 1.0f - smoothstep_f(0.20f, 0.48f, r);
}

} // namespace

extern "C" {

uint32_t or_hash32(uint32_t x) { return pcg_hash(x); }

float or_fbm_voxel(const cr_fbm_params* p, uint32_t ix, uint32_t iy, uint32_t iz)
{
    float cx = (float)ix + 0.5f, cy = (float)iy + 0.5f, cz = (float)iz + 0.5f;
    float ux = cx / (float)p->size[0], uy = cy / (float)p->size[1], uz = cz / (float)p->size[2];
    float sh = shape_fn(*p, ux, uy, uz);
    if (sh <= 0.0f) return 0.0f;
    float v = fbm(*p, cx * p->voxelSize, cy * p->voxelSize, cz * p->voxelSize) * sh;
    return smoothstep_f(p->thresholdLo, p->thresholdHi, v);
}

void or_fbm_grid(const cr_fbm_params* p, float* out, int32_t threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int64_t nz = p->size[2], ny = p->size[1], nx = p->size[0];
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t z = 0; z < nz; z++)
        for (int64_t y = 0; y < ny; y++)
            for (int64_t x = 0; x < nx; x++)
                out[x + nx * (y + ny * z)] = or_fbm_voxel(p, (uint32_t)x, (uint32_t)y, (uint32_t)z);
}

uint16_t or_f32_to_f16(float f)
{
    uint32_t b = o_f2u(f);
    uint32_t sign = (b >> 16) & 0x8000u;
    uint32_t a = b & 0x7fffffffu;
    if (a >= 0x7f800000u) return (uint16_t)(sign | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (a < 0x38800000u) {                                   /* half subnormal range */
        float v = o_u2f(a) * 16777216.0f;                     /* exact scaling by 2^24 */
        return (uint16_t)(sign | (uint32_t)rintf(v));         /* round half to even */
    }
    uint32_t e = (a >> 23) - 112u;
    uint32_t h = (e << 10) | ((a & 0x7fffffu) >> 13);
    uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
    return (uint16_t)(sign | h);
}

float or_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

float or_grid_max(const float* grid, uint64_t n)
{
    float m = 0.0f; /* T() */
    for (uint64_t i = 0; i < n; i++)
        if (m < grid[i]) m = grid[i];
    return m;
}

void or_brick_majorants(const float* grid, const uint32_t dims[3], uint32_t brick, float* out, uint32_t md[3])
{
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2], b = brick;
    md[0] = (uint32_t)((nx + b - 1) / b);
    md[1] = (uint32_t)((ny + b - 1) / b);
    md[2] = (uint32_t)((nz + b - 1) / b);
    for (int64_t bz = 0; bz < md[2]; bz++)
        for (int64_t by = 0; by < md[1]; by++)
            for (int64_t bx = 0; bx < md[0]; bx++) {
                float m = 0.0f;
                int64_t x0 = std::max<int64_t>(bx * b - 1, 0), x1 = std::min<int64_t>(bx * b + b, nx - 1);
                int64_t y0 = std::max<int64_t>(by * b - 1, 0), y1 = std::min<int64_t>(by * b + b, ny - 1);
                int64_t z0 = std::max<int64_t>(bz * b - 1, 0), z1 = std::min<int64_t>(bz * b + b, nz - 1);
                for (int64_t z = z0; z <= z1; z++)
                    for (int64_t y = y0; y <= y1; y++)
                        for (int64_t x = x0; x <= x1; x++) {
                            float v = grid[x + nx * (y + ny * z)];
                            if (v > m) m = v;
                        }
                out[bx + md[0] * (by + md[1] * bz)] = m;
            }
}

} // extern "C"
