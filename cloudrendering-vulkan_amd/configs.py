"""Named workloads (scene, camera, light, render settings).

  c1  128^3 FP32 fBm cumulus (6 octaves, lacunarity 2, gain 0.5), extinction
      scaled to 64 per unit box; albedo 0.9, HG g = 0.85, sun along
      normalize(1,-1,0) plus a constant sky; 256x256, 64 spp, <= 16 bounces.
  c2  1024x512x1024 FP16 anvil (8 octaves) in 4^3 bricks, 16^3 brick
      majorants, 256x128x256 shadow volume; albedo 0.999, g = 0.85;
      1920x1080, 256 spp, <= 256 bounces, ratio-tracked shadow rays.
  c3  photon beams traced on the c1 volume, 2^22 segments, radius 0.002 of
      the box; 30-bit Morton + 4x8-bit sort + Karras + refit; 1080p gather.
  c4  64 independent jobs for one batched launch: FP16 cumulus grids with a
      log-uniform side in [64, 512], square images with a log-uniform side in
      [128, 1024], log-uniform 16..512 spp, g ~ U[0.6, 0.9], the sun uniform
      on the upper hemisphere; brick-majorant tracking, ratio-tracked shadow
      rays, albedo 0.9, <= 16 bounces.

The camera mirrors the reference defaults (UniformBuffers.h:17-26: position
(0,0,-800), near plane 50, looking down +z) scaled to a unit-size cloud
(baseScaling 1 instead of 1000): position (0,0,-0.8), near plane 0.05.
"""
import ctypes as C
import math

from . import structs as S
from ._lib import load


def msvc_seeds(count, srand_seed=0):
    """Per-frame seeds: std::rand() after std::srand(0) (main.cpp:722, :949)."""
    out = (C.c_int32 * count)()
    load().cr_msvc_rand_sequence(srand_seed, count, out)
    return list(out)


def frames(count, first=0, ref_frame_count=False, srand_seed=0):
    """PushConstants for frames first..first+count-1.

    frameCount starts at 0 so the accumulation is the exact mean of the frames
    (the reference starts at 1 over an uncleared image); ref_frame_count=True starts at 1 as main.cpp:262 does."""
    seeds = msvc_seeds(first + count, srand_seed)[first:]
    base = 1 if ref_frame_count else 0
    return [S.PushConstants(0.0, seeds[i], base + first + i, 0.0, 0) for i in range(count)]


def camera(width, height, fov=90.0, position=(0.0, 0.0, -0.8), near=0.05, rotation=None):
    lib = load()
    cam = S.CameraProperties()
    lib.cr_camera_default(C.byref(cam))
    cam.position[:] = position
    cam.nearPlane = near
    lib.cr_camera_set_resolution(C.byref(cam), width, height)
    f = C.c_float(fov)
    lib.cr_camera_set_fov(C.byref(cam), C.byref(f))
    if rotation is not None:
        r = (C.c_float * 2)(*rotation)
        lib.cr_camera_set_rotation(C.byref(cam), r)
    return cam


def parameters(g=0.85, light_intensity=5.0):
    lib = load()
    p = S.Parameters()
    lib.cr_parameters_default(C.byref(p))
    p.lightIntensity = light_intensity
    lib.cr_parameters_set_phase_g(C.byref(p), g)
    return p


def shadow_volume(cloud, light_dir=(1.0, -1.0, 0.0), axis_count=256):
    lib = load()
    s = S.ShadowVolumeProperties()
    lib.cr_shadow_volume_default(C.byref(s))
    s.voxelAxisCount = axis_count
    d = (C.c_float * 3)(*light_dir)
    lo = (C.c_float * 3)(*cloud.bounds[0][:3])
    hi = (C.c_float * 3)(*cloud.bounds[1][:3])
    lib.cr_shadow_volume_set_light(C.byref(s), d, lo, hi)
    return s


def render_ext(albedo=1.0, bounce_cap=S.CR_UNBOUNDED, sky=None, nee=S.CR_NEE_SHADOW_VOLUME,
               tracking=S.CR_TRACK_GLOBAL_MAJORANT, sv_dims=(0, 0, 0), majorant_brick=0):
    e = S.RenderExt()
    load().cr_render_ext_default(C.byref(e))
    e.albedo = albedo
    e.bounceCap = bounce_cap
    if sky is not None:
        e.skyMode = S.CR_SKY_CONSTANT
        e.skyColor[:] = tuple(sky) + (1.0,) * (4 - len(sky))
    e.neeMode = nee
    e.trackingMode = tracking
    e.svDims[:] = sv_dims
    e.majorantBrick = majorant_brick
    return e


SKY = (0.55, 0.65, 0.85, 1.0)


def fbm(size=(128, 128, 128), voxel_size=1.0 / 128.0, seed=1, octaves=6, shape=S.CR_SHAPE_CUMULUS,
        lo=0.16, hi=0.36, base_frequency=4.0):
    p = S.FbmParams()
    load().cr_fbm_params_default(C.byref(p))
    p.size[:] = size
    p.voxelSize = voxel_size
    p.seed = seed
    p.octaves = octaves
    p.lacunarity = 2.0
    p.gain = 0.5
    p.baseFrequency = base_frequency
    p.shape = shape
    p.thresholdLo = lo
    p.thresholdHi = hi
    return p


class Scene:
    """Everything one render needs, in reference structs."""

    def __init__(self, name, fbm_params, storage, grid_layout, base_scaling, density_scaling, cam, params, ext,
                 sv_axis, spp, light_dir=(1.0, -1.0, 0.0)):
        self.name = name
        self.fbm = fbm_params
        self.storage = storage
        self.grid_layout = grid_layout
        self.base_scaling = base_scaling
        self.density_scaling = density_scaling
        self.cam = cam
        self.params = params
        self.ext = ext
        self.sv_axis = sv_axis
        self.spp = spp
        self.light_dir = light_dir

    @property
    def width(self):
        return 2 * self.cam.halfWidth

    @property
    def height(self):
        return 2 * self.cam.halfHeight

    def setup(self, ctx, build_sv=True):
        """Generate the grid on the device and bind every uniform (the
        reference's UpdateCloudData + UpdateShadowVolume sequence,
        main.cpp:223-335)."""
        c = S.CloudProperties()
        load().cr_cloud_default(C.byref(c))
        c.baseScaling = self.base_scaling
        c.densityScaling = self.density_scaling
        ctx.set_cloud(c)
        ctx.set_grid_layout(self.grid_layout)
        ctx.generate_fbm(self.fbm, self.storage)
        cloud = ctx.cloud()
        ctx.set_camera(self.cam)
        ctx.set_parameters(self.params)
        ctx.set_render_ext(self.ext)
        svp = shadow_volume(cloud, self.light_dir, self.sv_axis)
        ctx.set_shadow_volume(svp)
        if build_sv:
            ctx.sv_build()
        ctx.clear_image()
        return cloud, svp


def c1(width=256, height=256, spp=64, size=128, sv_axis=256):
    vs = 1.0 / size
    return Scene("c1", fbm((size, size, size), vs, seed=1, octaves=6), S.CR_DTYPE_F32, S.CR_GRID_DENSE, 1.0, 64.0,
                 camera(width, height, position=(0.0, 0.0, -0.2)), parameters(0.85),
                 render_ext(albedo=0.9, bounce_cap=16, sky=SKY), sv_axis, spp)


def c2(width=1920, height=1080, spp=256, size=(1024, 512, 1024), storage=S.CR_DTYPE_F16, layout=S.CR_GRID_BRICK4,
       nee=S.CR_NEE_RATIO_TRACKING, tracking=S.CR_TRACK_BRICK_DDA, bounce_cap=256):
    vs = 1.0 / size[0]
    svd = (256, 128, 256) if tuple(size) == (1024, 512, 1024) else (max(2, size[0] // 4), max(2, size[1] // 4), max(2, size[2] // 4))
    return Scene("c2", fbm(tuple(size), vs, seed=7, octaves=8, shape=S.CR_SHAPE_ANVIL, lo=0.16, hi=0.36),
                 storage, layout, 1.0, 64.0, camera(width, height, position=(0.0, 0.0, -0.15)), parameters(0.85),
                 render_ext(albedo=0.999, bounce_cap=bounce_cap, sky=SKY, nee=nee, tracking=tracking, sv_dims=svd),
                 256, spp)


def c4(n_jobs=64, seed=4, grid_side=(64, 512), image_side=(128, 1024), spp=(16, 512), camera_z=-0.2):
    """The jobs of a batched render (a list of Scene), drawn from a seeded RNG."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def log_uniform(lo, hi):
        return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))

    jobs = []
    for k in range(n_jobs):
        side = log_uniform(*grid_side)
        img = log_uniform(*image_side)
        img += img & 1  # the reference's images are 2 halfWidth x 2 halfHeight
        n = log_uniform(*spp)
        g = float(rng.uniform(0.6, 0.9))
        cos_t = float(rng.uniform(0.0, 1.0))   # sun above the horizon: light travels downward
        phi = float(rng.uniform(0.0, 2.0 * math.pi))
        sin_t = math.sqrt(max(0.0, 1.0 - cos_t * cos_t))
        light = (sin_t * math.cos(phi), -cos_t, sin_t * math.sin(phi))
        if cos_t < 1e-3:
            light = (light[0], -1e-3, light[2])
        cx, cy = (float(v) for v in rng.uniform(-0.1, 0.1, 2))
        jobs.append(Scene(f"c4/{k}", fbm((side, side, side), 1.0 / side, seed=100 + k, octaves=6), S.CR_DTYPE_F16,
                          S.CR_GRID_BRICK4, 1.0, 64.0, camera(img, img, position=(cx, cy, camera_z)), parameters(g),
                          render_ext(albedo=0.9, bounce_cap=16, sky=SKY, nee=S.CR_NEE_RATIO_TRACKING,
                                     tracking=S.CR_TRACK_BRICK_DDA), 64, n, light_dir=light))
    return jobs


def beam_params(num_paths, seed=38, radius=0.002, albedo=0.9, light_dir=(1.0, -1.0, 0.0), intensity=5.0,
                max_interactions=64):
    p = S.BeamTraceParams()
    load().cr_beam_trace_params_default(C.byref(p))
    n = math.sqrt(sum(x * x for x in light_dir))
    p.numPaths = num_paths
    p.seed = seed
    p.radius = radius
    p.albedo = albedo
    p.lightDirection[:] = (light_dir[0] / n, light_dir[1] / n, light_dir[2] / n, 0.0)
    p.lightIntensity = intensity
    p.maxInteractions = max_interactions
    return p
