"""GPU parity of the grid generator, shadow volume and path tracer.

Every comparison is bit-exact: the kernels evaluate the same IEEE binary32
operation sequence and RNG draws as the oracle (oracle/oracle_pt.cpp,
oracle/oracle_grid.cpp), which restates shaders/PathTracer.comp and
shaders/ShadowVolume.comp.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import assert_bitexact, oracle_scene_from_ctx

pytestmark = pytest.mark.gpu


def small_c1(cfg, S, w=40, h=24, size=48, sv_axis=40, **ext_kw):
    sc = cfg.c1(width=w, height=h, size=size, sv_axis=sv_axis)
    ext = dict(albedo=0.9, bounce_cap=16, sky=cfg.SKY)
    ext.update(ext_kw)
    sc.ext = cfg.render_ext(**ext)
    return sc


def skip_reference(maj, ds, bs):
    """(majorant * ds) / bs in float32 for a non-empty brick, -min(16, max-norm
    distance to the nearest non-empty brick) for an empty one (brute force)."""
    occ = np.argwhere(maj > 0)
    out = ((maj.astype(np.float32) * np.float32(ds)) / np.float32(bs)).astype(np.float32)
    for idx in np.argwhere(maj == 0):
        d = 16 if len(occ) == 0 else min(16, int(np.abs(occ - idx).max(axis=1).min()))
        out[tuple(idx)] = -float(d)
    return out


@pytest.mark.parametrize("storage,layout", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_fbm_grid_bitexact(ctx, cfg, S, O, storage, layout):
    p = cfg.fbm((37, 29, 23), 1.0 / 37, seed=5, octaves=6)
    ctx.set_render_ext(cfg.render_ext())
    ctx.set_grid_layout(layout)
    ctx.generate_fbm(p, storage)
    gpu = ctx.download_grid()
    ref = O.fbm_grid(p)
    if storage == S.CR_DTYPE_F16:
        ref = ref.astype(np.float16).astype(np.float32)  # numpy rounds to nearest even
    assert_bitexact(gpu, ref, f"fbm grid storage={storage} layout={layout}")
    maj, _ = O.brick_majorants(ref)
    assert_bitexact(ctx.majorants(), maj, "brick majorants")
    for b in (4, 8, 32):
        ctx.set_render_ext(cfg.render_ext(majorant_brick=b))
        maj, _ = O.brick_majorants(ref, b)
        assert_bitexact(ctx.majorants(), maj, f"brick majorants, {b}^3 bricks")
        cl = ctx.cloud()
        assert_bitexact(ctx.skip_grid(), skip_reference(maj, cl.densityScaling, cl.baseScaling), f"skip grid, {b}^3 bricks")
    # a new density scale rescales the non-empty bricks only
    cl = ctx.cloud()
    cl.densityScaling = 37.5
    ctx.set_cloud(cl)
    assert_bitexact(ctx.skip_grid(), skip_reference(maj, 37.5, cl.baseScaling), "skip grid after cr_set_cloud")
    c = ctx.cloud()
    assert c.maxExtinction == max(float(ref.max()), 0.01)
    assert tuple(c.voxelCount) == (37, 29, 23, 0)


def test_grid_upload_roundtrip(ctx, S, tmp_path):
    rng = np.random.default_rng(3)
    g = rng.random((7, 6, 5), dtype=np.float32)
    ctx.set_grid_layout(S.CR_GRID_BRICK4)
    ctx.upload_grid(g, 0.01)
    assert_bitexact(ctx.download_grid(), g, "upload x-fastest")
    # .xyz file order: z fastest (Grid3D<float>::Load transposes, Grid3D.h:73-85)
    ctx.upload_grid(np.ascontiguousarray(g.transpose(2, 1, 0)), 0.01, layout=S.CR_LAYOUT_Z_FASTEST)
    assert_bitexact(ctx.download_grid(), g, "upload z-fastest")
    path = str(tmp_path / "g.xyz")
    ctx.save_xyz(path)
    ctx.set_grid_layout(S.CR_GRID_DENSE)
    ctx.load_xyz(path)
    assert_bitexact(ctx.download_grid(), g, "save/load .xyz")
    c = ctx.cloud()
    # SetCloudProperties (main.cpp:144-163) with baseScaling from the context
    sx = np.float32(0.01 * 5 * c.baseScaling)
    assert c.bounds[0][0] == -sx / 2 and c.bounds[1][0] == sx / 2 and c.bounds[0][2] == 0.0
    assert c.maxExtinction == max(float(g.max()), 0.01)


@pytest.mark.parametrize("dims", [(0, 0, 0), (24, 17, 30)])
def test_shadow_volume_bitexact(ctx, cfg, S, O, dims):
    sc = small_c1(cfg, S, sv_axis=33, sv_dims=dims)
    cloud, svp = sc.setup(ctx)
    gpu = ctx.sv_download()
    s, grid, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext, with_sv=False)
    ref = O.sv_build(s)
    assert gpu.shape == ref.shape
    assert_bitexact(gpu, ref, f"shadow volume dims={dims}")
    assert 0.0 < gpu.min() and gpu.max() <= 1.0


def render_both(ctx, O, sc, svp, frames, init=None):
    if init is not None:
        ctx.upload_image(init)
    else:
        ctx.clear_image()
    ctx.pt_stats_reset()
    ctx.pt_render(frames)
    gpu = ctx.download_image(sc.height, sc.width)
    st = ctx.pt_stats()
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    ref, cnt = O.pt_render(s, frames, accum=init)
    return gpu, ref, st, cnt


def test_pt_reference_mode_bitexact(ctx, cfg, S, O):
    """Reference semantics: albedo 1, no bounce cap, gradient sky, shadow-volume
    NEE, global majorant, frameCount starting at 1 over a non-empty image
    (PathTracer.comp:366-425, main.cpp:262)."""
    sc = small_c1(cfg, S, w=70, h=30, size=40, sv_axis=36)
    sc.ext = cfg.render_ext()
    sc.params = cfg.parameters(0.85)
    sc.fbm.thresholdLo, sc.fbm.thresholdHi = 0.3, 0.45  # thinner medium: albedo-1 walks stay short
    cloud, svp = sc.setup(ctx)
    init = np.random.default_rng(1).random((sc.height, sc.width, 4), dtype=np.float32)
    fr = cfg.frames(3, ref_frame_count=True)
    gpu, ref, st, cnt = render_both(ctx, O, sc, svp, fr, init)
    assert_bitexact(gpu, ref, "reference-mode image")
    assert (st.paths, st.vertices, st.lookups, st.svLookups) == (cnt.paths, cnt.vertices, cnt.lookups, cnt.svLookups)
    assert st.paths == sc.width * sc.height * 3


@pytest.mark.parametrize("variant", ["c1", "ratio", "dda", "dda_ratio", "no_nee", "isotropic", "f16_brick", "dda_b8",
                                     "dda_b32"])
def test_pt_extensions_bitexact(ctx, cfg, S, O, variant):
    kw = {}
    if variant in ("ratio", "dda_ratio"):
        kw["nee"] = S.CR_NEE_RATIO_TRACKING
    if variant in ("dda", "dda_ratio", "f16_brick", "dda_b8", "dda_b32"):
        kw["tracking"] = S.CR_TRACK_BRICK_DDA
    if variant in ("dda_b8", "dda_b32"):
        kw["majorant_brick"] = int(variant[5:])
    if variant == "no_nee":
        kw["nee"] = S.CR_NEE_NONE
    sc = small_c1(cfg, S, **kw)
    if variant == "isotropic":
        sc.params = cfg.parameters(0.0)
        assert sc.params.isotropic == 1
    if variant == "f16_brick":
        sc.storage, sc.grid_layout = S.CR_DTYPE_F16, S.CR_GRID_BRICK4
    cloud, svp = sc.setup(ctx)
    fr = cfg.frames(4)
    gpu, ref, st, cnt = render_both(ctx, O, sc, svp, fr)
    assert_bitexact(gpu, ref, f"image ({variant})")
    assert (st.paths, st.vertices, st.lookups, st.svLookups, st.brickVisits) == \
        (cnt.paths, cnt.vertices, cnt.lookups, cnt.svLookups, cnt.brickVisits)
    assert np.isfinite(gpu).all()


def test_fbm_anvil_bitexact(ctx, cfg, S, O):
    """The production grid generator (8-octave anvil profile, FP16 in 4^3
    bricks, 16^3 brick majorants: the 1024x512x1024 workload's settings) at
    96x48x96 equals the oracle's generator voxel for voxel, and so do its brick
    majorants and skip grid."""
    sc = cfg.c2(width=32, height=18, size=(96, 48, 96))
    assert sc.fbm.shape == S.CR_SHAPE_ANVIL and sc.fbm.octaves == 8
    sc.setup(ctx, build_sv=False)
    gpu = ctx.download_grid()
    ref = O.fbm_grid(sc.fbm).astype(np.float16).astype(np.float32)
    assert_bitexact(gpu, ref, "anvil fBm grid (FP16, 4^3 bricks)")
    assert 0.0 < float(ref.mean()) < 0.5 and ref.max() == 1.0
    maj, _ = O.brick_majorants(ref, 16)
    assert_bitexact(ctx.majorants(), maj, "16^3 brick majorants")
    cl = ctx.cloud()
    assert_bitexact(ctx.skip_grid(), skip_reference(maj, cl.densityScaling, cl.baseScaling), "skip grid")


def test_fbm_anvil_f32_bitexact(ctx, cfg, S, O):
    """The anvil generator's float32 values, before any FP16 rounding."""
    sc = cfg.c2(width=32, height=18, size=(96, 48, 96), storage=S.CR_DTYPE_F32, layout=S.CR_GRID_DENSE)
    sc.setup(ctx, build_sv=False)
    assert_bitexact(ctx.download_grid(), O.fbm_grid(sc.fbm), "anvil fBm grid (FP32)")


def test_f16_storage_rounds_ties_to_even(ctx, S):
    """FP16 storage rounds to nearest, ties to even: every value halfway
    between two consecutive positive FP16 numbers (subnormal and normal), and
    its float32 neighbours, against numpy's IEEE conversion."""
    h = np.arange(0, 0x7BFF, dtype=np.uint16)
    lo = h.view(np.float16).astype(np.float64)
    hi = (h + 1).view(np.float16).astype(np.float64)
    mid = ((lo + hi) / 2).astype(np.float32)
    vals = np.concatenate([mid, np.nextafter(mid, np.float32(0)), np.nextafter(mid, np.float32(1e9))])
    n = len(vals)
    pad = (-n) % 1024
    vals = np.concatenate([vals, np.zeros(pad, np.float32)]).reshape(-1, 32, 32)
    for layout in (S.CR_GRID_DENSE, S.CR_GRID_BRICK4):
        ctx.set_grid_layout(layout)
        ctx.upload_grid(vals, 0.01, storage=S.CR_DTYPE_F16)
        assert_bitexact(ctx.download_grid(), vals.astype(np.float16).astype(np.float32), f"FP16 rounding, layout {layout}")


def c2_small(cfg, w=48, h=28, size=(64, 32, 64)):
    """The production workload's render settings (FP16 anvil in 4^3 bricks,
    16^3 brick-majorant DDA, ratio-tracked shadow rays, albedo 0.999, <= 256
    bounces, per-axis shadow-volume dims) on a smaller grid and image."""
    sc = cfg.c2(width=w, height=h, size=size)
    return sc


@pytest.mark.parametrize("n_frames", [4, 20], ids=["few_frame_launch", "many_frame_launch"])
def test_pt_c2_shape_bitexact(ctx, cfg, S, O, n_frames):
    """The benchmarked kernel instantiation (FP16 + 4^3 bricks + brick DDA +
    ratio tracking) with the benchmark's albedo, bounce cap and generator,
    against the oracle (PathTracer.comp:318-361, 366-425 plus the extensions).
    Launches of at most 16 frames run the instantiation with gated pixel
    claims, longer ones (the benchmark's 256) the other: both are checked."""
    sc = c2_small(cfg)
    assert sc.storage == S.CR_DTYPE_F16 and sc.grid_layout == S.CR_GRID_BRICK4
    assert sc.ext.trackingMode == S.CR_TRACK_BRICK_DDA and sc.ext.neeMode == S.CR_NEE_RATIO_TRACKING
    assert sc.ext.bounceCap == 256 and abs(sc.ext.albedo - 0.999) < 1e-7
    cloud, svp = sc.setup(ctx)
    fr = cfg.frames(n_frames)
    gpu, ref, st, cnt = render_both(ctx, O, sc, svp, fr)
    assert_bitexact(gpu, ref, "c2-shaped image")
    assert (st.paths, st.vertices, st.lookups, st.svLookups, st.brickVisits) == \
        (cnt.paths, cnt.vertices, cnt.lookups, cnt.svLookups, cnt.brickVisits)
    # the medium is dense enough that paths scatter many times
    assert cnt.vertices > 3 * cnt.paths and cnt.brickVisits > cnt.lookups


def test_pt_c2_full_size_region_bitexact(ctx, cfg, S, O):
    """The 1024x512x1024 FP16 workload itself: sampled voxels of the generated
    1 GiB grid equal the oracle's generator, the 64x32x64 brick majorants equal
    the oracle's over the downloaded grid, and a 64x48-pixel block at the
    centre of two 1920x1080 frames is bit-exact against the oracle."""
    import ctypes as C
    sc = cfg.c2()
    cloud, svp = sc.setup(ctx)
    grid = ctx.download_grid()
    assert grid.shape == (1024, 512, 1024)
    rng = np.random.default_rng(12)
    idx = rng.integers(0, (1024, 512, 1024), size=(4000, 3))
    p = sc.fbm
    vals = np.array([O.lib.or_fbm_voxel(C.byref(p), int(x), int(y), int(z)) for x, y, z in idx], np.float32)
    assert_bitexact(grid[idx[:, 2], idx[:, 1], idx[:, 0]], vals.astype(np.float16).astype(np.float32),
                    "sampled voxels of the 1 GiB anvil grid")
    maj = ctx.majorants()
    assert maj.shape == (64, 32, 64)
    assert_bitexact(maj, O.brick_majorants(grid, 16)[0], "64x32x64 brick majorants")
    ctx.clear_image()
    fr = cfg.frames(2)
    ctx.pt_render(fr)
    gpu = ctx.download_image(1080, 1920)
    g16 = grid.astype(np.float16)
    del grid
    s = O.scene(g16, ctx.cloud(), sc.cam, sc.params, svp, sc.ext, sv=ctx.sv_download(), majorants=maj)
    x0, y0, w, h = 928, 516, 64, 48
    ref, cnt = O.pt_render_region(s, fr, x0, y0, w, h)
    assert_bitexact(gpu[y0:y0 + h, x0:x0 + w], ref, "1080p block of the 1 GiB workload")
    assert cnt.vertices > 2 * cnt.paths
    assert np.isfinite(gpu).all()


def test_pt_frames_split_equals_single_launch(ctx, cfg, S):
    """n frames in one persistent launch == the same frames one launch each
    (20 frames: the one launch runs the many-frame instantiation, the single
    frames the one with gated pixel claims)."""
    sc = small_c1(cfg, S, w=32, h=32)
    sc.setup(ctx)
    fr = cfg.frames(20)
    ctx.clear_image()
    ctx.pt_render(fr)
    a = ctx.download_image(32, 32)
    ctx.clear_image()
    for f in fr:
        ctx.pt_render([f])
    b = ctx.download_image(32, 32)
    assert_bitexact(a, b, "split launches")


def test_pt_white_furnace(ctx, cfg, S):
    """lightIntensity 0, constant sky c, albedo 1: every path escapes with
    weight 1, so every pixel equals c exactly for any density."""
    sc = small_c1(cfg, S, w=24, h=24, albedo=1.0, bounce_cap=S.CR_UNBOUNDED)
    sc.params = cfg.parameters(0.85, light_intensity=0.0)
    sc.fbm.thresholdLo, sc.fbm.thresholdHi = 0.3, 0.45
    sc.setup(ctx)
    ctx.pt_render(cfg.frames(2))
    img = ctx.download_image(24, 24)
    np.testing.assert_array_equal(img, np.broadcast_to(np.array(cfg.SKY, np.float32), img.shape))


def test_pt_errors(ctx, cfg, S, pkg):
    sc = small_c1(cfg, S)
    sc.setup(ctx, build_sv=False)
    ctx.set_shadow_volume(cfg.shadow_volume(ctx.cloud(), (1, -1, 0), 16))
    with pytest.raises(pkg.CrError) as e:
        ctx.pt_render(cfg.frames(1))
    assert e.value.status == S.CR_ERR_STATE
    bad = cfg.render_ext(albedo=1.5)
    with pytest.raises(pkg.CrError) as e:
        ctx.set_render_ext(bad)
    assert e.value.status == S.CR_ERR_INVALID_ARGUMENT
