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


def test_pt_frames_split_equals_single_launch(ctx, cfg, S):
    """n frames in one persistent launch == the same frames one launch each."""
    sc = small_c1(cfg, S, w=32, h=32)
    sc.setup(ctx)
    fr = cfg.frames(6)
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
