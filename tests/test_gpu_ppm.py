"""GPU parity of photon mapping (RenderTechniquePPM) with the oracle
(oracle/oracle_ppm.cpp, which restates shaders/PPM_PT.comp and
shaders/PPM_PE.comp): the deposits and the photon grid byte for byte, the
estimate's image bit for bit, and the technique's per-frame loop (radius
schedule, a new photon map per frame seed)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import assert_bitexact, oracle_scene_from_ctx

pytestmark = pytest.mark.gpu


def setup(ctx, cfg, S, pkg, storage=None, layout=None, w=40, h=24, size=32, cells=100):
    """The c1 cloud at reduced size with the photon grid on its box (SetBounds)
    and the reference's photon-map scales divided by its 1000-unit scene:
    march step and initial radius 1/100 of the box, sky-ray step 1/10."""
    lib = pkg.load()
    sc = cfg.c1(width=w, height=h, size=size, sv_axis=48)
    sc.ext = cfg.render_ext(sky=cfg.SKY)
    if storage is not None:
        sc.storage, sc.grid_layout = storage, layout
    cloud, svp = sc.setup(ctx)
    pm = S.PhotonMapProperties()
    lib.cr_photon_map_default(C.byref(pm))
    pm.voxelCount[:] = (cells, cells, cells)
    lib.cr_photon_map_set_bounds(C.byref(pm), cloud.bounds[0], cloud.bounds[1])
    pm.stepSize = 0.01
    ctx.set_photon_map(pm)
    p = S.PpmParams()
    lib.cr_ppm_params_default(C.byref(p))
    p.numPaths = 4000
    p.ambientStep = 0.1
    return sc, cloud, svp, pm, p


def oracle_map(O, ctx, sc, svp, pm, p, seed):
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    photons, cells, n = O.ppm_trace(s, pm, p, seed)
    n_cells = pm.voxelCount[0] * pm.voxelCount[1] * pm.voxelCount[2]
    arrivals, start, sorted_ = O.ppm_build(photons, cells, n, n_cells)
    return s, n, arrivals, start, sorted_


@pytest.mark.parametrize("variant", ["f32_dense", "f16_brick", "coarse", "long_paths"])
def test_ppm_trace_bitexact(ctx, cfg, S, O, pkg, variant):
    """coarse: a 12^3 photon grid, where many cells receive more than the 32
    photons they keep.  long_paths: a 30 times denser cloud and no absorption,
    so paths scatter until they leave the box, with more than 64 deposits per
    path on average."""
    storage = (S.CR_DTYPE_F16, S.CR_GRID_BRICK4) if variant == "f16_brick" else (None, None)
    sc, cloud, svp, pm, p = setup(ctx, cfg, S, pkg, *storage, cells=12 if variant == "coarse" else 100)
    if variant == "long_paths":
        cloud.densityScaling *= 30.0
        ctx.set_cloud(cloud)
        ctx.sv_build()
        pm.absorption = 0.0
        ctx.set_photon_map(pm)
        p.numPaths = 400
    n = ctx.ppm_trace(38, p)
    arrivals, photons, nd = ctx.ppm_download()
    s, rn, rarr, rstart, rsorted = oracle_map(O, ctx, sc, svp, pm, p, 38)
    assert nd == n == rn and n > 10000
    if variant == "long_paths":
        assert n > 64 * p.numPaths
    np.testing.assert_array_equal(arrivals, rarr)
    assert bytes(photons)[:48 * n] == bytes(rsorted)[:48 * n], "photon grid differs"
    if variant == "coarse":
        assert (arrivals > 32).sum() > 100


@pytest.mark.parametrize("cells", [100, 12])
def test_ppm_estimate_bitexact(ctx, cfg, S, O, pkg, cells):
    """The estimate over the photon grid, bit for bit; on the 12^3 grid the
    query sphere meets cells that keep only their first 32 photons."""
    sc, cloud, svp, pm, p = setup(ctx, cfg, S, pkg, cells=cells)
    ctx.ppm_trace(7719, p)
    s, rn, rarr, rstart, rsorted = oracle_map(O, ctx, sc, svp, pm, p, 7719)
    m = O.photon_map(rarr, rstart, rsorted)
    init = np.random.default_rng(3).random((sc.height, sc.width, 4), dtype=np.float32)
    ctx.upload_image(init)
    seeds = cfg.msvc_seeds(3)
    frames = [S.PushConstants(0.0, seeds[i], 1 + i, 0.01, 0) for i in range(3)]
    ctx.ppm_estimate(frames, p)
    gpu = ctx.download_image(sc.height, sc.width)
    ref = O.ppm_estimate(s, pm, m, p, frames, accum=init)
    assert_bitexact(gpu, ref, "photon-mapping estimate")
    assert np.isfinite(gpu).all() and not np.array_equal(gpu, init)


def test_ppm_technique_frames(ctx, cfg, S, O, pkg):
    """RenderTechniquePPM.RecordDrawCommands per frame: the radius schedule of
    UpdateRadius, a photon map traced with the frame's seed, the estimate
    blended in -- equal to the oracle doing the same steps."""
    sc, cloud, svp, pm, p = setup(ctx, cfg, S, pkg, w=24, h=16)
    tech = pkg.RenderTechniquePPM(ctx, None, initialRadius=0.01, params=p)
    ctx.clear_image()
    frames = cfg.frames(3, ref_frame_count=True)
    tech.RecordDrawCommands(frames)
    gpu = ctx.download_image(sc.height, sc.width)
    lib = pkg.load()
    acc = np.zeros((sc.height, sc.width, 4), np.float32)
    r = np.float32(0.01)
    for f in frames:
        r = np.float32(lib.cr_ppm_update_radius(f.frameCount, S.CR_PPM_ALPHA, 0.01, float(r)))
        s, rn, rarr, rstart, rsorted = oracle_map(O, ctx, sc, svp, pm, p, f.seed)
        fr = S.PushConstants(0.0, f.seed, f.frameCount, float(r), 0)
        acc = O.ppm_estimate(s, pm, O.photon_map(rarr, rstart, rsorted), p, [fr], accum=acc)
    assert_bitexact(gpu, acc, "three frames of the PPM technique")
    assert tech.radius == r < 0.01


def test_ppm_errors(ctx, cfg, S, pkg):
    lib = pkg.load()
    sc, cloud, svp, pm, p = setup(ctx, cfg, S, pkg)
    ctx.set_photon_map(pm)                      # a new grid: no map until traced
    fr = (S.PushConstants * 1)(S.PushConstants(0.0, 38, 0, 0.01, 0))
    assert lib.cr_ppm_estimate(ctx.h, fr, 1, C.byref(p)) == S.CR_ERR_STATE
    ctx.ppm_trace(38, p)
    bad = (S.PushConstants * 1)(S.PushConstants(0.0, 38, 0, 0.0, 0))
    assert lib.cr_ppm_estimate(ctx.h, bad, 1, C.byref(p)) == S.CR_ERR_INVALID_ARGUMENT
    pm2 = S.PhotonMapProperties.from_buffer_copy(pm)
    pm2.voxelCount[0] = 0
    assert lib.cr_set_photon_map(ctx.h, C.byref(pm2)) == S.CR_ERR_INVALID_ARGUMENT
    p2 = S.PpmParams.from_buffer_copy(p)
    p2.numPaths = 0
    assert lib.cr_ppm_trace(ctx.h, 1, C.byref(p2), None) == S.CR_ERR_INVALID_ARGUMENT
