"""Batched rendering: many path-tracing jobs of uneven sizes in one persistent
launch (recorded into a HIP graph), checked pixel for pixel against the oracle
(oracle/oracle_pt.cpp, PathTracer.comp:318-361, 366-425 with the brick-DDA and
ratio-tracking extensions) and against rendering each job alone with
cr_pt_render.  Covers the job table (uneven grids, images, frame counts, phase
functions, sun directions), the launch shapes the autotuner chooses among,
graph replays with new frame lists, and the error paths.
"""
import numpy as np
import pytest

from gpu_util import assert_bitexact, oracle_scene_from_ctx

pytestmark = pytest.mark.gpu


def small_jobs(cfg, n=5, seed=11):
    """A scaled-down c4: grids 12..40 voxels, images 8..40 px, 1..6 frames."""
    return cfg.c4(n_jobs=n, seed=seed, grid_side=(12, 40), image_side=(8, 40), spp=(1, 6))


def render_alone(ctx, sc, frames, majorant_brick=None, O=None):
    """The job rendered alone with cr_pt_render; with O, also the oracle's
    render of the same job (same grid, majorants and uniforms)."""
    if majorant_brick is not None:
        sc.ext.majorantBrick = majorant_brick
    _, svp = sc.setup(ctx, build_sv=False)
    ctx.pt_render(frames)
    img = ctx.download_image(sc.height, sc.width)
    if O is None:
        return img
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext, with_sv=False)
    return img, O.pt_render(s, frames)[0]


def build_batch(ctx, cfg, jobs, first=0, O=None):
    """Render every job alone (the expected images) and, with O, with the
    oracle; then move the job into the batch."""
    ctx.batch_clear()
    expected, oracle = [], []
    for sc in jobs:
        fr = cfg.frames(sc.spp, first=first)
        if O is None:
            expected.append(render_alone(ctx, sc, fr))
        else:
            img, ref = render_alone(ctx, sc, fr, O=O)
            expected.append(img)
            oracle.append(ref)
        ctx.batch_add(fr)
    return (expected, oracle) if O is not None else expected


@pytest.mark.parametrize("shape", [(256, 0, 8), (128, 2, 16), (64, 1, 32)])
def test_batch_equals_oracle_and_single_renders(ctx, cfg, S, O, shape):
    """Every job of a scaled-down batched workload (FP16 bricks, brick DDA,
    ratio-tracked shadow rays, g in [0.6, 0.9], random suns) is bit-exact
    against the oracle and against its cr_pt_render."""
    jobs = small_jobs(cfg)
    assert all(sc.storage == S.CR_DTYPE_F16 and sc.grid_layout == S.CR_GRID_BRICK4 and
               sc.ext.trackingMode == S.CR_TRACK_BRICK_DDA and sc.ext.neeMode == S.CR_NEE_RATIO_TRACKING
               for sc in jobs)
    expected, oracle = build_batch(ctx, cfg, jobs, O=O)
    assert ctx.batch_size() == len(jobs)
    ctx.batch_set_tuning(*shape)
    ctx.batch_stats_reset()
    ctx.batch_render()
    for k, sc in enumerate(jobs):
        got = ctx.batch_download(k)
        assert got.shape == (sc.height, sc.width, 4)
        what = f"job {k} ({sc.fbm.size[0]}^3 grid, {sc.width}^2 px, {sc.spp} frames)"
        assert_bitexact(got, oracle[k], what + " vs the oracle")
        assert_bitexact(got, expected[k], what + " vs cr_pt_render")
    st = ctx.batch_stats()
    assert st.paths == sum(sc.width * sc.height * sc.spp for sc in jobs)
    assert st.launches == 1
    ctx.batch_clear()


def test_batch_c4_full_size_vs_oracle(ctx, cfg, S, O):
    """The 64-job batched workload at its real sizes (FP16 grids 64^3..512^3,
    images 128^2..1024^2, 16..512 frames, one launch): a 16x16-pixel block at
    the centre of every job's image is bit-exact against the oracle's render
    of the same job over all of its frames."""
    jobs = cfg.c4()
    assert len(jobs) == 64
    ctx.batch_clear()
    refs, blocks = [], []
    for k, sc in enumerate(jobs):
        _, svp = sc.setup(ctx, build_sv=False)
        fr = cfg.frames(sc.spp)
        s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext, with_sv=False)
        w = h = 16
        x0, y0 = (sc.width - w) // 2, (sc.height - h) // 2
        refs.append(O.pt_render_region(s, fr, x0, y0, w, h)[0])
        blocks.append((x0, y0, w, h))
        del s
        ctx.batch_add(fr)
    ctx.batch_stats_reset()
    ctx.batch_render()
    st = ctx.batch_stats()
    assert st.launches == 1 and st.paths == sum(sc.width * sc.height * sc.spp for sc in jobs)
    for k, sc in enumerate(jobs):
        x0, y0, w, h = blocks[k]
        got = ctx.batch_download(k)
        assert np.isfinite(got).all()
        assert_bitexact(got[y0:y0 + h, x0:x0 + w], refs[k],
                        f"job {k} ({sc.fbm.size[0]}^3 grid, {sc.width}^2 px, {sc.spp} frames)")
    ctx.batch_clear()


def test_batch_graph_replay_with_new_frames(ctx, cfg):
    """Second render (graph replay) with the next frames of every job equals
    rendering all frames of the job alone in one launch."""
    jobs = small_jobs(cfg, n=4, seed=5)
    ctx.batch_clear()
    for sc in jobs:
        sc.setup(ctx, build_sv=False)
        ctx.batch_add(cfg.frames(sc.spp))
    ctx.batch_render()
    for k, sc in enumerate(jobs):
        ctx.batch_set_frames(k, cfg.frames(2, first=sc.spp))
    ctx.batch_render()
    for k, sc in enumerate(jobs):
        exp = render_alone(ctx, sc, cfg.frames(sc.spp + 2))
        assert_bitexact(ctx.batch_download(k), exp, f"job {k} after replay")
    assert ctx.batch_stats().launches >= 2
    ctx.batch_clear()


def test_batch_autotune(ctx, cfg):
    jobs = small_jobs(cfg, n=3, seed=21)
    ctx.batch_clear()
    for sc in jobs:
        sc.setup(ctx, build_sv=False)
        ctx.batch_add(cfg.frames(sc.spp))
    t, rep = ctx.batch_autotune(True)
    assert t.blockThreads in (64, 128, 256) and t.tileWidth in (8, 16, 32) and t.blocksPerCU >= 1
    assert rep.candidates >= 3 and 0.0 < rep.bestMs <= rep.defaultMs
    got_t = ctx.batch_tuning()
    assert (got_t.blockThreads, got_t.blocksPerCU, got_t.tileWidth) == (t.blockThreads, t.blocksPerCU, t.tileWidth)
    ctx.batch_clear_images()
    ctx.batch_render()
    for k, sc in enumerate(jobs):
        b = rep.jobBrick[k]
        assert b in (8, 16, 32)
        exp = render_alone(ctx, sc, cfg.frames(sc.spp), majorant_brick=b)
        assert_bitexact(ctx.batch_download(k), exp, f"job {k} with the tuned {b}^3 bricks")
    ctx.batch_clear()


def test_batch_errors(ctx, cfg, S, pkg):
    lib = pkg.load()
    ctx.batch_clear()
    assert lib.cr_batch_render(ctx.h) == S.CR_ERR_STATE
    jobs = small_jobs(cfg, n=2, seed=3)
    jobs[0].setup(ctx, build_sv=False)
    ctx.batch_add(cfg.frames(1))
    # the context no longer holds a grid
    fr = (S.PushConstants * 1)(*cfg.frames(1))
    assert lib.cr_batch_add(ctx.h, fr, 1, None) == S.CR_ERR_STATE
    # a job with another NEE mode cannot join
    jobs[1].ext.neeMode = S.CR_NEE_NONE
    jobs[1].setup(ctx, build_sv=False)
    assert lib.cr_batch_add(ctx.h, fr, 1, None) == S.CR_ERR_INVALID_ARGUMENT
    bad = S.BatchTuning(96, 0, 8, 0)
    assert lib.cr_batch_set_tuning(ctx.h, bad) == S.CR_ERR_INVALID_ARGUMENT
    assert lib.cr_batch_download(ctx.h, 7, None, None, None) == S.CR_ERR_INVALID_ARGUMENT
    ctx.batch_clear()
    assert ctx.batch_size() == 0


def test_batch_capacity(ctx, cfg, S, pkg):
    lib = pkg.load()
    ctx.batch_clear()
    sc = small_jobs(cfg, n=1, seed=9)[0]
    sc.fbm = cfg.fbm((8, 8, 8), 1.0 / 8, seed=1)
    sc.cam = cfg.camera(8, 8)
    fr = cfg.frames(1)
    for _ in range(S.CR_BATCH_MAX_JOBS):
        sc.setup(ctx, build_sv=False)
        ctx.batch_add(fr)
    sc.setup(ctx, build_sv=False)
    arr = (S.PushConstants * 1)(*fr)
    assert lib.cr_batch_add(ctx.h, arr, 1, None) == S.CR_ERR_CAPACITY
    ctx.batch_render()
    ref = ctx.batch_download(0)
    for k in range(1, S.CR_BATCH_MAX_JOBS, 9):
        assert_bitexact(ctx.batch_download(k), ref, f"identical job {k}")
    ctx.batch_clear()
