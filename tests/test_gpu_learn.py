"""Learned in-scattering renderer on the GPU.

* record + continue(spp = 1) reproduces cr_pt_render bit for bit (the two
  passes are the path tracer cut at the first collision);
* the training data set is exactly the continuation targets and the stencil
  features of the recorded collisions, in record order;
* with shadow-volume NEE the first collision's next-event term is left out of
  the targets and added back exactly by the learned render (bit for bit);
* a render of n frames in one call equals the same frames over several calls;
* stencil features against the NumPy oracle (bf16 tolerance);
* the gain: a network trained on path-traced samples, rendered at 4 spp, has
  a lower relMSE against a converged path-traced reference than the plain
  path tracer at the same 4 spp.
"""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

from gpu_util import assert_bitexact

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bf16_to_f32(u):
    return (np.asarray(u, np.uint32) << 16).view(np.float32)


def scene(cfg, S, size=40, w=48, h=40, seed=1, nee=None, tracking=None, light=(1.0, -1.0, 0.0)):
    sc = cfg.c1(width=w, height=h, size=size, sv_axis=48)
    sc.fbm.seed = seed
    sc.light_dir = light
    sc.ext = cfg.render_ext(albedo=0.9, bounce_cap=16, sky=cfg.SKY,
                            nee=S.CR_NEE_RATIO_TRACKING if nee is None else nee,
                            tracking=S.CR_TRACK_BRICK_DDA if tracking is None else tracking)
    return sc


def compose_host(records, cont, frames):
    """The running mean of compose_kernel / PathTracer.comp:420-424 in float32."""
    F, H, W, _ = records.shape
    acc = np.zeros((H, W, 4), np.float32)
    for f in range(F):
        rec = records[f]
        L = np.where(rec[..., 7:8] > 0.5, cont[f], rec[..., :4]).astype(np.float32)
        n = np.float32(frames[f].frameCount)
        d = np.float32(frames[f].frameCount + 1)
        acc = ((L + acc * n) / d).astype(np.float32)
    return acc


@pytest.mark.parametrize("variant", ["dda_ratio", "global_sv"])
def test_record_continue_reproduces_render(ctx, cfg, S, variant):
    if variant == "dda_ratio":
        sc = scene(cfg, S)
    else:
        sc = scene(cfg, S, nee=S.CR_NEE_SHADOW_VOLUME, tracking=S.CR_TRACK_GLOBAL_MAJORANT)
    sc.setup(ctx)
    fr = cfg.frames(3)
    ctx.pt_render(fr)
    ref = ctx.download_image(sc.height, sc.width)
    rec = ctx.pt_record(fr, sc.height, sc.width)
    flags = rec[..., 7]
    assert set(np.unique(flags)) <= {0.0, 1.0} and 0 < flags.sum() < flags.size
    cont = ctx.pt_continue(rec, 1).reshape(rec.shape[:3] + (4,))
    assert np.all(cont[flags == 0] == 0.0)
    assert_bitexact(compose_host(rec, cont, fr), ref, f"record + continue ({variant})")


def test_collected_dataset_is_targets_and_features(ctx, cfg, S):
    sc = scene(cfg, S, w=24, h=20, seed=3)
    sc.setup(ctx)
    ctx.learn_dataset_clear()
    fr = cfg.frames(2)
    n = ctx.learn_collect(fr, 8, 0)
    rec = ctx.pt_record(fr, sc.height, sc.width).reshape(-1, 8)
    hits = rec[rec[:, 7] > 0.5]
    assert n == len(hits) == ctx.learn_dataset_size() and n > 0
    X, T = ctx.learn_dataset()
    assert_bitexact(T, ctx.pt_continue(hits, 8)[:, :3].copy(), "targets")
    pts = np.ascontiguousarray(hits)
    Xf = np.empty((n, 224), np.uint16)
    ctx.check(ctx.lib.cr_learn_features(ctx.h, pts.ctypes.data, n, Xf.ctypes.data), "features")
    assert np.array_equal(X, Xf)
    # every k-th collision when capped
    ctx.learn_dataset_clear()
    m = ctx.learn_collect(fr, 8, max(1, n // 3))
    k = (n + n // 3 - 1) // (n // 3)
    assert m == (n + k - 1) // k
    X2, T2 = ctx.learn_dataset()
    assert np.array_equal(X2, X[::k]) and np.array_equal(T2, T[::k])
    ctx.learn_dataset_clear()


def test_shadow_volume_first_nee_is_exact(ctx, cfg, S, pkg):
    """With shadow-volume NEE the first collision's next-event term is a
    function of the recorded position and direction: the data set leaves it
    out of the targets and the learned render adds it back exactly.  With a
    bounce cap of 1 and a black sky a continuation is exactly that term, so
    (2 continuations, an exact mean) every target is 0, and a network that
    predicts 0 renders exactly the plain tracer's image (RGB)."""
    sc = cfg.c1(width=24, height=20, size=40, sv_axis=48)
    sc.fbm.seed = 3
    sc.ext = cfg.render_ext(albedo=0.9, bounce_cap=1, sky=(0.0, 0.0, 0.0), nee=S.CR_NEE_SHADOW_VOLUME,
                            tracking=S.CR_TRACK_GLOBAL_MAJORANT)
    sc.setup(ctx)
    ctx.learn_dataset_clear()
    fr = cfg.frames(2)
    n = ctx.learn_collect(fr, 2, 0)
    rec = ctx.pt_record(fr, sc.height, sc.width).reshape(-1, 8)
    hits = rec[rec[:, 7] > 0.5]
    assert n == len(hits) > 0
    cont = ctx.pt_continue(hits, 2)[:, :3]
    assert cont.max() > 0.0
    X, T = ctx.learn_dataset()
    assert not T.any(), "targets must leave out the first collision's next-event term"
    # a network that outputs 0: the learned render is the exact term alone
    lib = pkg.load()
    tot = C.c_uint32()
    lib.cr_mlp_param_layout(None, None, None, None, C.byref(tot))
    zeros = np.zeros(tot.value, np.float32)
    ctx.check(lib.cr_mlp_set_params(ctx.h, zeros.ctypes.data, tot.value), "zero network")
    fr = cfg.frames(3)
    ctx.clear_image()
    ctx.pt_render(fr)
    plain = ctx.download_image(sc.height, sc.width)
    ctx.clear_image()
    ctx.learn_render(fr)
    learned = ctx.download_image(sc.height, sc.width)
    assert plain[..., :3].max() > 0.0
    assert_bitexact(learned[..., :3].copy(), plain[..., :3].copy(), "learned render with a zero network")
    # the data set keeps one kind of target
    sc.ext.neeMode = S.CR_NEE_RATIO_TRACKING
    ctx.set_render_ext(sc.ext)
    assert ctx.lib.cr_learn_collect(ctx.h, (S.PushConstants * 2)(*fr[:2]), 2, 2, 0, None) != 0
    ctx.learn_dataset_clear()


def test_learned_render_is_the_same_in_any_number_of_calls(ctx, cfg, S):
    """n frames of a learned render in one call give the image that the same
    frames give over several calls: the composition blends frame by frame in
    frame order, and a record's features and network output do not depend on
    the other records of its call (renderers pick the frames per call for
    speed alone)."""
    sc = scene(cfg, S, seed=5)
    sc.setup(ctx)
    ctx.mlp_init(3)
    fr = cfg.frames(12)
    ctx.clear_image()
    ctx.learn_render(fr)
    one = ctx.download_image(sc.height, sc.width)
    ctx.clear_image()
    for a, b in ((0, 5), (5, 6), (6, 12)):
        ctx.learn_render(fr[a:b])
    parts = ctx.download_image(sc.height, sc.width)
    assert one[..., :3].max() > 0.0
    assert_bitexact(parts, one, "learned render of 12 frames in three calls vs one")


def test_features_match_oracle(ctx, cfg, S):
    spec = importlib.util.spec_from_file_location("oracle_mlp", os.path.join(ROOT, "oracle", "oracle_mlp.py"))
    M = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(M)
    sc = scene(cfg, S, size=36, seed=4, light=(0.3, -1.0, 0.2))
    cloud, svp = sc.setup(ctx)
    ctx.learn_build_mips()
    grid = ctx.download_grid()
    mips = M.mip_pyramid(grid)
    rng = np.random.default_rng(5)
    b0 = np.array(cloud.bounds[0][:3]), np.array(cloud.bounds[1][:3])
    x = rng.uniform(b0[0], b0[1], (64, 3)).astype(np.float32)
    om = rng.standard_normal((64, 3)).astype(np.float32)
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    pts = np.zeros((64, 8), np.float32)
    pts[:, :3] = x
    pts[:, 3:6] = om
    X = np.empty((64, 224), np.uint16)
    ctx.check(ctx.lib.cr_learn_features(ctx.h, pts.ctypes.data, 64, X.ctypes.data), "features")
    got = bf16_to_f32(X)
    light = np.array(svp.lightDirection[:3])
    h0 = (cloud.bounds[1][0] - cloud.bounds[0][0]) / grid.shape[2]
    sscale = cloud.densityScaling / cloud.baseScaling
    for i in range(64):
        ref = M.features(mips, cloud.bounds[0], cloud.bounds[1], sscale, h0, x[i], om[i], light)
        np.testing.assert_allclose(got[i], ref, rtol=1e-2, atol=2e-3, err_msg=f"point {i}")


def rel_mse(img, ref, eps=1e-2):
    a, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    return float(np.mean((a - r) ** 2 / (r * r + eps)))


def test_trained_network_beats_plain_tracer_at_equal_spp(ctx, cfg, S):
    """Training data from 6 seeded clouds, a held-out cloud for the
    comparison; reference = 1024-frame plain render."""
    ctx.learn_dataset_clear()
    lights = [(1.0, -1.0, 0.0), (0.4, -1.0, 0.3), (-0.6, -1.0, 0.2)]
    for k in range(6):
        sc = scene(cfg, S, size=40, w=48, h=48, seed=20 + k, light=lights[k % 3])
        sc.setup(ctx)
        ctx.learn_collect(cfg.frames(16, srand_seed=100 + k), 48, 6000)
    n = ctx.learn_dataset_size()
    assert n > 20000
    ctx.mlp_init(3)
    first = ctx.learn_train(20, 1024, 1e-3, seed=2)
    last = ctx.learn_train(1500, 1024, 1e-3, seed=2)
    assert last < 0.7 * first

    sc = scene(cfg, S, size=40, w=48, h=48, seed=7, light=lights[0])
    sc.setup(ctx)
    ctx.clear_image()
    ctx.pt_render(cfg.frames(1024, srand_seed=7))
    ref = ctx.download_image(sc.height, sc.width)
    ctx.clear_image()
    ctx.pt_render(cfg.frames(4))
    plain = ctx.download_image(sc.height, sc.width)
    ctx.clear_image()
    ctx.learn_render(cfg.frames(4))
    learned = ctx.download_image(sc.height, sc.width)
    assert np.isfinite(learned).all()
    e_plain, e_learned = rel_mse(plain, ref), rel_mse(learned, ref)
    print(f"relMSE at 4 spp: plain {e_plain:.4g}, learned {e_learned:.4g} ({n} training samples, loss {last:.4g})")
    assert e_learned < e_plain
    ctx.learn_dataset_clear()


def draw_rows(seed, step, B, n):
    """The data-set row each batch row of a training step takes: a PCG-style
    hash of (seed, step, row) modulo the data-set size, in uint32 arithmetic."""
    r = np.arange(B, dtype=np.uint64)
    M = np.uint64(0xFFFFFFFF)
    st = ((np.uint64(seed * 0x9E3779B9 & 0xFFFFFFFF)) ^ np.uint64(step * 0x85EBCA6B & 0xFFFFFFFF)
          ^ ((r * np.uint64(0xC2B2AE35)) & M))
    st = (st * np.uint64(747796405) + np.uint64(2891336453)) & M
    w = (((st >> ((st >> np.uint64(28)) + np.uint64(4))) ^ st) * np.uint64(277803737)) & M
    w = (w >> np.uint64(22)) ^ w
    return (w % np.uint64(n)).astype(np.int64)


def test_training_draws_the_stated_rows(pkg, cfg, S):
    """cr_learn_train's step s (counted from the context's first) trains on
    data-set rows draw_rows(seed, s, B, n), exactly as cr_mlp_train_step does
    on those rows handed over from the host: the parameters after one drawn
    step equal, bit for bit, those after one step on the host-gathered batch."""
    c = pkg.Context(0)
    try:
        lib = c.lib
        sc = scene(cfg, S, w=24, h=20, seed=3)
        sc.setup(c)
        n = c.learn_collect(cfg.frames(2), 8, 0)
        assert n > 50
        X, T = c.learn_dataset()
        c.check(lib.cr_mlp_init(c.h, 5), "init")
        total = C.c_uint32()
        lib.cr_mlp_param_layout(None, None, None, None, C.byref(total))
        P0 = np.zeros(total.value, np.float32)
        c.check(lib.cr_mlp_get_params(c.h, P0.ctypes.data, total.value), "params")
        B, seed = 300, 17
        c.check(lib.cr_learn_train(c.h, 1, B, 1e-3, seed, None), "drawn step")
        P1 = np.zeros_like(P0)
        c.check(lib.cr_mlp_get_params(c.h, P1.ctypes.data, total.value), "params")
        idx = draw_rows(seed, 0, B, n)
        Xb = np.ascontiguousarray(X[idx])
        Tb = np.ascontiguousarray(T[idx])
        c.check(lib.cr_mlp_set_params(c.h, P0.ctypes.data, total.value), "reset")
        loss = C.c_float()
        c.check(lib.cr_mlp_train_step(c.h, Xb.ctypes.data, Tb.ctypes.data, B, 1e-3, C.byref(loss), None), "host step")
        P2 = np.zeros_like(P0)
        c.check(lib.cr_mlp_get_params(c.h, P2.ctypes.data, total.value), "params")
        assert not np.array_equal(P1, P0)
        assert np.array_equal(P1.view(np.uint32), P2.view(np.uint32))
    finally:
        c.close()
