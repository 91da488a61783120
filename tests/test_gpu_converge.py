"""Convergence of the production trackers to the reference technique.

The reference technique is shaders/PathTracer.comp's estimator: delta
tracking against the global majorant (PathTracer.comp:318-361) with
shadow-volume next-event estimation (PathTracer.comp:121-127, 405-411).  The
global-majorant kernel reproduces it bit for bit (test_gpu_pt.py), so a
16384-spp render of it is the converged reference image.  Ratio-tracked
shadow rays estimate the exact transmittance toward the sun, where the shadow
volume holds a trilinear interpolation of a marched transmittance, so their
images have a slightly different expectation; they are judged against a
16384-spp global-majorant render with the same ratio-tracked shadow rays.

The brick-majorant DDA tracker samples the same free-flight distribution with
tighter majorants, so its image has the expectation of the global-majorant
render with the same next-event estimator.  At 4096 spp it must match within
relMSE <= 1e-3 (mean over pixels and RGB of (I - R)^2 / (R^2 + 1e-2)) and per
pixel within 3 standard errors.  Each render is the mean of 32 independent
batches (distinct random frame seeds, each batch's running mean starting at
frameCount 0); a render's per-pixel standard error is the batches' sample
standard deviation over sqrt(32), and the z score of a pixel-channel is the
difference of the two renders over the root sum of their squared standard
errors.

The z score needs a standard error that the batches estimate well.  Where
collisions are rare (the sky around the cloud) most batches hold no event at
all and one event in one batch of one render moves that pixel by ~2e-5 with a
standard error of ~4e-7: |z| ~ 14 without any bias.  So a pixel-channel enters
the per-pixel gate only if in at least one of the two renders at least half
of the batches differ from the batch median; the others, rare-event pixels,
are held to a relative difference of 5e-3.  Under the null hypothesis the
gate's |z| > 3 fraction is about 0.3-0.6% and its mean z^2 about 1.05 (t
distributions with ~60 degrees of freedom, skewed batches); the test allows
1% and 1.15, and a null pair (the reference technique against itself with
other seeds) calibrates both.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BATCHES = 32
REF_SPP = 16384
TEST_SPP = 4096
OUT3_MAX = 0.01
MEAN_Z2_MAX = 1.15
RARE_REL_MAX = 5e-3
# distinct frame seeds (PathTracer.comp:368 seeds a pixel's stream with seed * id), drawn at random rather than
# consecutive so that frames of one pixel do not start from states that differ in a few bits
SEEDS = np.random.default_rng(2026).choice(1 << 20, REF_SPP + TEST_SPP, replace=False) + 1
# The c2-shaped scene (albedo 0.999, <= 256 bounces) has long, heavy-tailed
# paths: a batch of 128 paths is too skewed for the batches' sample deviation
# to estimate its standard error (measured: mean z^2 1.26 for 4096 spp against
# 16384, where null pairs of 16384 spp give 1.00 and 0.96), so its renders use
# 16384 spp (batches of 512 paths) on both sides.
C2_SEEDS = np.random.default_rng(77).choice(1 << 22, 2 * REF_SPP, replace=False) + 1


def batch_frames(S, seeds):
    """One batch: the given seeds, frameCount 0, 1, ... so the image holds the
    exact mean of the batch's frames."""
    return [S.PushConstants(0.0, int(s), i, 0.0, 0) for i, s in enumerate(seeds)]


def render_batches(ctx, S, sc, seeds):
    """The 32 batch images (float64, RGB)."""
    per = len(seeds) // BATCHES
    imgs = np.empty((BATCHES, sc.height, sc.width, 3), np.float64)
    for k in range(BATCHES):
        ctx.clear_image()
        ctx.pt_render(batch_frames(S, seeds[k * per:(k + 1) * per]))
        imgs[k] = ctx.download_image(sc.height, sc.width)[..., :3]
    return imgs


def rel_mse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


def gate(test_b, ref_b):
    """relMSE, the z scores of the gated pixel-channels, and the largest
    relative difference among the rare-event ones."""
    a, b = test_b.mean(0), ref_b.mean(0)
    sa = test_b.std(0, ddof=1) / np.sqrt(BATCHES)
    sb = ref_b.std(0, ddof=1) / np.sqrt(BATCHES)
    den = np.sqrt(sa * sa + sb * sb)

    def off_median(x):
        return (x != np.median(x, axis=0)).sum(0)

    gated = (np.maximum(off_median(test_b), off_median(ref_b)) >= BATCHES // 2) & (den > 0)
    z = (a - b)[gated] / den[gated]
    rare = ~gated
    rel = float((np.abs(a - b)[rare] / np.maximum(np.abs(b[rare]), 1e-6)).max()) if rare.any() else 0.0
    return rel_mse(a, b), z, rel, int(rare.sum())


def report(what, e, z, rel, n_rare, seconds):
    out3 = float(np.mean(np.abs(z) > 3.0))
    z2 = float(np.mean(z * z))
    print(f"{what} ({seconds:.2f} s): relMSE {e:.3e}; {z.size} gated pixel-channels: |z| > 3 for "
          f"{100 * out3:.3f}%, mean z^2 {z2:.3f}, max |z| {float(np.max(np.abs(z))):.2f}; {n_rare} rare-event "
          f"pixel-channels, largest relative difference {rel:.2e}", flush=True)
    return out3, z2


def c1_scene(cfg, S, tracking, nee, albedo=0.9):
    sc = cfg.c1(width=128, height=128)
    sc.ext = cfg.render_ext(albedo=albedo, bounce_cap=16, sky=cfg.SKY, tracking=tracking, nee=nee)
    return sc


def c2_scene(cfg, S, tracking):
    """The production workload's settings (FP16 anvil in 4^3 bricks, albedo
    0.999, <= 256 bounces, ratio-tracked shadow rays) at 96x48x96 voxels and
    96x54 pixels."""
    sc = cfg.c2(width=96, height=54, size=(96, 48, 96))
    sc.ext.trackingMode = tracking
    return sc


@pytest.fixture(scope="module")
def references(ctx, cfg, S):
    """16384-spp global-majorant renders: c1 with shadow-volume NEE (the
    reference technique) and with ratio tracking, and the c2-shaped scene with
    ratio tracking."""
    out = {}
    for key, sc in (("c1/sv", c1_scene(cfg, S, S.CR_TRACK_GLOBAL_MAJORANT, S.CR_NEE_SHADOW_VOLUME)),
                    ("c1/rt", c1_scene(cfg, S, S.CR_TRACK_GLOBAL_MAJORANT, S.CR_NEE_RATIO_TRACKING)),
                    ("c2/rt", c2_scene(cfg, S, S.CR_TRACK_GLOBAL_MAJORANT))):
        sc.setup(ctx)
        t0 = time.time()
        out[key] = render_batches(ctx, S, sc, (C2_SEEDS if key.startswith("c2") else SEEDS)[:REF_SPP])
        assert np.isfinite(out[key]).all()
        print(f"reference {key}, {REF_SPP} spp in {BATCHES} batches: {time.time() - t0:.2f} s", flush=True)
    return out


def check(ctx, S, sc, ref_b, what, seeds=None):
    sc.setup(ctx)
    t0 = time.time()
    test_b = render_batches(ctx, S, sc, SEEDS[REF_SPP:] if seeds is None else seeds)   # disjoint from the reference's
    assert np.isfinite(test_b).all()
    e, z, rel, n_rare = gate(test_b, ref_b)
    out3, z2 = report(what, e, z, rel, n_rare, time.time() - t0)
    return e, out3, z2, rel, z.size


@pytest.mark.parametrize("variant", ["null", "dda_sv", "dda_ratio", "c2_null", "c2_dda_ratio"])
def test_converges_to_reference(ctx, cfg, S, references, variant):
    """null: the reference technique against itself (other seeds), which
    calibrates the gate; dda_sv: brick DDA + shadow-volume NEE against the
    reference technique; dda_ratio: brick DDA + ratio tracking against global
    majorant + ratio tracking (4096 spp against 16384); c2_null, c2_dda_ratio:
    on the c2-shaped scene, the global-majorant render and the benchmarked
    combination against the global-majorant reference (16384 spp each)."""
    G, D = S.CR_TRACK_GLOBAL_MAJORANT, S.CR_TRACK_BRICK_DDA
    SV, RT = S.CR_NEE_SHADOW_VOLUME, S.CR_NEE_RATIO_TRACKING
    sc, key = {"null": (c1_scene(cfg, S, G, SV), "c1/sv"),
               "dda_sv": (c1_scene(cfg, S, D, SV), "c1/sv"),
               "dda_ratio": (c1_scene(cfg, S, D, RT), "c1/rt"),
               "c2_null": (c2_scene(cfg, S, G), "c2/rt"),
               "c2_dda_ratio": (c2_scene(cfg, S, D), "c2/rt")}[variant]
    seeds = C2_SEEDS[REF_SPP:] if variant.startswith("c2") else None
    e, out3, z2, rel, n = check(ctx, S, sc, references[key], variant, seeds)
    assert n > 1000
    assert e <= 1e-3
    assert out3 <= OUT3_MAX
    assert 0.9 <= z2 <= MEAN_Z2_MAX
    assert rel <= RARE_REL_MAX


def test_per_pixel_gate_rejects_a_biased_render(ctx, cfg, S, references):
    """The same gate, applied to the reference technique with the albedo
    lowered from 0.9 to 0.89, must fail: the per-pixel test has the power to
    see a bias of that size at these sample counts, where the mean relMSE
    bound does not (that render's relMSE against the reference is ~8.5e-4)."""
    sc = c1_scene(cfg, S, S.CR_TRACK_GLOBAL_MAJORANT, S.CR_NEE_SHADOW_VOLUME, albedo=0.89)
    e, out3, z2, rel, n = check(ctx, S, sc, references["c1/sv"], "reference technique at albedo 0.89")
    assert e <= 1e-3
    assert out3 > OUT3_MAX and z2 > MEAN_Z2_MAX


def test_shadow_volume_and_ratio_tracking_differ_visibly(ctx, cfg, S, references):
    """Why ratio tracking has its own reference: the shadow volume's
    discretised transmittance gives the image a slightly different
    expectation, which the gate sees (mean z^2 well above its null value)
    though the mean bound does not."""
    sc = c1_scene(cfg, S, S.CR_TRACK_BRICK_DDA, S.CR_NEE_RATIO_TRACKING)
    e, out3, z2, rel, n = check(ctx, S, sc, references["c1/sv"], "brick DDA + ratio tracking vs the SV reference")
    assert e <= 1e-3
    assert z2 > MEAN_Z2_MAX
