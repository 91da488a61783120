"""Convergence of the production trackers to the reference technique.

The reference technique is shaders/PathTracer.comp's estimator: delta
tracking against the global majorant (PathTracer.comp:318-361) with
shadow-volume next-event estimation (PathTracer.comp:121-127, 405-411).  The
global-majorant kernel reproduces it bit for bit (test_gpu_pt.py), so a
16384-spp render of it is the converged reference image.

The brick-majorant DDA tracker samples the same free-flight distribution with
tighter majorants, so with the same shadow-volume NEE its image has the same
expectation.  At 4096 spp it must match the reference render within relMSE
<= 1e-3 (mean over pixels and RGB of (I - R)^2 / (R^2 + 1e-2)) and per pixel
within 3 standard errors.  Each render is the mean of 32 independent batches
(distinct random frame seeds, each batch's running mean starting at frameCount 0);
the per-pixel standard error of a render is the batches' sample standard
deviation over sqrt(32), and the z score of a pixel-channel is the difference
of the two renders over the root sum of their squared standard errors.  Under
the null hypothesis |z| > 3 holds for about 0.3-0.6% of the pixel-channels (a
t distribution with ~60 degrees of freedom); the test allows 1%.

Ratio-tracked shadow rays estimate the exact transmittance toward the sun,
where the shadow volume holds a trilinear interpolation of a marched
transmittance, so that image has a (small) different expectation; it is held
to the relMSE bound only, and its share of 3-sigma outliers is printed.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BATCHES = 32
REF_SPP = 16384
TEST_SPP = 4096
# distinct frame seeds (PathTracer.comp:368 seeds a pixel's stream with seed * id), drawn at random rather than
# consecutive so that frames of one pixel do not start from states that differ in a few bits
SEEDS = np.random.default_rng(2026).choice(1 << 20, REF_SPP + TEST_SPP, replace=False) + 1


def batch_frames(S, seeds):
    """One batch: the given seeds, frameCount 0, 1, ... so the image holds the
    exact mean of the batch's frames."""
    return [S.PushConstants(0.0, int(s), i, 0.0, 0) for i, s in enumerate(seeds)]


def render_batches(ctx, S, sc, seeds):
    per = len(seeds) // BATCHES
    imgs = np.empty((BATCHES, sc.height, sc.width, 3), np.float64)
    for k in range(BATCHES):
        ctx.clear_image()
        ctx.pt_render(batch_frames(S, seeds[k * per:(k + 1) * per]))
        imgs[k] = ctx.download_image(sc.height, sc.width)[..., :3]
    return imgs.mean(0), imgs.std(0, ddof=1) / np.sqrt(BATCHES)


def rel_mse(img, ref, eps=1e-2):
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


def z_scores(a, se_a, b, se_b):
    """z of every pixel-channel with spread in either render, and the mask of those."""
    den = np.sqrt(se_a * se_a + se_b * se_b)
    live = den > 0
    return (a - b)[live] / den[live], live


@pytest.fixture(scope="module")
def reference(ctx, cfg, S):
    sc = cfg.c1(width=128, height=128)
    sc.setup(ctx)
    t0 = time.time()
    ref, se = render_batches(ctx, S, sc, SEEDS[:REF_SPP])
    print(f"reference technique, {REF_SPP} spp in {BATCHES} batches: {time.time() - t0:.2f} s", flush=True)
    assert np.isfinite(ref).all()
    return sc, ref, se


def compare(img, se, ref, se_ref):
    """relMSE, the z scores of the pixel-channels with spread, and the mask of those."""
    z, live = z_scores(img, se, ref, se_ref)
    # a pixel-channel with no spread (every path escapes without a collision) has one value per render, equal
    # up to the rounding of the running mean, which depends on the batch length
    flat = ~live
    assert np.all(np.abs(img - ref)[flat] <= 1e-4 * np.abs(ref[flat])), "constant pixels differ"
    return rel_mse(img, ref), z, live


def report(what, e, z, live, seconds):
    out3 = float(np.mean(np.abs(z) > 3.0))
    print(f"{what}, {TEST_SPP} spp ({seconds:.2f} s): relMSE {e:.3e}; of {z.size} pixel-channels with spread "
          f"({int((~live).sum())} without), |z| > 3 for {100 * out3:.3f}%, mean z^2 {float(np.mean(z * z)):.3f}, "
          f"max |z| {float(np.max(np.abs(z))):.2f}", flush=True)
    return out3


@pytest.mark.parametrize("nee", ["shadow_volume", "ratio_tracking"])
def test_brick_dda_converges_to_reference_technique(ctx, cfg, S, reference, nee):
    _, ref, se_ref = reference
    sc = cfg.c1(width=128, height=128)
    sc.ext = cfg.render_ext(albedo=0.9, bounce_cap=16, sky=cfg.SKY, tracking=S.CR_TRACK_BRICK_DDA,
                            nee=S.CR_NEE_SHADOW_VOLUME if nee == "shadow_volume" else S.CR_NEE_RATIO_TRACKING)
    sc.setup(ctx)
    t0 = time.time()
    img, se = render_batches(ctx, S, sc, SEEDS[REF_SPP:])  # seeds disjoint from the reference's
    assert np.isfinite(img).all()
    e, z, live = compare(img, se, ref, se_ref)
    out3 = report(f"brick DDA + {nee}", e, z, live, time.time() - t0)
    assert e <= 1e-3
    if nee == "shadow_volume":
        assert out3 <= 0.01


def test_per_pixel_gate_rejects_a_biased_render(ctx, cfg, S, reference):
    """The same gate, applied to the reference technique with the albedo
    lowered from 0.9 to 0.89, must fail: the per-pixel test has the power to
    see a bias of that size at these sample counts, where the mean relMSE
    bound does not (that render's relMSE against the reference is 8.5e-4)."""
    _, ref, se_ref = reference
    sc = cfg.c1(width=128, height=128)
    sc.ext = cfg.render_ext(albedo=0.89, bounce_cap=16, sky=cfg.SKY)
    sc.setup(ctx)
    t0 = time.time()
    img, se = render_batches(ctx, S, sc, SEEDS[REF_SPP:])
    e, z, live = compare(img, se, ref, se_ref)
    out3 = report("reference technique at albedo 0.89", e, z, live, time.time() - t0)
    assert out3 > 0.01
