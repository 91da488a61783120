"""Pins of the oracle functions the learned path and the photon-beam gather
rest on (CPU only).

Each check is fixed by mathematics independent of the oracle's own formula:
central differences for gradients, Kingma & Ba's closed forms for Adam,
affine density ramps (which trilinear filtering and 2x2x2 mip averaging
reproduce exactly) for the stencil geometry, the textbook Henyey-Greenstein
and biweight forms for one beam, the single-scattering integral estimated by
an independent estimator (the path tracer) for the beam gather, and the
projected area of a box for sun emission.  A plausible slip (a dropped
factor, a transposed weight, a wrong bias correction, sin^2 for sin, a
missing 1/r) fails at least one of them.
"""
import ctypes as C
import importlib.util
import math
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f3 = C.c_float * 3


def load_mlp(float64=False):
    """oracle/oracle_mlp.py; with float64 a private copy whose float32 casts
    are float64 (so central differences are not drowned by rounding)."""
    spec = importlib.util.spec_from_file_location("oracle_mlp_pin", os.path.join(ROOT, "oracle", "oracle_mlp.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    if float64:
        class NP64:
            def __getattr__(self, k):
                return np.float64 if k == "float32" else getattr(np, k)
        m.np = NP64()
    return m


# --------------------------------------------------------------- network
def test_mlp_gradients_central_differences():
    """loss_and_grads (oracle/oracle_mlp.py) on a shrunk 7 -> 6 -> 5 -> 3
    network against float64 central differences of its own loss, every
    parameter."""
    M = load_mlp(float64=True)
    M.LAYERS = [(7, 6), (6, 5), (5, 3)]
    rng = np.random.default_rng(7)
    params = []
    for i, o in M.LAYERS:
        params += [rng.standard_normal((i, o)) * 0.7, rng.standard_normal(o) * 0.3]
    x = rng.standard_normal((9, 7))
    target = rng.uniform(0.0, 3.0, (9, 3))
    loss, grads = M.loss_and_grads(params, x, target)
    # the loss is the mean squared error against log(1 + target)
    y = x
    for li in range(3):
        y = y @ params[2 * li] + params[2 * li + 1]
        if li < 2:
            y = np.maximum(y, 0.0)
    assert abs(loss - float(np.mean((y - np.log1p(target)) ** 2))) < 1e-12
    eps = 1e-6
    for pi, p in enumerate(params):
        fd = np.zeros_like(p)
        for idx in np.ndindex(p.shape):
            old = p[idx]
            p[idx] = old + eps
            lp, _ = M.loss_and_grads(params, x, target)
            p[idx] = old - eps
            lm, _ = M.loss_and_grads(params, x, target)
            p[idx] = old
            fd[idx] = (lp - lm) / (2 * eps)
        assert grads[pi].shape == p.shape
        np.testing.assert_allclose(grads[pi], fd, rtol=1e-5, atol=1e-8, err_msg=f"parameter {pi}")
        assert np.abs(fd).max() > 1e-4   # every layer carries gradient here


def test_adam_closed_forms():
    """adam_step (Kingma & Ba 2015, bias-corrected): with m = v = 0 and a
    constant gradient g every step moves each parameter by exactly
    -lr * g / (|g| + eps); and two hand-evaluated steps with g = +1, -1."""
    M = load_mlp()
    lr, eps = 1e-3, 1e-8
    g = np.array([0.5, -2.0, 1e-3, -3e-6], np.float32)
    p = [np.zeros(4, np.float32)]
    m = [np.zeros(4, np.float32)]
    v = [np.zeros(4, np.float32)]
    for step in range(1, 6):
        M.adam_step(p, [g], m, v, step, lr=lr, eps=eps)
        np.testing.assert_allclose(p[0], -step * lr * g / (np.abs(g) + eps), rtol=2e-5)
    # hand-evaluated: b1 = 0.9, b2 = 0.999
    #   step 1, g = +1: m = 0.1, v = 0.001, mhat = 1, vhat = 1 -> p = -lr / (1 + eps)
    #   step 2, g = -1: m = 0.09 - 0.1 = -0.01, v = 0.000999 + 0.001 = 0.001999,
    #           mhat = -0.01 / 0.19, vhat = 0.001999 / 0.001999 = 1 -> p += lr * 0.01 / 0.19 / (1 + eps)
    p = [np.zeros(1, np.float32)]
    m = [np.zeros(1, np.float32)]
    v = [np.zeros(1, np.float32)]
    M.adam_step(p, [np.ones(1, np.float32)], m, v, 1, lr=lr, eps=eps)
    assert abs(m[0][0] - 0.1) < 1e-7 and abs(v[0][0] - 0.001) < 1e-9
    np.testing.assert_allclose(p[0][0], -lr / (1 + eps), rtol=1e-6)
    M.adam_step(p, [-np.ones(1, np.float32)], m, v, 2, lr=lr, eps=eps)
    assert abs(m[0][0] + 0.01) < 1e-7 and abs(v[0][0] - 0.001999) < 1e-9
    np.testing.assert_allclose(p[0][0], -lr / (1 + eps) + lr * (0.01 / 0.19) / (1 + eps), rtol=1e-5)


def test_mip_pyramid_is_block_mean():
    """Level k+1 is the mean of the 2x2x2 blocks of level k; an edge block of
    an odd-sized level averages the voxels that exist (brute force)."""
    M = load_mlp()
    rng = np.random.default_rng(3)
    g = rng.random((8, 12, 16)).astype(np.float32)
    mips = M.mip_pyramid(g, 4)
    np.testing.assert_allclose(mips[1], g.reshape(4, 2, 6, 2, 8, 2).mean(axis=(1, 3, 5)), rtol=1e-6)
    np.testing.assert_allclose(mips[2], g.reshape(2, 4, 3, 4, 4, 4).mean(axis=(1, 3, 5)), rtol=1e-6)
    h = rng.random((5, 3, 7)).astype(np.float32)
    m1 = M.mip_pyramid(h, 2)[1]
    assert m1.shape == (3, 2, 4)
    for z, y, x in np.ndindex(m1.shape):
        blk = h[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2]
        assert abs(m1[z, y, x] - blk.mean()) < 1e-6


def _feature_setup(M):
    b0, b1 = np.array([-0.5, 0.0, -0.25]), np.array([0.5, 1.0, 0.75])
    L = np.array([0.3, -1.0, 0.45])
    L /= np.linalg.norm(L)
    return b0, b1, L


def test_features_constant_density():
    """Constant density rho: a tap is rho * sigma_scale * h_k inside the box
    and 0 outside; at the centre levels 0..4 (h_k <= 1/4) lie wholly inside
    and levels 6, 7 (h_k >= 1) keep only the centre tap; then cos(omega, -L),
    its square and six zeros."""
    M = load_mlp()
    b0, b1, L = _feature_setup(M)
    rho, s, h0 = 0.5, 7.0, 1.0 / 64
    mips = M.mip_pyramid(np.full((32, 32, 32), rho, np.float32))
    x = (b0 + b1) / 2
    omega = np.array([0.0, 0.6, 0.8])
    f = M.features(mips, b0, b1, s, h0, x, omega, L)
    assert f.shape == (224,)
    for k in range(8):
        taps = f[27 * k:27 * (k + 1)]
        full = np.float32(rho * s * h0 * 2 ** k)
        if k <= 4:
            np.testing.assert_allclose(taps, full, rtol=1e-6)
        elif k >= 6:
            np.testing.assert_allclose(taps[13], full, rtol=1e-6)
            assert (np.delete(taps, 13) == 0).all()
    c = float(np.dot(omega, -L))
    assert abs(f[216] - c) < 1e-6 and abs(f[217] - c * c) < 1e-6
    assert (f[218:] == 0).all()
    # a point outside the box sees only zeros at the finest level
    f = M.features(mips, b0, b1, s, h0, b1 + 0.1, omega, L)
    assert (f[:27] == 0).all()


def test_features_stencil_geometry_from_affine_ramps():
    """Densities affine in one coordinate are reproduced exactly by trilinear
    filtering and by 2x2x2 averaging, so a tap of the ramp in axis a returns
    that coordinate of the tap's position.  Three ramps recover every tap
    position: the stencil at level k is x + h_k (i e_x + j e_y + l e_z) with
    e_z = -L (toward the sun), e_x, e_y, e_z orthonormal, h_k = h_0 2^k, and
    i fastest, then j, then l."""
    M = load_mlp()
    b0, b1, L = _feature_setup(M)
    N = 64
    s, h0 = 3.0, 1.0 / 64
    centres = (np.arange(N) + 0.5) / N
    pos = np.zeros((4, 27, 3))
    x = b0 + np.array([0.47, 0.52, 0.5]) * (b1 - b0)
    omega = np.array([1.0, 0.0, 0.0])
    for a in range(3):
        ramp = 0.25 + centres            # density = 0.25 + u_a, u_a the normalised coordinate
        shape = [1, 1, 1]
        shape[2 - a] = N
        g = np.broadcast_to(ramp.reshape(shape), (N, N, N)).astype(np.float64)
        mips = M.mip_pyramid(g)
        f = M.features(mips, b0, b1, s, h0, x, omega, L)
        for k in range(4):
            u = f[27 * k:27 * (k + 1)].astype(np.float64) / (s * h0 * 2 ** k) - 0.25
            pos[k, :, a] = b0[a] + u * (b1[a] - b0[a])
    ez = -L
    for k in range(4):
        h = h0 * 2 ** k
        p = pos[k].reshape(3, 3, 3, 3)   # [l, j, i, axis]
        np.testing.assert_allclose(p[1, 1, 1], x, atol=2e-6)
        ex = (p[1, 1, 2] - p[1, 1, 0]) / (2 * h)
        ey = (p[1, 2, 1] - p[1, 0, 1]) / (2 * h)
        ez_got = (p[2, 1, 1] - p[0, 1, 1]) / (2 * h)
        np.testing.assert_allclose(ez_got, ez, atol=1e-4)
        basis = np.stack([ex, ey, ez_got])
        np.testing.assert_allclose(basis @ basis.T, np.eye(3), atol=2e-4)
        for l, j, i in np.ndindex(3, 3, 3):
            np.testing.assert_allclose(p[l, j, i], x + h * ((i - 1) * ex + (j - 1) * ey + (l - 1) * ez), atol=1e-5)


# ------------------------------------------------------------ beam gather
def unit_cloud(S, pkg, grid, density_scaling):
    c = S.CloudProperties()
    pkg.load().cr_cloud_default(C.byref(c))
    nz, ny, nx = grid.shape
    c.baseScaling, c.densityScaling = 1.0, density_scaling
    c.maxExtinction = max(float(grid.max()), 0.01)
    c.voxelCount[:] = (nx, ny, nz, 0)
    c.bounds[0][:] = (-0.5, -0.5, 0.0, 0.0)
    c.bounds[1][:] = (0.5, 0.5, 1.0, 0.0)
    return c


def hg(g, cos_t):
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos_t) ** 1.5)


@pytest.mark.parametrize("angle_deg,g", [(90.0, 0.0), (90.0, 0.6), (55.0, 0.6), (125.0, -0.3)])
def test_gather_single_beam_closed_form(O, S, pkg, cfg, angle_deg, g):
    """One beam crossing the centre pixel's ray at distance d (Jarosz et al.
    2011, beam x beam with a 1D biweight kernel):
    L = K(d/r)/r * albedo * sigma * Phi * T_cam * T_beam(s) * HG(cos) / sin,
    K(q) = 15/16 (1 - q^2)^2, evaluated by hand in float64, with T_beam the
    fraction of the 16 stored free paths that reach the crossing's distance
    from the start of the parent beam."""
    rho, sigma_scale, albedo, power, r, d = 0.7, 3.0, 0.8, 2.5, 0.04, 0.013
    grid = np.full((8, 8, 8), rho, np.float32)
    c = unit_cloud(S, pkg, grid, sigma_scale)
    cam = cfg.camera(16, 16, position=(0.0, 0.0, -0.8))
    assert tuple(cam.forward[:3]) == (0.0, 0.0, 1.0)
    svp = cfg.shadow_volume(c, tuple(cam.forward[:3]), 4)
    params = cfg.parameters(g)
    s = O.scene(grid, c, cam, params, svp, cfg.render_ext(), sv=np.full((4, 4, 4), 0.5, np.float32))
    # the beam lies in the plane y = d, at angle theta to +z, and crosses z = 0.4 at x = 0
    th = math.radians(angle_deg)
    bd = np.array([math.sin(th), 0.0, math.cos(th)])
    cross = np.array([0.0, d, 0.4])
    half, off = 0.05, 0.3
    b = (S.PhotonBeam * 1)()
    b[0].startPos[:] = tuple(cross - half * bd)
    b[0].endPos[:] = tuple(cross + half * bd)
    b[0].radius = r
    b[0].dataIdx = 0
    b[0]._padding_beam[0] = int(np.float32(off).view(np.uint32))
    dat = (S.PhotonBeamData * 1)()
    dat[0].power[:] = (power, power, power, power)
    soff = off + half                                   # distance of the crossing along the parent beam
    dists = [soff - 0.01] * 5 + [soff + 0.01] * 11       # 11 of 16 free paths reach it
    dat[0].trDistances[:] = dists
    nodes = (S.TreeNode * 1)()
    bs = O.beam_scene(s, b, dat, nodes, 1, 0, 0, albedo)
    got = O.gather_pixel(bs, 8, 8, 38, brute=True)
    q = d / r
    kern = 15.0 / 16.0 * (1 - q * q) ** 2 / r
    cos_t = float(np.dot(bd, [0.0, 0.0, -1.0]))
    sin_t = math.sqrt(1 - cos_t * cos_t)
    expect = kern * albedo * rho * sigma_scale * power * 0.5 * (11 / 16) * hg(g, cos_t) / sin_t
    np.testing.assert_allclose(got, expect, rtol=2e-5)
    # beyond the radius, or behind the camera, nothing
    b[0].startPos[1] = b[0].endPos[1] = 1.01 * r
    assert (O.gather_pixel(bs, 8, 8, 38, brute=True) == 0).all()


def _sorted_copy(S, beams, perm, n):
    raw = np.frombuffer(beams, np.uint8, count=48 * n).reshape(n, 48)[perm]
    out = (S.PhotonBeam * n)()
    C.memmove(out, raw.tobytes(), 48 * n)
    return out


def test_gather_single_scattering_matches_path_tracer(O, S, pkg, cfg):
    """Both estimate single scattering of sunlight in a homogeneous box,
    integral of T_cam(t) albedo sigma HG T_sun(t) I dt along each camera ray:
    the path tracer with a bounce cap of 1, a black sky and ratio-tracked
    shadow rays; the beam gather over first-flight sun beams
    (maxInteractions = 1), with T_cam from a 256^3 shadow volume along the
    camera axis.  Over the central 24x24 pixels of a 480x480 image (rays
    within 3 degrees of that axis) the means agree within 3%: the shadow
    volume's Riemann sum is good to about 1% here and the beam estimate's
    noise is about 0.5%.  At this angle sin^2 for sin would be 13% off, a
    missing albedo 11%, a missing density scale 50%."""
    sigma, albedo, g, intensity = 1.5, 0.9, 0.3, 5.0
    grid = np.ones((16, 16, 16), np.float32)
    c = unit_cloud(S, pkg, grid, sigma)
    L = np.array([0.5, -1.0, -0.6])
    L = tuple(L / np.linalg.norm(L))
    W, B = 480, 24
    x0 = y0 = (W - B) // 2
    cam = cfg.camera(W, W, position=(0.0, 0.0, -5.0))
    params = cfg.parameters(g, light_intensity=intensity)
    ext = cfg.render_ext(albedo=albedo, bounce_cap=1, sky=(0.0, 0.0, 0.0, 0.0), nee=S.CR_NEE_RATIO_TRACKING)
    s_pt = O.scene(grid, c, cam, params, cfg.shadow_volume(c, L, 4), ext, sv=np.ones((4, 4, 4), np.float32))
    pt, _ = O.pt_render_region(s_pt, cfg.frames(256), x0, y0, B, B)
    p = cfg.beam_params(1 << 16, seed=77, radius=0.03, albedo=albedo, light_dir=L, intensity=intensity,
                        max_interactions=1)
    beams, data, nb, nd = O.trace_beams(s_pt, p, 1 << 23)
    assert nd == 1 << 16 and nb < 1 << 23
    keys, perm, nodes = O.lbvh_build(beams, c.bounds[0], c.bounds[1])
    sb = _sorted_copy(S, beams, perm, nb)
    svp_cam = cfg.shadow_volume(c, tuple(cam.forward[:3]), 256)
    s_g = O.scene(grid, c, cam, params, svp_cam, cfg.render_ext())
    s_g = O.with_sv(s_g, O.sv_build(s_g))
    bs = O.beam_scene(s_g, sb, data, nodes, nb, 0, 0, albedo)
    gat = np.array([[O.gather_pixel(bs, x, y, 38)[0] for x in range(x0, x0 + B)] for y in range(y0, y0 + B)])
    a = float(gat.mean())
    b = float(pt[..., 0].mean())
    assert (pt[..., 0] > 0).all() and b > 0.05
    assert abs(a / b - 1.0) < 0.03, (a, b)


def test_sun_emission_uniform_over_projected_area(O, S, pkg, cfg):
    """generateSunRay (PPB_PT.comp:396-430, entry face): photons start on the
    faces the light enters through, in proportion to each face's projected
    area |L_a| * (face area), uniformly within a face, and each carries
    I * A_proj / N so the emitted power is I per unit projected area."""
    from scipy import stats
    grid = np.ones((8, 8, 8), np.float32)
    c = unit_cloud(S, pkg, grid, 1.0)
    L = np.array([0.4, -1.0, 0.7])
    L = L / np.linalg.norm(L)
    n = 30000
    s = O.scene(grid, c, cfg.camera(8, 8), cfg.parameters(0.85), cfg.shadow_volume(c, tuple(L), 4),
                cfg.render_ext(), sv=np.ones((4, 4, 4), np.float32))
    p = cfg.beam_params(n, seed=91, radius=0.5, albedo=0.9, light_dir=tuple(L), intensity=3.0, max_interactions=1)
    beams, data, nb, nd = O.trace_beams(s, p, 1 << 20)
    assert nd == n
    b0 = np.array(c.bounds[0][:3], np.float64)
    b1 = np.array(c.bounds[1][:3], np.float64)
    ext = b1 - b0
    face_area = np.array([ext[1] * ext[2], ext[0] * ext[2], ext[0] * ext[1]])
    a_proj = float((np.abs(L) * face_area).sum())
    power = np.array([data[i].power[0] for i in range(nd)], np.float64)
    np.testing.assert_allclose(power, 3.0 * a_proj / n, rtol=2e-6)
    assert abs(power.sum() - 3.0 * a_proj) < 1e-4 * 3.0 * a_proj
    raw = np.frombuffer(beams, np.uint8, count=48 * nb).reshape(nb, 48)
    start = raw[:, 0:12].view(np.float32).astype(np.float64)
    offs = raw[:, 36:40].view(np.float32)[:, 0]
    didx = raw[:, 32:36].view(np.uint32)[:, 0]
    first = offs == 0.0
    assert np.array_equal(np.sort(didx[first]), np.arange(n))
    e = start[first]
    entry = np.where(L > 0, b0, b1)      # the light enters through the faces it points away from
    on_face = np.abs(e - entry) == 0.0
    assert on_face.any(axis=1).all()
    counts = np.array([on_face[:, a].sum() for a in range(3)], np.float64)
    expect = n * np.abs(L) * face_area / a_proj
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    assert chi2 < 13.8, (counts, expect)     # 2 dof, p = 0.001
    for a in range(3):
        sel = e[on_face[:, a] & ~on_face[:, [k for k in range(3) if k != a]].any(axis=1)]
        for k in range(3):
            if k == a:
                continue
            u = (sel[:, k] - b0[k]) / ext[k]
            assert stats.kstest(u, "uniform").pvalue > 1e-3, (a, k)
