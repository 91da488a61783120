"""Pins of the photon-mapping oracle (oracle/oracle_ppm.cpp), CPU only.

Each check is fixed independently of the oracle's code: the radius schedule
by its product form, the photon grid by a brute-force stable grouping, one
photon's contribution to the density estimate by the textbook formula
(Jensen & Christensen 1998: f * Phi / (4/3 pi r^3 sigma_s)), and the ray
march by the path tracer's single-scattering estimate of the same integral.
"""
import ctypes as C
import math

import numpy as np
import pytest

f3 = C.c_float * 3


def test_ppm_update_radius_product_form(pkg, S):
    """RenderTechniquePPM::UpdateRadius (RenderTechniquePPM.cpp:302-312):
    r_0 = r_1 = r_init, r_n = r_init * prod_{k=2..n} ((k - 1 + a) / k)^(1/3)
    with the exponent .33333."""
    lib = pkg.load()
    a = float(np.float32(S.CR_PPM_ALPHA))
    for f in (0, 1):
        assert lib.cr_ppm_update_radius(f, a, 0.01, 7.0) == np.float32(0.01)
    r = np.float32(0.01)
    exact = 0.01
    for n in range(2, 400):
        r = np.float32(lib.cr_ppm_update_radius(n, a, 0.01, float(r)))
        exact *= ((n - 1 + a) / n) ** float(np.float32(0.33333))
        assert abs(r - exact) <= 1e-5 * exact, n
    assert abs(r / 0.01 - 400 ** (-(1 - a) / 3)) < 0.02   # r ~ n^(-(1 - alpha) / 3)


def test_photon_map_set_bounds(pkg, S):
    """SetBounds (UniformBuffers.h:225-243): a cube from the lower corner with
    the largest extent as side; voxelSize = side / voxelCount.x."""
    lib = pkg.load()
    pm = S.PhotonMapProperties()
    lib.cr_photon_map_default(C.byref(pm))
    assert tuple(pm.voxelCount) == (100, 100, 100) and pm.stepSize == 10.0 and pm.photonSize == 48
    assert abs(pm.lightDirection[0] - 2 ** -0.5) < 1e-7 and abs(pm.lightDirection[1] + 2 ** -0.5) < 1e-7
    b0 = (C.c_float * 4)(-0.5, 0.0, -0.25, 0.0)
    b1 = (C.c_float * 4)(0.5, 0.5, 1.75, 0.0)
    lib.cr_photon_map_set_bounds(C.byref(pm), b0, b1)
    assert list(pm.bounds[0]) == [-0.5, 0.0, -0.25, 0.0]
    assert list(pm.bounds[1]) == [1.5, 2.0, 1.75, 2.0]
    assert pm.voxelSize == np.float32(2.0) / np.float32(100.0)


def photon_array(S, n, rng):
    ph = (S.Photon * n)()
    raw = np.frombuffer(ph, np.float32).reshape(n, 12)
    raw[:] = rng.random((n, 12), dtype=np.float32)
    raw[:, 0] = np.arange(n)       # the emission index, to check the order
    return ph


def test_ppm_build_keeps_emission_order(O, S):
    """The grid groups deposits by cell, in emission order within a cell
    (a stable grouping); arrivals count every deposit of a cell."""
    rng = np.random.default_rng(5)
    n, n_cells = 5000, 64
    cells = rng.integers(0, n_cells, n).astype(np.uint32)
    cells[:100] = 7                  # one cell far over the 32 it keeps
    ph = photon_array(S, n, rng)
    arrivals, start, sorted_ = O.ppm_build(ph, cells, n, n_cells)
    np.testing.assert_array_equal(arrivals, np.bincount(cells, minlength=n_cells))
    got = np.frombuffer(sorted_, np.float32).reshape(n, 12)[:, 0].astype(np.int64)
    for c in range(n_cells):
        run = got[start[c]:start[c] + arrivals[c]]
        np.testing.assert_array_equal(run, np.flatnonzero(cells == c))


def unit_scene(O, S, pkg, cfg, density=1.0, sigma=2.0, g=0.0, sky=(0.0, 0.0, 0.0, 0.0), W=16, H=16, sv=None,
               camera_z=-0.8):
    lib = pkg.load()
    grid = np.full((16, 16, 16), density, np.float32)
    c = S.CloudProperties()
    lib.cr_cloud_default(C.byref(c))
    c.baseScaling, c.densityScaling = 1.0, sigma
    c.maxExtinction = max(density, 0.01)
    c.voxelCount[:] = (16, 16, 16, 0)
    c.bounds[0][:] = (-0.5, -0.5, 0.0, 0.0)
    c.bounds[1][:] = (0.5, 0.5, 1.0, 0.0)
    cam = cfg.camera(W, H, position=(0.0, 0.0, camera_z))
    params = cfg.parameters(g)
    ext = cfg.render_ext(sky=sky)
    pm = S.PhotonMapProperties()
    lib.cr_photon_map_default(C.byref(pm))
    lib.cr_photon_map_set_bounds(C.byref(pm), c.bounds[0], c.bounds[1])
    pm.stepSize = 0.01
    svp = cfg.shadow_volume(c, (1.0, -1.0, 0.0), 4)
    s = O.scene(grid, c, cam, params, svp, ext, sv=np.ones((4, 4, 4), np.float32) if sv is None else sv)
    return s, pm, c, cam, params, svp, ext, grid


@pytest.mark.parametrize("g", [0.0, 0.7])
def test_ppm_query_single_photon_closed_form(O, S, pkg, cfg, g):
    """One photon at distance d < r from the query point contributes
    HG(dir, photon dir) * Phi / (4/3 pi r^3 sigma_s) exactly; at d > r, or as
    the 33rd photon of its cell, nothing."""
    s, pm, *_ = unit_scene(O, S, pkg, cfg, g=g)
    n_cells = 100 ** 3
    pos = np.array([0.013, 0.021, 0.5])
    r, scatter = 0.01, 3.0
    # photon travelling along (sin t cos p, sin t sin p, cos t)
    theta, phi = 1.1, 4.0
    pdir = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
    view = np.array([0.0, 0.6, -0.8])
    ph = (S.Photon * 1)()
    ph[0].position[:] = tuple(pos + np.array([0.004, -0.003, 0.005])) + (0.0,)
    ph[0].power[:] = (2.0, 1.0, 0.5, 1.0)
    ph[0].theta, ph[0].phi = theta, phi
    cell = np.array([0], np.uint32)
    # the cell of the photon
    idx = [int((ph[0].position[k] - pm.bounds[0][k]) / pm.voxelSize) for k in range(3)]
    cell[0] = idx[0] + (idx[1] + idx[2] * 100) * 100
    arrivals, start, sorted_ = O.ppm_build(ph, cell, 1, n_cells)
    m = O.photon_map(arrivals, start, sorted_)
    got = O.ppm_query(s, pm, m, pos, view, scatter, r)
    cos_t = float(np.dot(view, pdir))
    hg = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos_t) ** 1.5)
    expect = hg * np.array([2.0, 1.0, 0.5, 1.0]) / (4.0 / 3.0 * math.pi * r ** 3 * scatter)
    np.testing.assert_allclose(got, expect, rtol=1e-5)
    assert (O.ppm_query(s, pm, m, pos + np.array([0.02, 0.0, 0.0]), view, scatter, r) == 0).all()
    assert (O.ppm_query(s, pm, m, pos, view, 0.0, r) == 0).all()
    # 33 photons in one cell: only the first 32 count
    ph33 = (S.Photon * 33)()
    for i in range(33):
        ph33[i] = ph[0]
    cells33 = np.full(33, cell[0], np.uint32)
    arrivals, start, sorted_ = O.ppm_build(ph33, cells33, 33, n_cells)
    assert arrivals[cell[0]] == 33
    got33 = O.ppm_query(s, pm, O.photon_map(arrivals, start, sorted_), pos, view, scatter, r)
    np.testing.assert_allclose(got33, 32 * expect, rtol=1e-5)


def test_ppm_trace_deposits(O, S, pkg, cfg):
    """Deposits lie in the cloud, in their own cell, carry unit directions
    (theta in [0, pi], phi in [0, 2 pi)); sun photons carry I A / (N p) / 2
    (the reference's halving) in every channel; no deposit on a path's first
    collision, so a medium too thin to scatter twice deposits nothing."""
    s, pm, c, *_ = unit_scene(O, S, pkg, cfg, sigma=6.0)
    p = S.PpmParams()
    pkg.load().cr_ppm_params_default(C.byref(p))
    p.numPaths, p.ambientStep = 3000, 0.1
    photons, cells, n = O.ppm_trace(s, pm, p, 41)
    assert n > 1000
    raw = np.frombuffer(photons, np.float32, count=12 * n).reshape(n, 12)
    pos, power, phi, theta = raw[:, :3], raw[:, 4:8], raw[:, 8], raw[:, 9]
    assert (pos >= np.array(c.bounds[0][:3]) - 1e-6).all() and (pos <= np.array(c.bounds[1][:3]) + 1e-6).all()
    idx = np.clip(((pos - np.array(pm.bounds[0][:3])) / pm.voxelSize).astype(np.int64), 0, 99)
    np.testing.assert_array_equal(cells, idx[:, 0] + (idx[:, 1] + idx[:, 2] * 100) * 100)
    assert (theta >= 0).all() and (theta <= math.pi + 1e-6).all() and (phi >= 0).all() and (phi < 2 * math.pi).all()
    L = np.array(pm.lightDirection[:3], np.float64)
    area = float((np.abs(L) * np.array([1.0, 1.0, 1.0])).sum())   # unit cube faces
    sun = (power[:, 0] > 0) & np.isclose(power[:, 0], power[:, 2]) & np.isclose(power[:, 0], power[:, 3])
    np.testing.assert_allclose(power[sun, 0], 5.0 * area / p.numPaths / 0.5 * 0.5, rtol=1e-5)
    thin, *_ = unit_scene(O, S, pkg, cfg, sigma=1e-3)
    _, _, n_thin = O.ppm_trace(thin, pm, p, 41)
    assert n_thin == 0


def test_ppm_estimate_single_scattering_matches_path_tracer(O, S, pkg, cfg):
    """With no photons and a black sky the estimate is the ray-marched single
    scattering of sunlight through the shadow volume, whose integral the path
    tracer (bounce cap 1, black sky, shadow-volume NEE over the same shadow
    volume) estimates without discretisation.  The march is PPM_PE.comp's
    rectangle rule (each step's own term unattenuated), first order in the
    step: its excess halves with the step (3.7% at 1/100 of the box, 2.0% at
    1/200 here), and the Richardson extrapolation 2 E(h/2) - E(h) matches the
    path tracer within 1.5% on the mean of the central 8x8 pixels."""
    s, pm, c, cam, params, svp, ext, grid = unit_scene(O, S, pkg, cfg, sigma=2.0, g=0.5, W=48, H=48, camera_z=-3.0)
    svp = cfg.shadow_volume(c, (1.0, -1.0, 0.3), 128)
    s = O.scene(grid, c, cam, params, svp, ext)
    sv = O.sv_build(s)
    s = O.with_sv(s, sv)
    p = S.PpmParams()
    pkg.load().cr_ppm_params_default(C.byref(p))
    p.ambientStep = 0.1
    n_cells = 100 ** 3
    empty = O.photon_map(np.zeros(n_cells, np.uint32), np.zeros(n_cells, np.uint32), (S.Photon * 1)())
    fr = [S.PushConstants(0.0, sd, i, 0.01, 0) for i, sd in enumerate(cfg.msvc_seeds(16))]
    est = {}
    for h in (0.01, 0.005):
        pm.stepSize = h
        est[h] = float(O.ppm_estimate(s, pm, empty, p, fr)[20:28, 20:28, 0].mean())
    ext_pt = cfg.render_ext(albedo=1.0, bounce_cap=1, sky=(0.0, 0.0, 0.0, 0.0))
    s_pt = O.with_sv(O.scene(grid, c, cam, params, svp, ext_pt), sv)
    pt, _ = O.pt_render_region(s_pt, cfg.frames(2048), 20, 20, 8, 8)
    b = float(pt[..., 0].mean())
    assert b > 1e-3
    g1, g2 = est[0.01] - b, est[0.005] - b
    assert g1 > 0 and 1.5 < g1 / g2 < 2.5, (est, b)
    assert abs((2 * est[0.005] - est[0.01]) / b - 1.0) < 0.015, (est, b)
