"""Pins of the CPU oracle (CPU only).

The reference ships no tests, golden images or fixtures (Tests.cpp asserts
nothing and is never called), so the oracle is pinned by known-answer values
from the literature, closed forms, invariants and brute force on tiny inputs.
Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, swapped operand) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

f3 = C.c_float * 3


# ------------------------------------------------------------------- RNG
def test_xorshift32_marsaglia_kat(O):
    # Marsaglia, "Xorshift RNGs", JSS 8(14), 2003: (13,17,5) from 2463534242
    s, out = 2463534242, []
    for _ in range(3):
        s = O.lib.or_xorshift32(s)
        out.append(s)
    assert out == [723471715, 2497366906, 2064144800]


def py_xorshift(s):
    s ^= (s << 13) & 0xFFFFFFFF
    s ^= s >> 17
    s ^= (s << 5) & 0xFFFFFFFF
    return s


@pytest.mark.parametrize("seed,pid,expect", [
    (100, 1, [1736819584, 557851450, 2177451145]),
    (100, 2501, [968489561, 3635019040, 113007206]),
    (41, 2609, [1232270677, 3973286297, 570165833]),
    (100, 0, [0, 0, 0]),
])
def test_initialize_random_stream(O, seed, pid, expect):
    """initializeRandom(seed * id): state = seed*id mod 2^32, then
    (state % 7 + 2) warm-up draws (PathTracer.comp:235-241, :368)."""
    out = (C.c_uint32 * 3)()
    O.lib.or_rng_stream(seed, pid, 3, out)
    assert list(out) == expect
    s = (seed * pid) & 0xFFFFFFFF
    for _ in range(s % 7 + 2):
        s = py_xorshift(s)
    ref = []
    for _ in range(3):
        s = py_xorshift(s)
        ref.append(s)
    assert ref == expect


def test_msvc_rand(O):
    out = (C.c_int32 * 8)()
    O.lib.or_msvc_rand(1, 3, out)
    assert list(out)[:3] == [41, 18467, 6334]  # the well-known MSVC rand() after srand(1)
    O.lib.or_msvc_rand(0, 8, out)
    assert list(out) == [38, 7719, 21238, 2437, 8855, 11797, 8365, 32285]


# ------------------------------------------------------------------ math
def ulps(a, ref):
    a = np.asarray(a, np.float64)
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(a - ref) / sp


def test_logf_accuracy(O):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(20000).astype(np.float32), np.float32(2.0) ** -np.arange(0, 149, dtype=np.float32),
                         np.array([1.0, 0.5, 0.70710677, 1.4142135, 3.0e-39], np.float32)])
    xs = xs[xs > 0]
    got = np.array([O.lib.or_logf(float(x)) for x in xs], np.float32)
    ref = np.log(xs.astype(np.float64))
    u = ulps(got, ref)
    u[ref == 0] = np.abs(got[ref == 0])
    assert u.max() <= 2.0
    assert O.lib.or_logf(0.0) == -math.inf
    assert O.lib.or_logf(1.0) == 0.0


def test_expf_accuracy(O):
    xs = np.linspace(-86.0, 10.0, 20001).astype(np.float32)
    got = np.array([O.lib.or_expf(float(x)) for x in xs], np.float32)
    assert ulps(got, np.exp(xs.astype(np.float64))).max() <= 2.0
    assert O.lib.or_expf(0.0) == 1.0
    assert O.lib.or_expf(-100.0) == 0.0


def test_sincosf_accuracy(O):
    xs = np.linspace(0.0, 2 * np.pi, 20001).astype(np.float32)
    s, c = C.c_float(), C.c_float()
    es = ec = 0.0
    for x in xs:
        O.lib.or_sincosf(float(x), C.byref(s), C.byref(c))
        es = max(es, abs(s.value - math.sin(float(x))))
        ec = max(ec, abs(c.value - math.cos(float(x))))
    assert es < 2.5e-7 and ec < 2.5e-7


def ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got - ref) / ulp


def test_acos_atan2_accuracy(O):
    """o_acosf and o_atan2f (omath.h) within 3 ULP of float64 arccos / arctan2 (acos
    reaches 2.5 ULP just below 1, where sqrt((1-x)(1+x)) is small),
    every octant and the axes."""
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(-1.0, 1.0, 200000), [-1.0, 1.0, 0.0, 0.5, -0.5, 0.7071068]]).astype(np.float32)
    got = O.math_array(5, x)
    assert ulp_err(got, np.arccos(x.astype(np.float64))).max() <= 3.0
    y = (rng.standard_normal(200000) * 10.0 ** rng.uniform(-4, 4, 200000)).astype(np.float32)
    xx = (rng.standard_normal(200000) * 10.0 ** rng.uniform(-4, 4, 200000)).astype(np.float32)
    pairs = np.empty(400000, np.float32)
    pairs[0::2], pairs[1::2] = y, xx
    got = O.math_array(6, pairs)[0::2]
    ref = np.arctan2(y.astype(np.float64), xx.astype(np.float64))
    big = np.abs(ref) > 1e-30
    assert ulp_err(got[big], ref[big]).max() <= 3.0
    axes = np.array([0.0, 1.0, 1.0, 0.0, 0.0, -1.0, -1.0, 0.0, 0.0, 0.0], np.float32)
    np.testing.assert_allclose(O.math_array(6, axes)[0::2], [0.0, np.pi / 2, np.pi, -np.pi / 2, 0.0], rtol=1e-7)


def test_math_array_domain(O):
    """sin/cos are defined on |x| <= 2^17 and NaN beyond; the array entry
    point is the scalar functions element by element, sqrt is IEEE's."""
    x = np.array([131072.0, -131072.0, 131088.0, np.inf, -np.inf, np.nan, 3.0], np.float32)
    for op in (3, 4):
        r = O.math_array(op, x)
        assert np.isfinite(r[:2]).all() and np.isnan(r[2:6]).all()
        s, c = C.c_float(), C.c_float()
        O.lib.or_sincosf(3.0, C.byref(s), C.byref(c))
        assert r[6] == (s.value if op == 3 else c.value)
    y = np.random.default_rng(2).random(1000, dtype=np.float32) * 100
    assert np.array_equal(O.math_array(0, y), np.sqrt(y))
    assert np.array_equal(O.math_array(1, y), np.array([O.lib.or_logf(float(v)) for v in y], np.float32))
    assert np.array_equal(O.math_array(2, -y), np.array([O.lib.or_expf(float(-v)) for v in y], np.float32))


def test_f16_conversion_matches_ieee(O):
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.random(5000).astype(np.float32),
        (rng.random(2000) * 70000).astype(np.float32),
        (rng.random(2000) * 1e-4).astype(np.float32),   # half subnormals
        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                  1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf], np.float32),
    ])
    got = O.f32_to_f16(vals)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, ref)
    back = np.array([O.lib.or_f16_to_f32(int(h)) for h in ref], np.float32)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(back, vals.astype(np.float16).astype(np.float32))


# ---------------------------------------------------------------- Morton
def test_morton_kats(O):
    assert O.lib.or_expand_bits10(1023) == 0x09249249
    assert O.lib.or_expand_bits10(1) == 1 and O.lib.or_expand_bits10(2) == 8
    b0, b1 = f3(0, 0, 0), f3(1, 1, 1)
    assert O.lib.or_morton30(f3(0.99999, 0, 0), b0, b1) == 0x24924924  # x = 1023
    assert O.lib.or_morton30(f3(0, 0, 0.99999), b0, b1) == 0x09249249
    assert O.lib.or_morton30(f3(-5, 7, 0.5), b0, b1) == (O.lib.or_expand_bits10(1023) << 1) + O.lib.or_expand_bits10(512)
    # reference 24-bit form: (3,2,1) voxels of 256 units -> 53 (PPB_PT.comp:326-333)
    assert O.lib.or_morton24_ref(f3(3 * 256 + 1, 2 * 256 + 1, 1 * 256 + 1), f3(0, 0, 0)) == 53


def test_expand8_equals_expand10_below_256(O):
    """The reference's 8-bit expandBits is Karras's 10-bit function on [0,255]."""
    b0 = f3(0, 0, 0)
    for v in range(256):
        e10 = O.lib.or_expand_bits10(v)
        assert O.lib.or_morton24_ref(f3(v * 256 + 1, 0, 0), b0) == e10 << 2
        # interleave: bit k of v lands at bit 3k
        assert e10 == sum(((v >> k) & 1) << (3 * k) for k in range(8))


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("n", [1, 2, 17, 5000, 100003])
def test_radix_sort_is_stable_sort(O, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 30, n, dtype=np.uint32)
    keys[: n // 3] = keys[0]  # duplicates
    rng.shuffle(keys)
    ko, po = O.radix_sort(keys)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(po, ref)
    np.testing.assert_array_equal(ko, keys[ref])


# ------------------------------------------------------------- hierarchy
def tree_arrays(nodes, n):
    a = np.frombuffer(nodes, np.uint32).reshape(-1, 16)[: 2 * n - 1]
    return a[:, 0], a[:, 1], a[:, 2], a[:, 3]


@pytest.mark.parametrize("n,dups", [(2, 0), (3, 0), (8, 1), (1000, 0), (1000, 1), (4096, 0)])
def test_karras_equals_topdown_and_is_a_tree(O, n, dups):
    rng = np.random.default_rng(n + dups)
    keys = np.sort(rng.integers(0, 1 << 30, n, dtype=np.uint32))
    if dups:
        keys = np.sort(np.repeat(keys[: max(1, n // 4)], 4)[:n])
    a = np.frombuffer(O.karras(keys), np.uint8)
    b = np.frombuffer(O.karras(keys, topdown=True), np.uint8)
    assert (a == b).all()
    parent, left, right, leaf = tree_arrays(O.karras(keys), n)
    inner = n - 1
    assert parent[0] == 0xFFFFFFFF
    children = np.concatenate([left[:inner], right[:inner]])
    assert sorted(children.tolist()) == list(range(1, 2 * n - 1))  # every node but the root has one parent
    assert (leaf[inner:] == 1).all() and (leaf[:inner] == 0).all()
    for i in range(inner):
        assert parent[left[i]] == i and parent[right[i]] == i
    # leaves in order: an in-order walk visits leaf 0..n-1
    order, stack, cur = [], [], 0
    while stack or cur is not None:
        while cur is not None and cur < inner:
            stack.append(cur)
            cur = int(left[cur])
        if cur is not None:
            order.append(cur - inner)
        p = stack.pop() if stack else None
        cur = int(right[p]) if p is not None else None
        if p is None:
            break
    assert order == list(range(n))


def beams_random(S, n, seed):
    rng = np.random.default_rng(seed)
    beams = (S.PhotonBeam * n)()
    a = np.frombuffer(beams, np.uint8).reshape(n, 48)
    f, u = a.view(np.float32), a.view(np.uint32)
    st = rng.uniform(-0.4, 0.4, (n, 3)).astype(np.float32)
    st[:, 2] += 0.5
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f[:, 0:3] = st
    f[:, 4:7] = st + d * np.float32(0.003)
    f[:, 7] = 0.002
    u[:, 8] = 0
    return beams


def test_leaf_bounds_closed_form(O, S):
    b = S.PhotonBeam()
    b.startPos[:] = (0.1, 0.2, 0.3)
    b.endPos[:] = (0.5, 0.2, 0.3)
    b.radius = 0.25
    lo, hi = f3(), f3()
    O.lib.or_leaf_bounds(C.byref(b), lo, hi)
    # axis along x: the disk caps span +-r in y and z, nothing in x
    assert list(lo) == [np.float32(0.1), np.float32(0.2) - np.float32(0.25), np.float32(0.3) - np.float32(0.25)]
    assert list(hi) == [np.float32(0.5), np.float32(0.2) + np.float32(0.25), np.float32(0.3) + np.float32(0.25)]
    # a slanted cylinder: every point of both rims lies in the box
    b.endPos[:] = (0.4, 0.6, 0.1)
    O.lib.or_leaf_bounds(C.byref(b), lo, hi)
    pa, pb = np.array(b.startPos, np.float64), np.array(b.endPos, np.float64)
    ax = (pb - pa) / np.linalg.norm(pb - pa)
    u = np.cross(ax, [1.0, 0, 0]); u /= np.linalg.norm(u)
    v = np.cross(ax, u)
    for t in np.linspace(0, 2 * np.pi, 64):
        for c in (pa, pb):
            p = c + 0.25 * (np.cos(t) * u + np.sin(t) * v)
            assert (p >= np.array(lo) - 1e-6).all() and (p <= np.array(hi) + 1e-6).all()
    # and the box is tight: extremes are attained
    ext = 0.25 * np.sqrt(1 - ax ** 2)
    np.testing.assert_allclose(np.array(lo), np.minimum(pa, pb) - ext, rtol=1e-6)


def test_refit_unions(O, S):
    n = 3000
    beams = beams_random(S, n, 5)
    keys, perm, nodes = O.lbvh_build(beams, (-0.5, -0.5, 0.0), (0.5, 0.5, 1.0))
    a = np.frombuffer(nodes, np.uint8).reshape(-1, 64)
    u = a[:, :16].view(np.uint32)
    bnd = a[:, 16:48].view(np.float32).reshape(-1, 2, 4)
    inner = n - 1
    for i in range(inner):
        l, r = u[i, 1], u[i, 2]
        np.testing.assert_array_equal(bnd[i, 0], np.minimum(bnd[l, 0], bnd[r, 0]))
        np.testing.assert_array_equal(bnd[i, 1], np.maximum(bnd[l, 1], bnd[r, 1]))
    np.testing.assert_array_equal(bnd[0, 0], bnd[inner:, 0].min(axis=0))
    np.testing.assert_array_equal(bnd[0, 1], bnd[inner:, 1].max(axis=0))
    # leaf i is the box of the i-th beam in sorted order
    lo, hi = f3(), f3()
    for i in (0, 17, n - 1):
        O.lib.or_leaf_bounds(C.byref(beams[int(perm[i])]), lo, hi)
        assert list(bnd[inner + i, 0, :3]) == list(lo) and list(bnd[inner + i, 1, :3]) == list(hi)
    processed = a[:, 48:52].view(np.uint32)[:, 0]
    assert (processed[:inner] == 2).all() and (processed[inner:] == 0).all()


# ------------------------------------------------------------- rendering
def unit_cloud(S, pkg, grid, density_scaling=64.0):
    c = S.CloudProperties()
    pkg.load().cr_cloud_default(C.byref(c))
    nz, ny, nx = grid.shape
    c.baseScaling, c.densityScaling = 1.0, density_scaling
    c.maxExtinction = max(float(grid.max()), 0.01)
    c.voxelCount[:] = (nx, ny, nz, 0)
    c.bounds[0][:] = (-0.5, -0.5, 0.0, 0.0)
    c.bounds[1][:] = (0.5, 0.5, 1.0, 0.0)
    return c


def test_pt_density_zero_gives_background(O, S, pkg, cfg):
    g = np.ones((8, 8, 8), np.float32)
    c = unit_cloud(S, pkg, g, density_scaling=0.0)
    cam = cfg.camera(16, 12)
    ext = cfg.render_ext()
    s = O.scene(g, c, cam, cfg.parameters(0.85), cfg.shadow_volume(c), ext, sv=np.ones((4, 4, 4), np.float32))
    img, _ = O.pt_render(s, cfg.frames(1))
    pos, d, bg = f3(), f3(), (C.c_float * 4)()
    for (x, y) in [(0, 0), (5, 7), (15, 11)]:
        O.lib.or_camera_ray(C.byref(cam), x, y, pos, d)
        O.lib.or_background(C.byref(ext), d, bg)
        assert list(img[y, x]) == list(bg)


def test_background_gradient_stops(O, cfg):
    """sampleBackground: BG_COLORS[i] exactly at BG_DISTS[i] (PathTracer.comp:22-37, :104-113)."""
    ext = cfg.render_ext()
    out = (C.c_float * 4)()
    stops = {-1.0: (0.0, 0.0, 0.02, 1.0), 0.0: (0.7, 0.9, 1.0, 1.0), 1.0: (0.01, 0.1, 0.7, 1.0)}
    for y, col in stops.items():
        O.lib.or_background(C.byref(ext), f3(0.0, y, 0.0), out)
        np.testing.assert_allclose(list(out), col, rtol=1e-6)


def test_pt_white_furnace_oracle(O, S, pkg, cfg):
    rng = np.random.default_rng(2)
    g = rng.random((12, 12, 12), dtype=np.float32)
    c = unit_cloud(S, pkg, g, 20.0)
    ext = cfg.render_ext(albedo=1.0, sky=(0.3, 0.5, 0.7, 1.0))
    s = O.scene(g, c, cfg.camera(12, 10), cfg.parameters(0.6, light_intensity=0.0), cfg.shadow_volume(c), ext,
                sv=np.ones((4, 4, 4), np.float32))
    img, _ = O.pt_render(s, cfg.frames(3))
    np.testing.assert_array_equal(img, np.broadcast_to(np.array([0.3, 0.5, 0.7, 1.0], np.float32), img.shape))


@pytest.mark.parametrize("tracking", [0, 1])
def test_pt_transmittance_closed_form(O, S, pkg, cfg, tracking):
    """No scattering counted (bounce cap 0), no sun, constant sky c, homogeneous
    medium sigma: E[pixel] = c * exp(-sigma * chord) along the camera ray."""
    g = np.ones((16, 16, 16), np.float32)
    sigma = 3.0
    c = unit_cloud(S, pkg, g, sigma)
    cam = cfg.camera(8, 8, fov=60.0, position=(0.0, 0.0, -0.5))
    ext = cfg.render_ext(albedo=1.0, bounce_cap=0, sky=(1.0, 1.0, 1.0, 1.0), tracking=tracking)
    s = O.scene(g, c, cam, cfg.parameters(0.85, light_intensity=0.0), cfg.shadow_volume(c), ext,
                sv=np.ones((4, 4, 4), np.float32))
    nfr = 4000
    img, _ = O.pt_render(s, cfg.frames(nfr), threads=8)
    pos, d = f3(), f3()
    tmin, tmax = C.c_float(), C.c_float()
    exp_sum = var_sum = 0.0
    # pixel (0,0) has id 0: its xorshift state is 0 forever, zeta = 0 and the
    # ray always escapes (kept for stream compatibility)
    assert img[0, 0, 0] == 1.0
    img[0, 0, 0] = 0.0
    for y in range(8):
        for x in range(8):
            if x == 0 and y == 0:
                continue
            O.lib.or_camera_ray(C.byref(cam), x, y, pos, d)
            hit = O.lib.or_intersect_box(f3(-0.5, -0.5, 0.0), f3(0.5, 0.5, 1.0), pos, d, C.byref(tmin), C.byref(tmax))
            chord = (tmax.value - tmin.value) if hit and tmax.value > 0 else 0.0
            p = math.exp(-sigma * chord)
            exp_sum += p
            var_sum += p * (1 - p) / nfr
    got = float(img[..., 0].sum())
    assert abs(got - exp_sum) < 4.0 * math.sqrt(var_sum) + 1e-3


def test_hg_sampling_statistics(O, S, cfg):
    """scatterRay draws cos(theta) from Henyey-Greenstein about the incoming
    direction: E[cos] = g and the cos histogram matches the pdf."""
    g = 0.7
    p = cfg.parameters(g)
    n_in = f3(0.36, -0.48, 0.8)
    out = f3()
    st = C.c_uint32()
    cosines = []
    s = 12345
    for i in range(40000):
        O.lib.or_scatter_dir(C.byref(p), n_in, s, out, C.byref(st))
        s = st.value
        v = np.array(out, np.float64)
        cosines.append(float(np.dot(v, np.array(n_in, np.float64)) / np.linalg.norm(v)))
    cosines = np.array(cosines)
    assert abs(cosines.mean() - g) < 4 * cosines.std() / math.sqrt(len(cosines))
    edges = np.linspace(-1, 1, 21)
    hist, _ = np.histogram(cosines, edges)
    cdf = lambda m: (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * m) - 1 / (1 + g))
    expect = np.diff(cdf(edges)) * len(cosines)
    chi2 = ((hist - expect) ** 2 / expect).sum()
    assert chi2 < 45.0  # 19 dof, p ~ 1e-3


def test_phase_pdf_normalised(O, cfg):
    for g in (0.0, 0.3, 0.85, -0.5):
        p = cfg.parameters(g)
        xs, ws = np.polynomial.legendre.leggauss(400)
        tot = 0.0
        a = f3(0.0, 0.0, 1.0)
        for x, w in zip(xs, ws):
            b = f3(float(math.sqrt(1 - x * x)), 0.0, float(x))
            tot += w * O.lib.or_phase_pdf(C.byref(p), a, b)
        assert abs(2 * math.pi * tot - 1.0) < 2e-3


def test_shadow_volume_closed_form(O, S, pkg, cfg):
    """Homogeneous density: T(x,y,z) = exp(-h * sigma * (number of lattice
    points k <= z inside the box)) (ShadowVolume.comp:84-93)."""
    g = np.full((8, 8, 8), 0.5, np.float32)
    c = unit_cloud(S, pkg, g, 4.0)
    svp = cfg.shadow_volume(c, (0.3, -1.0, 0.2), 12)
    s = O.scene(g, c, cfg.camera(8, 8), cfg.parameters(0.85), svp, cfg.render_ext())
    sv = O.sv_build(s)
    P = svp
    b0, b1 = np.array(c.bounds[0][:3]), np.array(c.bounds[1][:3])
    sig = 0.5 * 4.0
    for (x, y) in [(6, 6), (3, 8), (0, 0), (11, 5)]:
        k = 0
        for z in range(12):
            p = np.array(P.bounds[0][:3]) + (x * np.array(P.right[:3]) + y * np.array(P.up[:3]) +
                                             z * np.array(P.lightDirection[:3])) * P.voxelSize
            k += int(((p >= b0 - 1e-7) & (p <= b1 + 1e-7)).all())
            assert abs(sv[z, y, x] - math.exp(-P.voxelSize * sig * k)) < 2e-5


def test_trilinear_affine_and_clamp(O):
    nz, ny, nx = 5, 6, 7
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = (0.25 * x + 0.5 * y - 0.125 * z + 1.0).astype(np.float32)
    dims = (C.c_uint32 * 3)(nx, ny, nz)
    rng = np.random.default_rng(3)
    for _ in range(200):
        u = rng.uniform(0.5, [nx - 0.5, ny - 0.5, nz - 0.5])
        n = f3(*(u / np.array([nx, ny, nz])))
        v = O.lib.or_trilinear_f32(g.ctypes.data, dims, n)
        ref = 0.25 * (u[0] - 0.5) + 0.5 * (u[1] - 0.5) - 0.125 * (u[2] - 0.5) + 1.0
        assert abs(v - ref) < 2e-6
    # clamp to edge: outside the texel centres the edge value holds
    assert O.lib.or_trilinear_f32(g.ctypes.data, dims, f3(-0.3, 0.0, 0.0)) == g[0, 0, 0]
    assert O.lib.or_trilinear_f32(g.ctypes.data, dims, f3(1.2, 1.0, 1.0)) == g[-1, -1, -1]


def test_fbm_generator_properties(O, cfg):
    p = cfg.fbm((40, 32, 24), 1.0 / 40, seed=9, octaves=6)
    a = O.fbm_grid(p)
    b = O.fbm_grid(p)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= 0.0 and a.max() <= 1.0
    assert a[0, 0, 0] == 0.0  # corners are outside the cumulus envelope
    for (x, y, z) in [(3, 4, 5), (20, 16, 12)]:
        assert O.lib.or_fbm_voxel(C.byref(p), x, y, z) == a[z, y, x]
    p2 = cfg.fbm((40, 32, 24), 1.0 / 40, seed=10, octaves=6)
    assert not np.array_equal(O.fbm_grid(p2), a)


def test_gather_bvh_equals_brute_force_oracle(O, S, pkg, cfg):
    rng = np.random.default_rng(4)
    g = rng.random((10, 10, 10), dtype=np.float32)
    c = unit_cloud(S, pkg, g, 8.0)
    cam = cfg.camera(12, 10, position=(0.0, 0.0, -0.3))
    svp = cfg.shadow_volume(c, tuple(cam.forward), 8)
    s = O.scene(g, c, cam, cfg.parameters(0.7), svp, cfg.render_ext(), sv=np.full((8, 8, 8), 0.8, np.float32))
    p = cfg.beam_params(60, seed=5, radius=0.05, albedo=0.9)
    beams, data, nb, nd = O.trace_beams(s, p, 1 << 16)
    keys, perm, nodes = O.lbvh_build(beams, c.bounds[0], c.bounds[1])
    sb = (S.PhotonBeam * nb)(*[beams[int(i)] for i in perm])
    bs = O.beam_scene(s, sb, data, nodes, nb, 0, 0, 0.9)
    nonzero = 0
    for y in range(10):
        for x in range(12):
            a = O.gather_pixel(bs, x, y, 38)
            b = O.gather_pixel(bs, x, y, 38, brute=True)
            np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-7)
            nonzero += a[0] > 0
    assert nonzero > 10


def test_gather_axis_parallel_rays_cull(O, S, pkg, cfg):
    """Rays with a direction component exactly 0 (the 12 x 10 camera's centre
    column, centre row and centre pixel, whose ray is (0, 0, 1)) still cull
    the boxes they lie outside of: intersectBox's slab formula gives NaN on
    such an axis, which would let the ray visit every node.  The culled
    traversal still equals the loop over every beam."""
    rng = np.random.default_rng(4)
    g = rng.random((10, 10, 10), dtype=np.float32)
    c = unit_cloud(S, pkg, g, 8.0)
    cam = cfg.camera(12, 10, position=(0.0, 0.0, -0.3))
    svp = cfg.shadow_volume(c, tuple(cam.forward), 8)
    s = O.scene(g, c, cam, cfg.parameters(0.7), svp, cfg.render_ext(), sv=np.full((8, 8, 8), 0.8, np.float32))
    p = cfg.beam_params(400, seed=5, radius=0.01, albedo=0.9)
    beams, data, nb, nd = O.trace_beams(s, p, 1 << 16)
    assert nb == 1 << 16
    keys, perm, nodes = O.lbvh_build(beams, c.bounds[0], c.bounds[1])
    sb = (S.PhotonBeam * nb)(*[beams[int(i)] for i in perm])
    bs = O.beam_scene(s, sb, data, nodes, nb, 0, 0, 0.9)
    for (x, y) in [(6, 5), (6, 1), (6, 8), (2, 5), (10, 5)]:
        r, visited, tested = O.gather_pixel(bs, x, y, 38, stats=True)
        assert visited < (nb - 1) // 100, ((x, y), visited)
        np.testing.assert_allclose(r, O.gather_pixel(bs, x, y, 38, brute=True), rtol=1e-4, atol=1e-7)


def test_ppb_update_radius_pinned(O):
    """UpdateRadius (RenderTechniquePPB.cpp:530-546) against what its formula
    fixes: restart at frames 0 and 1, a sum of exactly B ones for alpha = 1,
    the single-beam ratio (n - 1 + alpha) / n, the 56-term sum in float64,
    and ratios that round to 1 once c + j exceeds 2^24."""
    upd = O.lib.or_ppb_update_radius
    f32 = np.float32
    for f in (0, 1):
        assert upd(f, 56, 0.8, 0.002, 5.0) == f32(0.002)
    for f in (2, 3, 1000):
        for B in (1, 7, 56):
            assert upd(f, B, 1.0, 0.0, 0.25) == 0.25 * B
    prev = f32(0.01)
    for f in range(2, 3000):
        r = upd(f, 1, 0.8, 0.0, float(prev))
        exact = float(prev) * (f - 1 + float(f32(0.8))) / f
        assert abs(r - exact) <= 3e-7 * exact, f
        prev = f32(r)
    for f, B in ((2, 56), (3, 56), (40, 56), (9, 1000)):
        c = (f - 2) * B
        exact = 0.5 * sum((c + j + float(f32(0.8))) / (c + j + 1.0) for j in range(1, B + 1))
        assert abs(upd(f, B, 0.8, 0.0, 0.5) - exact) <= 4e-7 * B * exact, (f, B)
    # c = (2^27 - 2) * 56 wraps in 32 bits to 3221225360; past 2^24 each ratio rounds to 1
    assert upd(1 << 27, 56, 0.8, 0.0, 0.5) == 0.5 * 56


def test_sqrt_rounding_correction_exhaustive(O):
    """The FMA-residual correction the kernels apply to the hardware square
    root returns the correctly rounded root for every input in [2^-100,
    FLT_MAX] from each of the three start values within one step of it (all
    2^31-odd inputs; a few seconds with OpenMP).  The device side of the same
    claim is test_gpu_math.py's comparison with IEEE sqrt."""
    lo = int(np.float32(2.0 ** -100).view(np.uint32))
    hi = 0x7F7FFFFF
    assert O.lib.or_sqrt_correction_errors(lo, hi) == 0
