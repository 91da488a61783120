"""GPU parity of the photon-beam pipeline: beam producer, LBVH build
(Morton codes, sorted permutation, hierarchy, refit) and gather, bit-exact
against the oracle (oracle/oracle_lbvh.cpp, oracle/oracle_ppb.cpp)."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import assert_bitexact, oracle_scene_from_ctx

pytestmark = pytest.mark.gpu


def random_beams(S, n, seed, dup_frac=0.0):
    rng = np.random.default_rng(seed)
    beams = (S.PhotonBeam * n)()
    a = np.frombuffer(beams, dtype=np.uint8).reshape(n, 48)
    st = rng.uniform(-0.45, 0.45, (n, 3)).astype(np.float32)
    st[:, 2] += 0.5
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    en = (st + d * rng.uniform(0.0, 0.004, (n, 1))).astype(np.float32)
    if dup_frac > 0:
        k = int(n * dup_frac)
        st[:k] = st[0]
        en[:k] = en[0]
    f = a.view(np.float32)
    u = a.view(np.uint32)
    f[:, 0:3] = st
    f[:, 4:7] = en
    f[:, 7] = 0.002
    u[:, 8] = rng.integers(0, 4, n)
    u[:, 3] = 0
    data = (S.PhotonBeamData * 4)()
    for i in range(4):
        data[i].power[:] = (1.0, 1.0, 1.0, 1.0)
        data[i].trDistances[:] = tuple(np.linspace(0.001, 0.05, 16) * (i + 1))
    return beams, data


def assert_same_tree(gpu_nodes, ref_nodes, n):
    a = np.frombuffer(gpu_nodes, np.uint8).reshape(-1, 64)[: 2 * n - 1]
    b = np.frombuffer(ref_nodes, np.uint8).reshape(-1, 64)[: 2 * n - 1]
    if (a == b).all():
        return
    names = ["parent", "left", "right", "isLeaf", "lo.x", "lo.y", "lo.z", "lo.w", "hi.x", "hi.y", "hi.z", "hi.w",
             "processed", "p0", "p1", "p2"]
    au, bu = a.view(np.uint32), b.view(np.uint32)
    bad_nodes = np.nonzero((au != bu).any(axis=1))[0]
    i = int(bad_nodes[0])
    fields = [names[k] for k in range(16) if au[i, k] != bu[i, k]]
    raise AssertionError(f"{len(bad_nodes)} of {2 * n - 1} nodes differ; first node {i} (inner={n - 1}) fields "
                         f"{fields}: gpu={au[i].tolist()} oracle={bu[i].tolist()}")


def setup_c1(ctx, cfg, S, size=48):
    sc = cfg.c1(width=48, height=32, size=size, sv_axis=40)
    cloud, svp = sc.setup(ctx, build_sv=False)
    return sc, cloud, svp


@pytest.mark.parametrize("n,dup", [(1, 0.0), (2, 0.0), (3, 0.0), (1000, 0.0), (5000, 0.3), (70001, 0.0)])
def test_lbvh_bitexact_random(ctx, cfg, S, O, n, dup):
    sc, cloud, _ = setup_c1(ctx, cfg, S)
    beams, data = random_beams(S, n, seed=n, dup_frac=dup)
    ctx.upload_beams(beams, data)
    ctx.lbvh_build()
    keys, perm, nodes = ctx.lbvh_download()
    rk, rp, rn = O.lbvh_build(beams, cloud.bounds[0], cloud.bounds[1])
    codes = np.empty(n, np.uint32)
    O.lib.or_beam_morton30(beams, n, (C.c_float * 3)(*cloud.bounds[0][:3]), (C.c_float * 3)(*cloud.bounds[1][:3]),
                           codes.ctypes.data)
    assert np.array_equal(np.sort(keys), np.sort(codes)), "Morton codes differ (as multisets)"
    np.testing.assert_array_equal(keys, rk, err_msg="sorted keys")
    np.testing.assert_array_equal(perm, rp, err_msg="permutation")
    assert_same_tree(nodes, rn, n)
    # unique stable order: by key, ties by index
    codes = np.empty(n, np.uint32)
    O.lib.or_beam_morton30(beams, n, (C.c_float * 3)(*cloud.bounds[0][:3]), (C.c_float * 3)(*cloud.bounds[1][:3]),
                           codes.ctypes.data)
    np.testing.assert_array_equal(perm, np.argsort(codes, kind="stable"))


def test_trace_beams_bitexact_and_lbvh(ctx, cfg, S, O):
    sc, cloud, svp = setup_c1(ctx, cfg, S)
    p = cfg.beam_params(100, seed=38, radius=0.004, albedo=0.9, max_interactions=16)
    nb, nd = ctx.trace_beams(p, 1 << 20)
    gb, gd = ctx.download_beams()
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext, with_sv=False)
    rb, rd, rnb, rnd = O.trace_beams(s, p, 1 << 20)
    assert (nb, nd) == (rnb, rnd) and nb > 100
    assert bytes(gb) == bytes(rb)[: nb * 48], "beams differ"
    assert bytes(gd) == bytes(rd)[: nd * 80], "beam data differ"
    # truncation keeps the first max_beams segments in path order
    nb2, nd2 = ctx.trace_beams(p, nb // 3)
    gb2, gd2 = ctx.download_beams()
    assert nb2 == nb // 3 and bytes(gb2) == bytes(gb)[: nb2 * 48]
    # LBVH over the traced beams
    ctx.trace_beams(p, 1 << 20)
    ctx.lbvh_build()
    keys, perm, nodes = ctx.lbvh_download()
    rk, rp, rn = O.lbvh_build(gb, cloud.bounds[0], cloud.bounds[1])
    np.testing.assert_array_equal(keys, rk)
    np.testing.assert_array_equal(perm, rp)
    assert_same_tree(nodes, rn, nb)


@pytest.mark.parametrize("ref_compat,jitter", [(0, 0), (1, 0), (0, 1)])
def test_gather_bitexact(ctx, cfg, S, O, ref_compat, jitter):
    sc, cloud, _ = setup_c1(ctx, cfg, S)
    # camera transmittance: shadow volume along the camera forward (main.cpp:226-229)
    svp = cfg.shadow_volume(cloud, tuple(sc.cam.forward), 40)
    ctx.set_shadow_volume(svp)
    ctx.sv_build()
    p = cfg.beam_params(200, seed=11, radius=0.01, albedo=0.9, max_interactions=16)
    nb, nd = ctx.trace_beams(p, 1 << 20)
    ctx.lbvh_build()
    ext = S.GatherExt(ref_compat, jitter, 0.9, 0)
    ctx.clear_image()
    fr = cfg.frames(2)
    ctx.gather(fr, ext)
    gpu = ctx.download_image(sc.height, sc.width)
    st = ctx.gather_stats()
    gb, gd = ctx.download_beams()
    keys, perm, nodes = ctx.lbvh_download()
    sorted_beams = (S.PhotonBeam * nb)(*[gb[int(i)] for i in perm])
    for i in range(nb):
        sorted_beams[i].mortonCode = int(keys[i])
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    b = O.beam_scene(s, sorted_beams, gd, nodes, nb, ref_compat, jitter, 0.9)
    ref = O.gather_render(b, fr)
    assert_bitexact(gpu, ref, f"gather ref_compat={ref_compat} jitter={jitter}")
    # refCompat keeps PPB_PE.comp's missing guards (no t >= 0 test, a
    # degenerate beam direction is not rejected), so a pixel may be NaN
    # there exactly as in the reference; the corrected estimator is finite.
    if not ref_compat:
        assert np.isfinite(gpu).all()
    assert np.nanmax(gpu[..., 0]) > 0.0
    assert st.rays >= sc.width * sc.height * 2


def test_gather_bitexact_multi_workgroup_tree(ctx, cfg, S, O):
    """A tree of more than 16384 beams: its refit spans several workgroups and
    the nodes above their boundaries come from the second refit launch; the
    gather over it equals the oracle's bit for bit."""
    sc, cloud, _ = setup_c1(ctx, cfg, S)
    svp = cfg.shadow_volume(cloud, tuple(sc.cam.forward), 40)
    ctx.set_shadow_volume(svp)
    ctx.sv_build()
    p = cfg.beam_params(200, seed=21, radius=0.004, albedo=0.9, max_interactions=16)
    nb, nd = ctx.trace_beams(p, 1 << 20)
    assert nb > 2 * 16384
    ctx.lbvh_build()
    ctx.clear_image()
    fr = cfg.frames(1)
    ext = S.GatherExt(0, 0, 0.9, 0)
    ctx.gather(fr, ext)
    gpu = ctx.download_image(sc.height, sc.width)
    gb, gd = ctx.download_beams()
    keys, perm, nodes = ctx.lbvh_download()
    sorted_beams = (S.PhotonBeam * nb)(*[gb[int(i)] for i in perm])
    for i in range(nb):
        sorted_beams[i].mortonCode = int(keys[i])
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    b = O.beam_scene(s, sorted_beams, gd, nodes, nb, 0, 0, 0.9)
    ref = O.gather_render(b, fr)
    assert_bitexact(gpu, ref, f"gather over {nb} beams")
    assert np.isfinite(gpu).all() and gpu[..., 0].max() > 0.0


@pytest.mark.parametrize("keep", [1, 2, 3])
def test_gather_bitexact_tiny_trees(ctx, cfg, S, O, keep):
    """One beam (no internal node: the root box is tested directly), two and
    three beams: the gather equals the oracle's bit for bit."""
    sc, cloud, _ = setup_c1(ctx, cfg, S)
    svp = cfg.shadow_volume(cloud, tuple(sc.cam.forward), 40)
    ctx.set_shadow_volume(svp)
    ctx.sv_build()
    p = cfg.beam_params(40, seed=11, radius=0.05, albedo=0.9, max_interactions=16)
    ctx.trace_beams(p, 1 << 20)
    gb, gd = ctx.download_beams()
    # beams that light some pixel on their own (tried longest first), so the
    # comparison is not one of empty images
    fr = cfg.frames(2)
    lengths = [sum((gb[i].endPos[k] - gb[i].startPos[k]) ** 2 for k in range(3)) for i in range(len(gb))]
    pick = []
    for i in sorted(range(len(gb)), key=lambda i: -lengths[i])[:64]:
        ctx.upload_beams((S.PhotonBeam * 1)(gb[i]), gd)
        ctx.lbvh_build()
        ctx.clear_image()
        ctx.gather(fr[:1], S.GatherExt(0, 0, 0.9, 0))
        if ctx.download_image(sc.height, sc.width)[..., 0].max() > 0.0:
            pick.append(i)
            if len(pick) == keep:
                break
    assert len(pick) == keep
    ctx.upload_beams((S.PhotonBeam * keep)(*[gb[i] for i in pick]), gd)
    ctx.lbvh_build()
    ctx.clear_image()
    ctx.gather(fr, S.GatherExt(0, 0, 0.9, 0))
    gpu = ctx.download_image(sc.height, sc.width)
    sb, sd = ctx.download_beams()
    keys, perm, nodes = ctx.lbvh_download()
    sorted_beams = (S.PhotonBeam * keep)(*[sb[int(i)] for i in perm])
    for i in range(keep):
        sorted_beams[i].mortonCode = int(keys[i])
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    b = O.beam_scene(s, sorted_beams, sd, nodes, keep, 0, 0, 0.9)
    ref = O.gather_render(b, fr)
    assert_bitexact(gpu, ref, f"gather over {keep} beam(s)")
    assert gpu[..., 0].max() > 0.0


def test_gather_bvh_equals_brute_force(ctx, cfg, S, O):
    """The BVH sum equals the loop over every beam (up to reassociation)."""
    sc, cloud, _ = setup_c1(ctx, cfg, S)
    svp = cfg.shadow_volume(cloud, tuple(sc.cam.forward), 40)
    ctx.set_shadow_volume(svp)
    ctx.sv_build()
    p = cfg.beam_params(150, seed=3, radius=0.02, albedo=0.9, max_interactions=16)
    nb, nd = ctx.trace_beams(p, 1 << 20)
    ctx.lbvh_build()
    ctx.clear_image()
    fr = cfg.frames(1)
    ctx.gather(fr, S.GatherExt(0, 0, 0.9, 0))
    gpu = ctx.download_image(sc.height, sc.width)
    gb, gd = ctx.download_beams()
    keys, perm, nodes = ctx.lbvh_download()
    sorted_beams = (S.PhotonBeam * nb)(*[gb[int(i)] for i in perm])
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    b = O.beam_scene(s, sorted_beams, gd, nodes, nb, 0, 0, 0.9)
    for (x, y) in [(24, 16), (10, 20), (30, 8), (20, 25)]:
        brute = O.gather_pixel(b, x, y, fr[0].seed, brute=True)
        np.testing.assert_allclose(gpu[y, x], brute, rtol=2e-4, atol=1e-6)


def test_lbvh_full_size_bitexact(ctx, cfg, S, O):
    """The photon-beam workload at full size: 2^22 beams traced on the c1
    volume (seed 38, radius 0.002 x box side); sorted keys, permutation and all
    2^23 - 1 nodes equal the oracle's build of the same beams."""
    sc = cfg.c1(width=64, height=64)
    cloud, _ = sc.setup(ctx, build_sv=False)
    side = max(cloud.bounds[1][k] - cloud.bounds[0][k] for k in range(3))
    p = cfg.beam_params(1 << 19, seed=38, radius=0.002 * side, albedo=0.9, light_dir=sc.light_dir, max_interactions=64)
    nb, nd = ctx.trace_beams(p, 1 << 22)
    assert nb == 1 << 22
    ctx.lbvh_build()
    keys, perm, nodes = ctx.lbvh_download()
    beams, _ = ctx.download_beams()
    rk, rp, rn = O.lbvh_build(beams, cloud.bounds[0], cloud.bounds[1])
    np.testing.assert_array_equal(keys, rk, err_msg="sorted keys")
    np.testing.assert_array_equal(perm, rp, err_msg="permutation")
    assert_same_tree(nodes, rn, nb)
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)


def test_ppb_allocate_and_free(ctx, cfg, S, O, pkg):
    """RenderTechniquePPB::AllocateResources / FreeResources
    (RenderTechniquePPB.cpp:282-355): with every buffer reserved up front the
    trace and the build give the same beams and tree as with buffers grown on
    demand; after FreeResources the build has no beams (CR_ERR_STATE) until
    the next trace."""
    sc, cloud, svp = setup_c1(ctx, cfg, S)
    p = cfg.beam_params(100, seed=38, radius=0.004, albedo=0.9, max_interactions=16)
    tech = pkg.RenderTechniquePPB(ctx, 0.004)
    tech.FreeResources()
    nb, nd = ctx.trace_beams(p, 1 << 20)
    ctx.lbvh_build()
    k1, p1, n1 = ctx.lbvh_download()
    tech.FreeResources()
    with pytest.raises(pkg.CrError) as e:
        ctx.lbvh_build()
    assert e.value.status == S.CR_ERR_STATE
    tech.AllocateResources(p, 1 << 20)
    assert ctx.lib.cr_ppb_allocate(ctx.h, 1) == S.CR_ERR_CAPACITY
    nb2, nd2 = ctx.trace_beams(p, 1 << 20)
    ctx.lbvh_build()
    k2, p2, n2 = ctx.lbvh_download()
    assert (nb2, nd2) == (nb, nd)
    np.testing.assert_array_equal(k1, k2)
    np.testing.assert_array_equal(p1, p2)
    assert bytes(n1) == bytes(n2)
