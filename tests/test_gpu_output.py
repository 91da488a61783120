"""Progressive accumulation, image output and the .xyz density format,
through the C-ABI.

Progressive accumulation restates the reference's render loop: per frame
seed = std::rand() after std::srand(0) (main.cpp:722, :949), frameCount + 1
after every frame (main.cpp:728), and frameCount = 1 without clearing the
image after any change (UpdateShadowVolume, main.cpp:262, reached from every
parameter change at main.cpp:560-590).  Images are checked bit for bit
against the oracle over the same starting image.
"""
import struct

import numpy as np
import pytest

from gpu_util import assert_bitexact, oracle_scene_from_ctx

pytestmark = pytest.mark.gpu


def scene(cfg):
    sc = cfg.c1(width=40, height=24, size=32, sv_axis=32)   # the reference technique: global majorant, SV NEE
    return sc


def test_progressive_reproduces_the_reference_loop(ctx, cfg, S, O):
    sc = scene(cfg)
    cloud, svp = sc.setup(ctx)
    H, W = sc.height, sc.width
    init = np.random.default_rng(4).random((H, W, 4), dtype=np.float32)
    ctx.upload_image(init)
    ctx.pt_progressive_reset()                       # srand(0), frameCount from 1, no clear
    a = ctx.pt_progressive(3)
    b = ctx.pt_progressive(2)
    assert (a.frameCount, b.frameCount, b.framesRendered, b.restarts) == (4, 6, 5, 0)
    img = ctx.download_image(H, W)
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    ref, _ = O.pt_render(s, cfg.frames(5, ref_frame_count=True), accum=init)
    assert_bitexact(img, ref, "five frames of the reference loop over an uncleared image")

    # a parameter change restarts frameCount at 1 without clearing; the seeds go on with std::rand
    sc.params = cfg.parameters(0.6)
    ctx.set_parameters(sc.params)
    c = ctx.pt_progressive(2)
    assert (c.restarts, c.frameCount, c.framesRendered) == (1, 3, 2)
    seeds = cfg.msvc_seeds(7)[5:]
    frames = [S.PushConstants(0.0, seeds[i], 1 + i, 0.0, 0) for i in range(2)]
    s, _, _ = oracle_scene_from_ctx(O, ctx, sc.cam, sc.params, svp, sc.ext)
    ref2, _ = O.pt_render(s, frames, accum=img)
    img2 = ctx.download_image(H, W)
    assert_bitexact(img2, ref2, "two frames after the restart")
    assert not np.array_equal(img2, ref)

    # nothing changed: no restart
    d = ctx.pt_progressive(1)
    assert (d.restarts, d.frameCount, d.framesRendered) == (1, 4, 3)
    # a shadow-volume rebuild counts as a change (UpdateShadowVolume itself restarts the count)
    ctx.sv_build()
    e = ctx.pt_progressive(0)
    assert (e.restarts, e.frameCount, e.framesRendered) == (2, 1, 0)

    # exact-mean mode: restart at frameCount 0 on a cleared image
    ctx.pt_progressive_reset(0, 0, True)
    ctx.pt_progressive(4)
    ctx.clear_image()
    ctx.pt_render(cfg.frames(4))
    exact = ctx.download_image(H, W)
    ctx.pt_progressive_reset(0, 0, True)
    ctx.upload_image(init)
    ctx.pt_progressive(4)
    assert_bitexact(ctx.download_image(H, W), exact, "exact-mean restart")


def test_write_pfm(ctx, cfg, tmp_path):
    """Portable float map: 'PF', width and height, scale -1 (little-endian),
    then RGB float32 rows from the bottom of the image up."""
    sc = scene(cfg)
    sc.setup(ctx)
    ctx.pt_render(cfg.frames(2))
    img = ctx.download_image(sc.height, sc.width)
    path = tmp_path / "out.pfm"
    ctx.write_pfm(path)
    raw = path.read_bytes()
    header = b"PF\n%d %d\n-1.0\n" % (sc.width, sc.height)
    assert raw.startswith(header)
    data = np.frombuffer(raw[len(header):], "<f4")
    assert data.size == sc.width * sc.height * 3
    np.testing.assert_array_equal(data.reshape(sc.height, sc.width, 3), img[::-1, :, :3])


def test_xyz_file_built_byte_by_byte(ctx, S, tmp_path):
    """A .xyz file built byte by byte as Grid3D<float>::Load reads it
    (Grid3D.h:44-90): three little-endian u32 counts, three f64 voxel sizes,
    then f32 values with z fastest and x slowest, stored at
    x + sx*y + sx*sy*z.  SetCloudProperties (main.cpp:144-163) sizes the box
    as voxel size * count * baseScaling, centred in x and y, from 0 in z.
    cr_grid_save_xyz writes the same bytes back."""
    sx, sy, sz = 5, 3, 4
    vs = (0.01, 0.02, 0.03)
    vals = (np.arange(sx * sy * sz, dtype=np.float32) * np.float32(0.25) + np.float32(0.125))
    blob = struct.pack("<3I3d", sx, sy, sz, *vs) + vals.astype("<f4").tobytes()
    assert len(blob) == 36 + 4 * sx * sy * sz
    path = tmp_path / "byte.xyz"
    path.write_bytes(blob)
    ctx.set_grid_layout(S.CR_GRID_DENSE)
    ctx.load_xyz(str(path))
    g = ctx.download_grid()
    assert g.shape == (sz, sy, sx)
    k = 0
    for x in range(sx):
        for y in range(sy):
            for z in range(sz):
                assert g[z, y, x] == vals[k], (x, y, z)
                k += 1
    c = ctx.cloud()
    assert tuple(c.voxelCount[:3]) == (sx, sy, sz)
    size = [vs[i] * n * c.baseScaling for i, n in enumerate((sx, sy, sz))]
    np.testing.assert_allclose(list(c.bounds[0][:3]), [-size[0] / 2, -size[1] / 2, 0.0], rtol=1e-6)
    np.testing.assert_allclose(list(c.bounds[1][:3]), [size[0] / 2, size[1] / 2, size[2]], rtol=1e-6)
    assert c.maxExtinction == vals.max()
    out = tmp_path / "saved.xyz"
    ctx.save_xyz(str(out))
    assert out.read_bytes() == blob
    # a truncated file is an error, not a partial grid
    short = tmp_path / "short.xyz"
    short.write_bytes(blob[:-4])
    with pytest.raises(Exception):
        ctx.load_xyz(str(short))


def test_headless_driver_writes_the_same_image(ctx, cfg, S, pkg, tmp_path):
    """The plain C++ driver (driver/headless.cpp over include/cr/techniques.hpp,
    no HIP headers) renders the fBm cumulus through the progressive loop and
    writes a PFM; the same set-up through the Python binding gives the same
    image bit for bit."""
    import ctypes as C
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pkg.__file__), "headless")
    assert os.path.exists(exe), "build the driver (make -C cloudrendering-vulkan_amd)"
    W, H, n = 48, 32, 3
    out = tmp_path / "driver.pfm"
    r = subprocess.run([exe, str(n), str(W), str(H), "", str(out)], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    raw = out.read_bytes()
    header = b"PF\n%d %d\n-1.0\n" % (W, H)
    assert raw.startswith(header)
    img = np.frombuffer(raw[len(header):], "<f4").reshape(H, W, 3)[::-1]
    # the driver's set-up: reference render settings (cr_render_ext_default), default light intensity
    sc = cfg.c1(width=W, height=H)
    sc.ext = cfg.render_ext()
    p = S.Parameters()
    lib = pkg.load()
    lib.cr_parameters_default(C.byref(p))
    lib.cr_parameters_set_phase_g(C.byref(p), 0.85)
    sc.params = p
    sc.setup(ctx)
    ctx.clear_image()
    ctx.pt_render(cfg.frames(n))
    assert_bitexact(img, ctx.download_image(H, W)[..., :3], "headless driver image")
