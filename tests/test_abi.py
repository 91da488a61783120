"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, its struct layouts match the reference's, and its host-side setters
restate UniformBuffers.h.  No kernel is launched."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("abi.h",):
        text = open(os.path.join(ROOT, "include", "cr", h)).read()
        names += re.findall(r"CR_API\s+[\w\s\*]+?\b(cr_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.load()
    names = declared_symbols()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    from crmi355x import _lib
    assert sorted(_lib.SIGNATURES) == names


def test_struct_layouts(S):
    assert C.sizeof(S.PushConstants) == 24 and C.alignment(S.PushConstants) == 8
    assert S.CloudProperties.maxExtinction.offset == 48
    assert S.Parameters.isotropic.offset == 24
    assert S.ShadowVolumeProperties.basisChange.offset == 80
    assert S.ShadowVolumeProperties.voxelSize.offset == 148
    assert S.PhotonBeam.mortonCode.offset == 12 and S.PhotonBeam.dataIdx.offset == 32
    assert S.TreeNode.bounds.offset == 16 and S.TreeNode.processed.offset == 48
    assert C.sizeof(S.PhotonBeamData) == 80


def test_status_strings(pkg):
    lib = pkg.load()
    assert lib.cr_status_string(0) == b"CR_OK"
    assert lib.cr_status_string(6) == b"CR_ERR_STATE"
    assert lib.cr_version().startswith(b"cr-mi355x")


def test_create_without_gpu_reports_device_error(pkg):
    lib = pkg.load()
    h = C.c_void_p()
    s = lib.cr_create(0, C.byref(h))
    if s == 0:
        lib.cr_destroy(h)
        pytest.skip("a GPU is present")
    assert s == pkg.structs.CR_ERR_DEVICE
    assert lib.cr_create(0, None) == pkg.structs.CR_ERR_INVALID_ARGUMENT
    assert lib.cr_destroy(None) == pkg.structs.CR_ERR_INVALID_ARGUMENT


def test_parameters_set_phase_g(pkg, S):
    lib = pkg.load()
    p = S.Parameters()
    lib.cr_parameters_default(C.byref(p))
    assert p.maxRayBounces == 5 and p.lightIntensity == 5.0 and p.isotropic == 1  # UniformBuffers.h:83-90
    lib.cr_parameters_set_phase_g(C.byref(p), 1.5)   # ignored (UniformBuffers.h:95)
    assert p.isotropic == 1
    lib.cr_parameters_set_phase_g(C.byref(p), 0.99)  # clamped to 0.95
    assert p.phaseG == np.float32(0.95) and p.isotropic == 0
    assert p.phaseOnePlusG2 == np.float32(1) + np.float32(0.95) * np.float32(0.95)
    assert p.phaseOneOver2G == np.float32(0.5) / np.float32(0.95)
    lib.cr_parameters_set_phase_g(C.byref(p), -0.00001)
    assert p.isotropic == 1


def test_camera_setters(pkg, S):
    lib = pkg.load()
    c = S.CameraProperties()
    lib.cr_camera_default(C.byref(c))
    assert (c.halfWidth, c.halfHeight, c.nearPlane) == (400, 300, 50.0)
    assert abs(c.pixelSizeY - 2 * math.tan(math.radians(60)) * 50 / 600) < 1e-6
    lib.cr_camera_set_resolution(C.byref(c), 1920, 1080)
    fov = C.c_float(200.0)
    lib.cr_camera_set_fov(C.byref(c), C.byref(fov))
    assert fov.value == 150.0  # clamped (UniformBuffers.h:54)
    assert abs(c.pixelSizeY - 2 * math.tan(math.radians(75)) * 50 / 1080) < 1e-5
    # horizontal pixel size keeps a square frustum (the reference does the same)
    assert abs(c.pixelSizeX - c.pixelSizeY * 540 / 960) < 1e-6 * c.pixelSizeX
    rot = (C.c_float * 2)(100.0, 30.0)
    lib.cr_camera_set_rotation(C.byref(c), rot)
    assert rot[0] == 80.0
    f, r, u = (np.array(v, np.float64) for v in (c.forward, c.right, c.up))
    assert abs(np.linalg.norm(f) - 1) < 1e-6 and abs(f @ r) < 1e-6 and abs(f @ u) < 1e-6 and abs(r @ u) < 1e-6


def test_shadow_volume_set_light(pkg, S):
    lib = pkg.load()
    s = S.ShadowVolumeProperties()
    lib.cr_shadow_volume_default(C.byref(s))
    assert s.voxelAxisCount == 500 and list(s.lightDirection) == [1, -1, 0, 0]
    lo, hi = (C.c_float * 3)(-0.5, -0.25, 0.0), (C.c_float * 3)(0.5, 0.25, 1.0)
    lib.cr_shadow_volume_set_light(C.byref(s), (C.c_float * 3)(1, -1, 0), lo, hi)
    L = np.array(s.lightDirection[:3], np.float64)
    R = np.array(s.right[:3], np.float64)
    U = np.array(s.up[:3], np.float64)
    assert abs(np.linalg.norm(L) - 1) < 1e-6 and abs(L @ R) < 1e-6 and abs(L @ U) < 1e-6
    diag = math.sqrt(1 + 0.25 + 1)
    assert abs(s.voxelSize - diag / 499) < 1e-6  # 2R/(N-1), R = half diagonal
    center = (np.array(s.bounds[0][:3]) + np.array(s.bounds[1][:3])) / 2
    np.testing.assert_allclose(center, [0.0, 0.0, 0.5], atol=1e-6)
    # basisChange maps the light-space axes to the unit vectors
    M = np.array(s.basisChange, np.float64).reshape(4, 4).T  # column-major -> row-major
    np.testing.assert_allclose(M[:3, :3] @ R, [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(M[:3, :3] @ U, [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(M[:3, :3] @ L, [0, 0, 1], atol=1e-6)
    assert s.bounds[0][3] == 1.0 and s.bounds[1][3] == 1.0


def test_msvc_seed_sequence(pkg):
    out = (C.c_int32 * 8)()
    pkg.load().cr_msvc_rand_sequence(0, 8, out)
    assert list(out) == [38, 7719, 21238, 2437, 8855, 11797, 8365, 32285]


def test_ppb_update_radius_matches_oracle(pkg, O, S):
    """The library's UpdateRadius (a host function) equals the oracle bit for
    bit over a progressive run, for the reference's constants and others."""
    lib = pkg.load()
    for B, alpha, r0 in ((S.CR_PPB_BEAMS_PER_PASS, S.CR_PPB_ALPHA, 0.002), (1, 0.5, 0.01), (300, 0.9, 1.0)):
        a = b = 0.0
        for f in range(0, 400):
            a = lib.cr_ppb_update_radius(f, B, alpha, r0, a)
            b = O.lib.or_ppb_update_radius(f, B, alpha, r0, b)
            assert np.float32(a).tobytes() == np.float32(b).tobytes(), (B, f)
    assert np.array_equal(O.ppb_radii(5, 56, 0.8, 0.002)[:2], np.float32([0.002, 0.002]))


def test_configs_are_the_named_workloads(cfg, S):
    c1 = cfg.c1()
    assert (c1.width, c1.height, c1.spp) == (256, 256, 64)
    assert tuple(c1.fbm.size) == (128, 128, 128) and c1.fbm.octaves == 6
    assert c1.ext.albedo == np.float32(0.9) and c1.ext.bounceCap == 16
    assert c1.params.phaseG == np.float32(0.85) and c1.density_scaling == 64.0
    c2 = cfg.c2()
    assert (c2.width, c2.height, c2.spp) == (1920, 1080, 256)
    assert tuple(c2.fbm.size) == (1024, 512, 1024) and c2.fbm.octaves == 8
    assert c2.storage == S.CR_DTYPE_F16 and c2.ext.neeMode == S.CR_NEE_RATIO_TRACKING
    assert tuple(c2.ext.svDims) == (256, 128, 256) and c2.ext.bounceCap == 256
    fr = cfg.frames(3)
    assert [f.seed for f in fr] == [38, 7719, 21238] and [f.frameCount for f in fr] == [0, 1, 2]
