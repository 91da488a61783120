"""Regression fixtures of the CPU oracle: SHA-256 digests of its outputs on
small fixed inputs (fBm grids, brick majorants, shadow volume, path-traced
images in every tracking/NEE mode, a photon-beam trace, the LBVH of seeded
beams, a beam gather, the elementary functions).

The reference ships no fixtures; the oracle is pinned against closed forms,
known answers and brute force in test_oracle.py.  These digests freeze what
was pinned, so any later change to the oracle's arithmetic (and hence to what
the GPU kernels are compared against) is visible as a fixture change.

    python tests/golden/make_golden.py      # rewrites oracle_digests.json
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import cr_import  # noqa: E402


def digest(a):
    if isinstance(a, C.Array):
        b = bytes(a)
    else:
        b = np.ascontiguousarray(a).tobytes()
    return hashlib.sha256(b).hexdigest()


def cloud_props(pkg, grid, voxel_size, base=1.0, dens=64.0):
    """SetCloudProperties (main.cpp:144-163) on the host for a CPU-made grid."""
    c = pkg.structs.CloudProperties()
    pkg.load().cr_cloud_default(C.byref(c))
    c.baseScaling = base
    c.densityScaling = dens
    nz, ny, nx = grid.shape
    sx = np.float32(voxel_size * nx * base)
    sy = np.float32(voxel_size * ny * base)
    sz = np.float32(voxel_size * nz * base)
    c.maxExtinction = max(float(grid.max()), 0.01)
    c.voxelCount[:] = (nx, ny, nz, 0)
    c.bounds[0][:] = (-sx / 2, -sy / 2, 0.0, 0.0)
    c.bounds[1][:] = (sx / 2, sy / 2, sz, 0.0)
    return c


def compute():
    pkg = cr_import.package()
    cfg = cr_import.configs()
    O = cr_import.oracle()
    S = pkg.structs
    out = {}

    # elementary functions
    x = np.linspace(-30.0, 30.0, 4001, dtype=np.float32)
    pos = np.abs(x) + np.float32(1e-3)
    out["math/sqrt"] = digest(O.math_array(0, pos))
    out["math/log"] = digest(O.math_array(1, pos))
    out["math/exp"] = digest(O.math_array(2, x))
    out["math/sin"] = digest(O.math_array(3, x))
    out["math/cos"] = digest(O.math_array(4, x))

    # grids
    p_cu = cfg.fbm((32, 28, 24), 1.0 / 32, seed=5, octaves=6)
    g_cu = O.fbm_grid(p_cu)
    out["fbm/cumulus_32x28x24_seed5"] = digest(g_cu)
    p_an = cfg.fbm((24, 16, 24), 1.0 / 24, seed=7, octaves=8, shape=S.CR_SHAPE_ANVIL)
    out["fbm/anvil_24x16x24_seed7"] = digest(O.fbm_grid(p_an))
    for b in (4, 16):
        out[f"majorants/cumulus_b{b}"] = digest(O.brick_majorants(g_cu, b)[0])

    # path tracer, 24x16, 2 frames, in each mode
    cloud = cloud_props(pkg, g_cu, 1.0 / 32)
    cam = cfg.camera(24, 16, position=(0.0, 0.0, -0.2))
    svp = cfg.shadow_volume(cloud, (1.0, -1.0, 0.0), 24)
    frames = cfg.frames(2)
    modes = {
        "reference": (cfg.render_ext(), cfg.parameters(0.85)),
        "sv_global": (cfg.render_ext(albedo=0.9, bounce_cap=16, sky=cfg.SKY), cfg.parameters(0.85)),
        "ratio_dda": (cfg.render_ext(albedo=0.9, bounce_cap=16, sky=cfg.SKY, nee=S.CR_NEE_RATIO_TRACKING,
                                     tracking=S.CR_TRACK_BRICK_DDA), cfg.parameters(0.85)),
        "ratio_dda_b8_iso": (cfg.render_ext(albedo=0.9, bounce_cap=16, sky=cfg.SKY, nee=S.CR_NEE_RATIO_TRACKING,
                                            tracking=S.CR_TRACK_BRICK_DDA, majorant_brick=8), cfg.parameters(0.0)),
    }
    for name, (ext, params) in modes.items():
        s = O.scene(g_cu, cloud, cam, params, svp, ext)
        sv = O.sv_build(s)
        if name == "reference":
            out["sv/cumulus_24"] = digest(sv)
        O.with_sv(s, sv)
        img, cnt = O.pt_render(s, frames)
        out[f"pt/{name}"] = digest(img)
        out[f"pt/{name}/counters"] = [int(cnt.paths), int(cnt.vertices), int(cnt.lookups), int(cnt.svLookups),
                                      int(cnt.brickVisits)]
        if name == "reference":
            f16 = O.scene(g_cu.astype(np.float16), cloud, cam, params, svp, ext, sv=sv)
            out["pt/reference_f16"] = digest(O.pt_render(f16, frames)[0])

    # photon beams: trace, LBVH, gather
    s = O.scene(g_cu, cloud, cam, cfg.parameters(0.85), svp, modes["sv_global"][0])
    O.with_sv(s, O.sv_build(s))
    bp = cfg.beam_params(64, max_interactions=16)
    beams, data, nb, nd = O.trace_beams(s, bp, 1 << 16)
    out["ppb/trace/beams"] = digest(bytes(beams)[: nb * 48])
    out["ppb/trace/data"] = digest(bytes(data)[: nd * 80])
    keys, perm, nodes = O.lbvh_build(beams, cloud.bounds[0], cloud.bounds[1]) if nb else (None, None, None)
    if nb:
        out["ppb/lbvh/keys"] = digest(keys)
        out["ppb/lbvh/perm"] = digest(perm)
        out["ppb/lbvh/nodes"] = digest(bytes(nodes)[: (2 * nb - 1) * 64])
        sorted_beams = (S.PhotonBeam * nb)(*[beams[int(i)] for i in perm])
        for i in range(nb):
            sorted_beams[i].mortonCode = int(keys[i])
        b = O.beam_scene(s, sorted_beams, data, nodes, nb, 0, 0, 0.9)
        out["ppb/gather"] = digest(O.gather_render(b, cfg.frames(1)))
    rng = np.random.default_rng(3)
    rb = (S.PhotonBeam * 777)()
    a = np.frombuffer(rb, dtype=np.uint8).reshape(777, 48)
    f = a.view(np.float32)
    st = rng.uniform(-0.45, 0.45, (777, 3)).astype(np.float32)
    f[:, 0:3] = st
    f[:, 4:7] = st + rng.uniform(-0.01, 0.01, (777, 3)).astype(np.float32)
    f[:, 7] = 0.002
    k2, p2, n2 = O.lbvh_build(rb, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    out["lbvh/random777/keys"] = digest(k2)
    out["lbvh/random777/perm"] = digest(p2)
    out["lbvh/random777/nodes"] = digest(bytes(n2)[: (2 * 777 - 1) * 64])
    return out


if __name__ == "__main__":
    d = compute()
    with open(os.path.join(HERE, "oracle_digests.json"), "w") as fh:
        json.dump(d, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(f"wrote {len(d)} digests")
