"""Helpers shared by the GPU parity tests."""
import numpy as np


def first_mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    av = a.view(np.uint32) if a.dtype == np.float32 else a
    bv = b.view(np.uint32) if b.dtype == np.float32 else b
    ne = av != bv
    if a.dtype == np.float32 and b.dtype == np.float32:
        ne &= ~(np.isnan(a) & np.isnan(b))  # IEEE-754 leaves the NaN payload and sign open
    bad = np.argwhere(ne)
    if len(bad) == 0:
        return None
    i = tuple(bad[0])
    return f"{len(bad)} of {a.size} differ; first at {i}: gpu={a[i]!r} oracle={b[i]!r}"


def assert_bitexact(gpu, ref, what):
    m = first_mismatch(gpu, ref)
    assert m is None, f"{what}: {m}"


def oracle_scene_from_ctx(O, ctx, cam, params, svp, ext, with_sv=True):
    """Build the oracle scene from exactly what the device holds (grid,
    cloud properties, majorants, shadow volume)."""
    grid = ctx.download_grid()
    cloud = ctx.cloud()
    maj = ctx.majorants()
    sv = ctx.sv_download() if with_sv else None
    return O.scene(grid, cloud, cam, params, svp, ext, sv=sv, majorants=maj), grid, sv
