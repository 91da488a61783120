"""cr_pt_render_static, the allocation-free path-tracer entry point, against
cr_pt_render through the C-ABI: two calls in a row on one GPU (the second
with a smaller image, fewer frames and another kernel instantiation, so the
shared scratch array holds the first call's contents), each bit-exact against
cr_pt_render over the same starting image; then the limits."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import assert_bitexact

pytestmark = pytest.mark.gpu


class DeviceImage:
    """A device buffer from the HIP runtime the library itself links (the
    caller's allocation, made before the allocation-free call)."""

    def __init__(self, host):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def download(self, shape):
        out = np.empty(shape, np.float32)
        assert out.nbytes == self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def test_pt_render_static_equals_cr_pt_render(ctx, cfg, S, pkg):
    lib = pkg.load()
    c2 = cfg.c2(width=64, height=40, size=(48, 24, 48))           # FP16 bricks, brick DDA, ratio tracking
    c1 = cfg.c1(width=24, height=16, size=24, sv_axis=24)         # FP32 dense, global majorant, shadow-volume NEE
    acc = None
    for sc, nfr in ((c2, 6), (c1, 2)):
        w, h = sc.width, sc.height
        sc.setup(ctx)
        fr = cfg.frames(nfr, first=3)
        init = np.random.default_rng(w).random((h, w, 4), dtype=np.float32)
        ctx.upload_image(init)
        ctx.pt_render(fr)
        ref = ctx.download_image(h, w)
        if acc is not None:
            acc.free()
        acc = DeviceImage(init)
        arr = (S.PushConstants * nfr)(*fr)
        st = lib.cr_pt_render_static(ctx.h, arr, nfr, acc.ptr)
        assert st == 0, lib.cr_last_error(ctx.h)
        ctx.synchronize()
        assert_bitexact(acc.download((h, w, 4)), ref, f"{sc.name} {w}x{h}, {nfr} frames: static entry point vs cr_pt_render")
    n = S.CR_PT_STATIC_MAX_FRAMES + 1
    big = (S.PushConstants * n)(*cfg.frames(n))
    assert lib.cr_pt_render_static(ctx.h, big, n, acc.ptr) == S.CR_ERR_CAPACITY
    assert b"CR_PT_STATIC_MAX_FRAMES" in lib.cr_last_error(ctx.h)
    assert lib.cr_pt_render_static(ctx.h, big, 1, None) == S.CR_ERR_INVALID_ARGUMENT
    acc.free()
