"""MI355X (gfx950) volumetric hot path of jeanfilho/CloudRendering-Vulkan.

Python mirror of the reference's technique API (RenderTechnique.h:15-66)
over the C-ABI of libcrmi355x.so.  The host side of the reference is C++;
this module is the binding the tests, the bench and the driver use.  All
computation runs in the gfx950 kernels; there is no CPU fallback.

    ctx = Context()
    ctx.generate_fbm(fbm)                          # or ctx.upload_grid(...)
    sv = RenderTechniqueSV(ctx); sv.QueueUpdateShadowVolume(svp); sv.RecordDrawCommands()
    pt = RenderTechniquePT(ctx); pt.QueueUpdateCameraProperties(cam)
    pt.QueueUpdateParameters(params); pt.UpdateDescriptorSets()
    pt.RecordDrawCommands(frames)                  # n frames, one persistent launch
    img = ctx.download_image()
"""
import ctypes as C

import numpy as np

from . import structs
from ._lib import LibraryNotBuilt, load  # noqa: F401
from .structs import *  # noqa: F401,F403

__all__ = ["Context", "CrError", "RenderTechniquePT", "RenderTechniqueSV", "RenderTechniquePPB", "RenderTechniquePPM",
           "structs", "load", "LibraryNotBuilt"]


class CrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} ({status})")
        self.status = status


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """Owns one cr_context (device buffers resident in HBM, one HIP stream)."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        s = self.lib.cr_create(device, C.byref(h))
        if s != 0:
            raise CrError(s, f"cr_create(device={device}) failed: {self.lib.cr_status_string(s).decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.cr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, s, what):
        if s != 0:
            msg = self.lib.cr_last_error(self.h).decode()
            raise CrError(s, f"{what}: {self.lib.cr_status_string(s).decode()}: {msg}")

    @property
    def stream(self):
        return self.lib.cr_stream(self.h)

    def synchronize(self):
        self.check(self.lib.cr_synchronize(self.h), "cr_synchronize")

    # ---------------------------------------------------------------- grid
    def set_grid_layout(self, layout):
        self.check(self.lib.cr_set_grid_layout(self.h, layout), "cr_set_grid_layout")

    def upload_grid(self, data, voxel_size, layout=structs.CR_LAYOUT_X_FASTEST, storage=structs.CR_DTYPE_F32):
        """data: 3-D array indexed [z, y, x] (x fastest) or [x, y, z] with layout Z_FASTEST."""
        a = np.ascontiguousarray(data)
        if a.dtype == np.float32:
            src = structs.CR_DTYPE_F32
        elif a.dtype == np.float16:
            src = structs.CR_DTYPE_F16
        else:
            raise TypeError("grid must be float32 or float16")
        if layout == structs.CR_LAYOUT_X_FASTEST:
            nz, ny, nx = a.shape
        else:
            nx, ny, nz = a.shape
        vs = voxel_size if np.ndim(voxel_size) else (voxel_size,) * 3
        h = structs.GridHeader(nx, ny, nz, 0, float(vs[0]), float(vs[1]), float(vs[2]))
        self.check(self.lib.cr_grid_upload(self.h, C.byref(h), _ptr(a), src, layout, storage), "cr_grid_upload")

    def load_xyz(self, path, storage=structs.CR_DTYPE_F32):
        self.check(self.lib.cr_grid_load_xyz(self.h, path.encode(), storage), "cr_grid_load_xyz")

    def save_xyz(self, path):
        self.check(self.lib.cr_grid_save_xyz(self.h, path.encode()), "cr_grid_save_xyz")

    def generate_fbm(self, params, storage=structs.CR_DTYPE_F32):
        self.check(self.lib.cr_grid_generate_fbm(self.h, C.byref(params), storage), "cr_grid_generate_fbm")

    def download_grid(self):
        c = self.cloud()
        nx, ny, nz = c.voxelCount[0], c.voxelCount[1], c.voxelCount[2]
        out = np.empty((nz, ny, nx), np.float32)
        self.check(self.lib.cr_grid_download(self.h, _ptr(out)), "cr_grid_download")
        return out

    def majorants(self):
        d = (C.c_uint32 * 3)()
        self.check(self.lib.cr_grid_majorants(self.h, d, None), "cr_grid_majorants")
        out = np.empty((d[2], d[1], d[0]), np.float32)
        self.check(self.lib.cr_grid_majorants(self.h, d, _ptr(out)), "cr_grid_majorants")
        return out

    # ------------------------------------------------------------ uniforms
    def skip_grid(self):
        d = (C.c_uint32 * 3)()
        self.check(self.lib.cr_grid_majorants(self.h, d, None), "cr_grid_majorants")
        out = np.empty((d[2], d[1], d[0]), np.float32)
        self.check(self.lib.cr_grid_skip(self.h, _ptr(out)), "cr_grid_skip")
        return out

    def set_camera(self, cam):
        self.check(self.lib.cr_set_camera(self.h, C.byref(cam)), "cr_set_camera")

    def set_cloud(self, cloud):
        self.check(self.lib.cr_set_cloud(self.h, C.byref(cloud)), "cr_set_cloud")

    def cloud(self):
        c = structs.CloudProperties()
        self.check(self.lib.cr_get_cloud(self.h, C.byref(c)), "cr_get_cloud")
        return c

    def set_parameters(self, p):
        self.check(self.lib.cr_set_parameters(self.h, C.byref(p)), "cr_set_parameters")

    def set_shadow_volume(self, s):
        self.check(self.lib.cr_set_shadow_volume(self.h, C.byref(s)), "cr_set_shadow_volume")

    def set_render_ext(self, e):
        self.check(self.lib.cr_set_render_ext(self.h, C.byref(e)), "cr_set_render_ext")

    # -------------------------------------------------------- shadow volume
    def sv_build(self):
        self.check(self.lib.cr_sv_build(self.h), "cr_sv_build")

    def sv_dims(self):
        d = (C.c_uint32 * 3)()
        self.check(self.lib.cr_sv_dims(self.h, d), "cr_sv_dims")
        return tuple(d)

    def sv_download(self):
        dx, dy, dz = self.sv_dims()
        out = np.empty((dz, dy, dx), np.float32)
        self.check(self.lib.cr_sv_download(self.h, _ptr(out)), "cr_sv_download")
        return out

    # --------------------------------------------------------- path tracer
    def clear_image(self):
        self.check(self.lib.cr_pt_clear(self.h), "cr_pt_clear")

    def pt_render(self, frames):
        arr = (structs.PushConstants * len(frames))(*frames)
        self.check(self.lib.cr_pt_render(self.h, arr, len(frames)), "cr_pt_render")

    def image_shape(self, cam):
        return 2 * cam.halfHeight, 2 * cam.halfWidth

    def download_image(self, H, W):
        out = np.empty((H, W, 4), np.float32)
        self.check(self.lib.cr_pt_download(self.h, _ptr(out)), "cr_pt_download")
        return out

    def upload_image(self, img):
        a = np.ascontiguousarray(img, dtype=np.float32)
        self.check(self.lib.cr_pt_upload(self.h, _ptr(a)), "cr_pt_upload")

    # ------------------------------------------------------ photon mapping
    def set_photon_map(self, pm):
        self.check(self.lib.cr_set_photon_map(self.h, C.byref(pm)), "cr_set_photon_map")

    def ppm_trace(self, seed, params):
        n = C.c_uint32()
        self.check(self.lib.cr_ppm_trace(self.h, seed, C.byref(params), C.byref(n)), "cr_ppm_trace")
        return n.value

    def ppm_estimate(self, frames, params):
        arr = (structs.PushConstants * len(frames))(*frames)
        self.check(self.lib.cr_ppm_estimate(self.h, arr, len(frames), C.byref(params)), "cr_ppm_estimate")

    def ppm_download(self):
        """(arrivals per cell, photons sorted by cell in emission order)."""
        nc, nd = C.c_uint32(), C.c_uint32()
        self.check(self.lib.cr_ppm_download(self.h, C.byref(nc), C.byref(nd), None, None), "cr_ppm_download")
        arrivals = np.empty(nc.value, np.uint32)
        photons = (structs.Photon * max(nd.value, 1))()
        self.check(self.lib.cr_ppm_download(self.h, C.byref(nc), C.byref(nd), _ptr(arrivals), C.addressof(photons)),
                   "cr_ppm_download")
        return arrivals, photons, nd.value

    def pt_progressive_reset(self, srand_seed=0, first_frame_count=1, clear_on_restart=False):
        self.check(self.lib.cr_pt_progressive_reset(self.h, srand_seed, first_frame_count, int(clear_on_restart)),
                   "cr_pt_progressive_reset")

    def pt_progressive(self, n_frames):
        """The reference render loop's next n_frames frames (restarting the
        running mean after any change); returns the progressive state."""
        st = structs.ProgressiveState()
        self.check(self.lib.cr_pt_progressive(self.h, n_frames, C.byref(st)), "cr_pt_progressive")
        return st

    def write_pfm(self, path):
        self.check(self.lib.cr_pt_write_pfm(self.h, str(path).encode()), "cr_pt_write_pfm")

    def pt_stats(self):
        s = structs.PtStats()
        self.check(self.lib.cr_pt_stats_get(self.h, C.byref(s)), "cr_pt_stats_get")
        return s

    def pt_stats_reset(self):
        self.check(self.lib.cr_pt_stats_reset(self.h), "cr_pt_stats_reset")

    # ----------------------------------------------------- learned renderer
    def pt_record(self, frames, H, W):
        arr = (structs.PushConstants * len(frames))(*frames)
        out = np.empty((len(frames), H, W, 8), np.float32)
        self.check(self.lib.cr_pt_record(self.h, arr, len(frames), _ptr(out)), "cr_pt_record")
        return out

    def pt_continue(self, records, spp):
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 8)
        out = np.empty((len(r), 4), np.float32)
        self.check(self.lib.cr_pt_continue(self.h, _ptr(r), len(r), spp, _ptr(out)), "cr_pt_continue")
        return out

    def learn_build_mips(self):
        self.check(self.lib.cr_learn_build_mips(self.h), "cr_learn_build_mips")

    def learn_collect(self, frames, spp, max_points=0):
        arr = (structs.PushConstants * len(frames))(*frames)
        n = C.c_uint32()
        self.check(self.lib.cr_learn_collect(self.h, arr, len(frames), spp, max_points, C.byref(n)), "cr_learn_collect")
        return n.value

    def learn_dataset_size(self):
        n = C.c_uint64()
        self.check(self.lib.cr_learn_dataset_size(self.h, C.byref(n)), "cr_learn_dataset_size")
        return n.value

    def learn_dataset_clear(self):
        self.check(self.lib.cr_learn_dataset_clear(self.h), "cr_learn_dataset_clear")

    def learn_dataset(self, first=0, n=None):
        if n is None:
            n = self.learn_dataset_size() - first
        X = np.empty((n, 224), np.uint16)
        T = np.empty((n, 3), np.float32)
        self.check(self.lib.cr_learn_dataset_download(self.h, first, n, _ptr(X), _ptr(T)), "cr_learn_dataset_download")
        return X, T

    def learn_train(self, steps, batch, lr, seed=1):
        loss = C.c_float()
        self.check(self.lib.cr_learn_train(self.h, steps, batch, lr, seed, C.byref(loss)), "cr_learn_train")
        return loss.value

    def learn_render(self, frames):
        arr = (structs.PushConstants * len(frames))(*frames)
        self.check(self.lib.cr_learn_render(self.h, arr, len(frames)), "cr_learn_render")

    def mlp_init(self, seed=1):
        self.check(self.lib.cr_mlp_init(self.h, seed), "cr_mlp_init")

    def mlp_set_ema(self, decay):
        self.check(self.lib.cr_mlp_set_ema(self.h, decay), "cr_mlp_set_ema")

    # ------------------------------------------------------- batched render
    def batch_add(self, frames):
        arr = (structs.PushConstants * len(frames))(*frames)
        j = C.c_uint32()
        self.check(self.lib.cr_batch_add(self.h, arr, len(frames), C.byref(j)), "cr_batch_add")
        return j.value

    def batch_clear(self):
        self.check(self.lib.cr_batch_clear(self.h), "cr_batch_clear")

    def batch_size(self):
        n = C.c_uint32()
        self.check(self.lib.cr_batch_size(self.h, C.byref(n)), "cr_batch_size")
        return n.value

    def batch_set_frames(self, job, frames):
        arr = (structs.PushConstants * len(frames))(*frames)
        self.check(self.lib.cr_batch_set_frames(self.h, job, arr, len(frames)), "cr_batch_set_frames")

    def batch_clear_images(self):
        self.check(self.lib.cr_batch_clear_images(self.h), "cr_batch_clear_images")

    def batch_set_tuning(self, block_threads=256, blocks_per_cu=0, tile_width=8):
        t = structs.BatchTuning(block_threads, blocks_per_cu, tile_width, 0)
        self.check(self.lib.cr_batch_set_tuning(self.h, C.byref(t)), "cr_batch_set_tuning")

    def batch_tuning(self):
        t = structs.BatchTuning()
        self.check(self.lib.cr_batch_get_tuning(self.h, C.byref(t)), "cr_batch_get_tuning")
        return t

    def batch_render(self):
        self.check(self.lib.cr_batch_render(self.h), "cr_batch_render")

    def batch_download(self, job):
        w, h = C.c_uint32(), C.c_uint32()
        self.check(self.lib.cr_batch_download(self.h, job, None, C.byref(w), C.byref(h)), "cr_batch_download")
        out = np.empty((h.value, w.value, 4), np.float32)
        self.check(self.lib.cr_batch_download(self.h, job, _ptr(out), None, None), "cr_batch_download")
        return out

    def batch_stats(self):
        s = structs.PtStats()
        self.check(self.lib.cr_batch_stats_get(self.h, C.byref(s)), "cr_batch_stats_get")
        return s

    def batch_stats_reset(self):
        self.check(self.lib.cr_batch_stats_reset(self.h), "cr_batch_stats_reset")

    def batch_autotune(self, tune_bricks=True):
        t = structs.BatchTuning()
        r = structs.BatchTuneReport()
        self.check(self.lib.cr_batch_autotune(self.h, int(tune_bricks), C.byref(t), C.byref(r)), "cr_batch_autotune")
        return t, r

    # -------------------------------------------------------- photon beams
    def trace_beams(self, params, max_beams):
        nb, nd = C.c_uint32(), C.c_uint32()
        self.check(self.lib.cr_ppb_trace_beams(self.h, C.byref(params), max_beams, C.byref(nb), C.byref(nd)),
                   "cr_ppb_trace_beams")
        return nb.value, nd.value

    def upload_beams(self, beams, data):
        self.check(self.lib.cr_beams_upload(self.h, beams, len(beams), data, len(data)), "cr_beams_upload")

    def beams_count(self):
        n, nd = C.c_uint32(), C.c_uint32()
        self.check(self.lib.cr_beams_count(self.h, C.byref(n), C.byref(nd)), "cr_beams_count")
        return n.value, nd.value

    def download_beams(self):
        n, nd = self.beams_count()
        beams = (structs.PhotonBeam * n)()
        data = (structs.PhotonBeamData * nd)()
        self.check(self.lib.cr_beams_download(self.h, beams, data), "cr_beams_download")
        return beams, data

    def lbvh_build(self):
        self.check(self.lib.cr_lbvh_build(self.h), "cr_lbvh_build")

    def lbvh_download(self, nodes=True):
        n, _ = self.beams_count()
        keys = np.empty(n, np.uint32)
        perm = np.empty(n, np.uint32)
        tn = (structs.TreeNode * max(2 * n - 1, 1))() if nodes else None
        self.check(self.lib.cr_lbvh_download(self.h, _ptr(keys), _ptr(perm), tn), "cr_lbvh_download")
        return keys, perm, tn

    def lbvh_timings(self):
        t = (C.c_double * 5)()
        self.check(self.lib.cr_lbvh_timings(self.h, t), "cr_lbvh_timings")
        return dict(zip(["morton", "sort", "hierarchy", "refit", "pack"], list(t)))

    def gather(self, frames, ext=None):
        arr = (structs.PushConstants * len(frames))(*frames)
        self.check(self.lib.cr_ppb_gather(self.h, arr, len(frames), C.byref(ext) if ext is not None else None),
                   "cr_ppb_gather")

    def gather_stats(self):
        s = structs.GatherStats()
        self.check(self.lib.cr_gather_stats_get(self.h, C.byref(s)), "cr_gather_stats_get")
        return s


# --------------------------------------------------------------------------
# Technique objects: the reference's plugin interface (RenderTechnique.h).
# QueueUpdate* records a binding, UpdateDescriptorSets commits the queue to
# the context, RecordDrawCommands issues the technique's dispatches.
# --------------------------------------------------------------------------
class RenderTechnique:
    """RenderTechnique.h:15-66 without Vulkan descriptor sets."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._queue = []

    def QueueUpdateCloudData(self, cloud_properties, imageIdx=0):
        self._queue.append(("cloud", cloud_properties))

    def QueueUpdateCloudDataSampler(self, grid, imageIdx=0):
        # grid: dict(data=ndarray[z,y,x], voxel_size=..., storage=...) or None (keep the context grid)
        if grid is not None:
            self._queue.append(("grid", grid))

    def QueueUpdateCameraProperties(self, camera, imageIdx=0):
        self._queue.append(("camera", camera))

    def QueueUpdateParameters(self, parameters, imageIdx=0):
        self._queue.append(("parameters", parameters))

    def QueueUpdateShadowVolume(self, sv_properties, imageIdx=0):
        self._queue.append(("shadow_volume", sv_properties))

    def QueueUpdateShadowVolumeSampler(self, _sv_image=None, imageIdx=0):
        pass  # the shadow volume image is owned by the context

    def UpdateDescriptorSets(self):
        for kind, value in self._queue:
            if kind == "cloud":
                self.ctx.set_cloud(value)
            elif kind == "grid":
                self.ctx.upload_grid(value["data"], value["voxel_size"], storage=value.get("storage", 0))
            elif kind == "camera":
                self.ctx.set_camera(value)
            elif kind == "parameters":
                self.ctx.set_parameters(value)
            elif kind == "shadow_volume":
                self.ctx.set_shadow_volume(value)
        self._queue.clear()

    def GetRequiredSetCount(self):
        return 1


class RenderTechniqueSV(RenderTechnique):
    """RenderTechniqueSV (RenderTechniqueSV.cpp:102-113): builds the shadow volume."""

    def RecordDrawCommands(self, imageIndex=0):
        self.UpdateDescriptorSets()
        self.ctx.sv_build()


class RenderTechniquePT(RenderTechnique):
    """RenderTechniquePT (RenderTechniquePT.cpp:129-168): one call renders len(frames) frames."""

    def RecordDrawCommands(self, frames, imageIndex=0):
        self.UpdateDescriptorSets()
        if isinstance(frames, structs.PushConstants):
            frames = [frames]
        self.ctx.pt_render(list(frames))


class RenderTechniquePPB(RenderTechnique):
    """RenderTechniquePPB (RenderTechniquePPB.cpp:436-519): beams -> LBVH -> gather."""

    def __init__(self, ctx, initialRadius=0.002):
        super().__init__(ctx)
        self.initialRadius = initialRadius
        self.radius = initialRadius
        self.trace_params = None
        self.max_beams = 1 << 22

    def UpdateRadius(self, frameNumber):
        """RenderTechniquePPB::UpdateRadius (RenderTechniquePPB.cpp:530-546)."""
        self.radius = self.ctx.lib.cr_ppb_update_radius(frameNumber, structs.CR_PPB_BEAMS_PER_PASS,
                                                        structs.CR_PPB_ALPHA, self.initialRadius, self.radius)
        return self.radius

    def AllocateResources(self, trace_params=None, max_beams=1 << 22):
        """RenderTechniquePPB.cpp:282-355: every beam, sort and tree buffer for
        up to max_beams beams, allocated up front."""
        self.trace_params = trace_params
        self.max_beams = max_beams
        self.ctx.check(self.ctx.lib.cr_ppb_allocate(self.ctx.h, max_beams), "cr_ppb_allocate")

    def FreeResources(self):
        self.ctx.check(self.ctx.lib.cr_ppb_free(self.ctx.h), "cr_ppb_free")

    def RecordDrawCommands(self, frames, imageIndex=0, gather_ext=None, retrace=True):
        """One frame per call, as the reference: the radius of frames[0]'s
        frameCount (RenderTechniquePPB.cpp:439), beams, LBVH, gather."""
        self.UpdateDescriptorSets()
        if isinstance(frames, structs.PushConstants):
            frames = [frames]
        if retrace and self.trace_params is not None:
            self.trace_params.radius = self.UpdateRadius(frames[0].frameCount)
            self.ctx.trace_beams(self.trace_params, self.max_beams)
        self.ctx.lbvh_build()
        self.ctx.gather(list(frames), gather_ext)


class RenderTechniquePPM(RenderTechnique):
    """RenderTechniquePPM (RenderTechniquePPM.cpp:229-312): per frame the
    photon map is traced again with the frame's seed, then the estimate blends
    the frame into the accumulation image; the radius shrinks per frame as
    UpdateRadius does."""

    def __init__(self, ctx, photon_map_properties=None, initialRadius=10.0, params=None):
        super().__init__(ctx)
        self.initialRadius = initialRadius
        self.radius = initialRadius
        self.params = params if params is not None else structs.PpmParams()
        if params is None:
            ctx.lib.cr_ppm_params_default(C.byref(self.params))
        if photon_map_properties is not None:
            ctx.set_photon_map(photon_map_properties)

    def UpdateRadius(self, frameNumber):
        self.radius = self.ctx.lib.cr_ppm_update_radius(frameNumber, structs.CR_PPM_ALPHA, self.initialRadius,
                                                        self.radius)
        return self.radius

    def RecordDrawCommands(self, frames, imageIndex=0):
        self.UpdateDescriptorSets()
        if isinstance(frames, structs.PushConstants):
            frames = [frames]
        for f in frames:
            f = structs.PushConstants(f.time, f.seed, f.frameCount, self.UpdateRadius(f.frameCount), f.currentBuffer)
            self.ctx.ppm_trace(f.seed, self.params)
            self.ctx.ppm_estimate([f], self.params)
