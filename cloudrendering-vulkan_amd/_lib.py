"""Loader and ctypes signatures of libcrmi355x.so (include/cr/abi.h).

The library is the product: there is no Python or CPU fallback.  If the
shared object is missing, loading raises ``LibraryNotBuilt``; build it with
``python -c "import __graft_entry__ as g; g.build()"`` or ``make -C
cloudrendering-vulkan_amd``.
"""
import ctypes as C
import os

from . import structs as S

HERE = os.path.dirname(os.path.abspath(__file__))
# CR_LIB_PATH: another build of the same sources (the host-sanitizer build)
LIB_PATH = os.environ.get("CR_LIB_PATH") or os.path.join(HERE, "libcrmi355x.so")

P = C.POINTER
vp = C.c_void_p
ctx_p = C.c_void_p
st = C.c_int32

# name -> (restype, argtypes)
SIGNATURES = {
    "cr_status_string": (C.c_char_p, [st]),
    "cr_last_error": (C.c_char_p, [ctx_p]),
    "cr_version": (C.c_char_p, []),
    "cr_create": (st, [C.c_int, P(ctx_p)]),
    "cr_destroy": (st, [ctx_p]),
    "cr_synchronize": (st, [ctx_p]),
    "cr_stream": (vp, [ctx_p]),
    "cr_camera_default": (None, [P(S.CameraProperties)]),
    "cr_camera_set_resolution": (None, [P(S.CameraProperties), C.c_int, C.c_int]),
    "cr_camera_set_fov": (None, [P(S.CameraProperties), P(C.c_float)]),
    "cr_camera_set_rotation": (None, [P(S.CameraProperties), P(C.c_float)]),
    "cr_parameters_default": (None, [P(S.Parameters)]),
    "cr_parameters_set_phase_g": (None, [P(S.Parameters), C.c_float]),
    "cr_cloud_default": (None, [P(S.CloudProperties)]),
    "cr_shadow_volume_default": (None, [P(S.ShadowVolumeProperties)]),
    "cr_shadow_volume_set_light": (None, [P(S.ShadowVolumeProperties), P(C.c_float), P(C.c_float), P(C.c_float)]),
    "cr_msvc_rand_sequence": (None, [C.c_uint32, C.c_uint32, P(C.c_int32)]),
    "cr_set_grid_layout": (st, [ctx_p, C.c_int]),
    "cr_grid_upload": (st, [ctx_p, P(S.GridHeader), vp, C.c_int, C.c_int, C.c_int]),
    "cr_grid_load_xyz": (st, [ctx_p, C.c_char_p, C.c_int]),
    "cr_grid_save_xyz": (st, [ctx_p, C.c_char_p]),
    "cr_grid_download": (st, [ctx_p, vp]),
    "cr_grid_majorants": (st, [ctx_p, P(C.c_uint32), vp]),
    "cr_grid_skip": (st, [ctx_p, vp]),
    "cr_fbm_params_default": (None, [P(S.FbmParams)]),
    "cr_grid_generate_fbm": (st, [ctx_p, P(S.FbmParams), C.c_int]),
    "cr_set_camera": (st, [ctx_p, P(S.CameraProperties)]),
    "cr_set_cloud": (st, [ctx_p, P(S.CloudProperties)]),
    "cr_get_cloud": (st, [ctx_p, P(S.CloudProperties)]),
    "cr_set_parameters": (st, [ctx_p, P(S.Parameters)]),
    "cr_set_shadow_volume": (st, [ctx_p, P(S.ShadowVolumeProperties)]),
    "cr_render_ext_default": (None, [P(S.RenderExt)]),
    "cr_set_render_ext": (st, [ctx_p, P(S.RenderExt)]),
    "cr_sv_build": (st, [ctx_p]),
    "cr_sv_dims": (st, [ctx_p, P(C.c_uint32)]),
    "cr_sv_download": (st, [ctx_p, vp]),
    "cr_pt_clear": (st, [ctx_p]),
    "cr_pt_render": (st, [ctx_p, P(S.PushConstants), C.c_uint32]),
    "cr_pt_render_static": (st, [ctx_p, P(S.PushConstants), C.c_uint32, vp]),
    "cr_pt_progressive_reset": (st, [ctx_p, C.c_uint32, C.c_uint32, C.c_int]),
    "cr_pt_progressive": (st, [ctx_p, C.c_uint32, P(S.ProgressiveState)]),
    "cr_pt_write_pfm": (st, [ctx_p, C.c_char_p]),
    "cr_pt_download": (st, [ctx_p, vp]),
    "cr_pt_upload": (st, [ctx_p, vp]),
    "cr_pt_stats_get": (st, [ctx_p, P(S.PtStats)]),
    "cr_pt_stats_reset": (st, [ctx_p]),
    "cr_batch_tuning_default": (None, [P(S.BatchTuning)]),
    "cr_batch_add": (st, [ctx_p, P(S.PushConstants), C.c_uint32, P(C.c_uint32)]),
    "cr_batch_clear": (st, [ctx_p]),
    "cr_batch_size": (st, [ctx_p, P(C.c_uint32)]),
    "cr_batch_set_frames": (st, [ctx_p, C.c_uint32, P(S.PushConstants), C.c_uint32]),
    "cr_batch_clear_images": (st, [ctx_p]),
    "cr_batch_set_tuning": (st, [ctx_p, P(S.BatchTuning)]),
    "cr_batch_get_tuning": (st, [ctx_p, P(S.BatchTuning)]),
    "cr_batch_render": (st, [ctx_p]),
    "cr_batch_download": (st, [ctx_p, C.c_uint32, vp, P(C.c_uint32), P(C.c_uint32)]),
    "cr_batch_stats_get": (st, [ctx_p, P(S.PtStats)]),
    "cr_batch_stats_reset": (st, [ctx_p]),
    "cr_batch_autotune": (st, [ctx_p, C.c_int, P(S.BatchTuning), P(S.BatchTuneReport)]),
    "cr_beam_trace_params_default": (None, [P(S.BeamTraceParams)]),
    "cr_ppb_update_radius": (C.c_float, [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float]),
    "cr_ppb_trace_beams": (st, [ctx_p, P(S.BeamTraceParams), C.c_uint32, P(C.c_uint32), P(C.c_uint32)]),
    "cr_beams_upload": (st, [ctx_p, vp, C.c_uint32, vp, C.c_uint32]),
    "cr_beams_count": (st, [ctx_p, P(C.c_uint32), P(C.c_uint32)]),
    "cr_beams_download": (st, [ctx_p, vp, vp]),
    "cr_ppb_allocate": (st, [ctx_p, C.c_uint32]),
    "cr_ppb_free": (st, [ctx_p]),
    "cr_ppm_params_default": (None, [P(S.PpmParams)]),
    "cr_photon_map_default": (None, [P(S.PhotonMapProperties)]),
    "cr_photon_map_set_bounds": (None, [P(S.PhotonMapProperties), P(C.c_float), P(C.c_float)]),
    "cr_set_photon_map": (st, [ctx_p, P(S.PhotonMapProperties)]),
    "cr_ppm_update_radius": (C.c_float, [C.c_uint32, C.c_float, C.c_float, C.c_float]),
    "cr_ppm_trace": (st, [ctx_p, C.c_int32, P(S.PpmParams), P(C.c_uint32)]),
    "cr_ppm_estimate": (st, [ctx_p, P(S.PushConstants), C.c_uint32, P(S.PpmParams)]),
    "cr_ppm_download": (st, [ctx_p, P(C.c_uint32), P(C.c_uint32), vp, vp]),
    "cr_lbvh_build": (st, [ctx_p]),
    "cr_lbvh_download": (st, [ctx_p, vp, vp, vp]),
    "cr_lbvh_timings": (st, [ctx_p, P(C.c_double)]),
    "cr_gather_ext_default": (None, [P(S.GatherExt)]),
    "cr_ppb_gather": (st, [ctx_p, P(S.PushConstants), C.c_uint32, P(S.GatherExt)]),
    "cr_gather_stats_get": (st, [ctx_p, P(S.GatherStats)]),
    "cr_mlp_param_layout": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
    "cr_mlp_init": (st, [ctx_p, C.c_uint32]),
    "cr_mlp_set_params": (st, [ctx_p, vp, C.c_uint32]),
    "cr_mlp_get_params": (st, [ctx_p, vp, C.c_uint32]),
    "cr_mlp_set_ema": (st, [ctx_p, C.c_float]),
    "cr_mlp_get_ema": (st, [ctx_p, vp, C.c_uint32]),
    "cr_mlp_forward": (st, [ctx_p, vp, C.c_uint32, vp]),
    "cr_mlp_train_step": (st, [ctx_p, vp, vp, C.c_uint32, C.c_float, P(C.c_float), vp]),
    "cr_mlp_probe": (st, [ctx_p, vp, vp, vp]),
    "cr_math_probe": (st, [ctx_p, C.c_int, vp, C.c_uint32, vp]),
    "cr_learn_build_mips": (st, [ctx_p]),
    "cr_learn_features": (st, [ctx_p, vp, C.c_uint32, vp]),
    "cr_pt_record": (st, [ctx_p, P(S.PushConstants), C.c_uint32, vp]),
    "cr_pt_continue": (st, [ctx_p, vp, C.c_uint32, C.c_uint32, vp]),
    "cr_learn_collect": (st, [ctx_p, P(S.PushConstants), C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_uint32)]),
    "cr_learn_dataset_size": (st, [ctx_p, P(C.c_uint64)]),
    "cr_learn_dataset_clear": (st, [ctx_p]),
    "cr_learn_dataset_download": (st, [ctx_p, C.c_uint64, C.c_uint32, vp, vp]),
    "cr_learn_train": (st, [ctx_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, P(C.c_float)]),
    "cr_learn_render": (st, [ctx_p, P(S.PushConstants), C.c_uint32]),
}


class LibraryNotBuilt(RuntimeError):
    pass


_lib = None


def load(path=None):
    """Load libcrmi355x.so once and attach the ctypes signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibraryNotBuilt(f"{p} is missing: the gfx950 library has not been built")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
