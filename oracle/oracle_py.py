"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU restatement of the reference
hot path.  Imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the CPU baseline, never as the
thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    import subprocess
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


class Oracle:
    def __init__(self, S, lib_path=None):
        """S: the structs module of the package (ctypes mirrors of the ABI);
        lib_path: another build of liboracle's sources (default oracle/liboracle.so)."""
        if lib_path is None and not os.path.exists(LIB):
            build()
        self.S = S
        self.lib = L = C.CDLL(lib_path or LIB)
        P = C.POINTER
        vp = C.c_void_p

        class OrScene(C.Structure):
            _fields_ = [("grid_f32", vp), ("grid_f16", vp), ("dims", C.c_uint32 * 3), ("majorants", vp),
                        ("mdims", C.c_uint32 * 3), ("sv", vp), ("svdims", C.c_uint32 * 3),
                        ("cam", S.CameraProperties), ("cloud", S.CloudProperties), ("params", S.Parameters),
                        ("svp", S.ShadowVolumeProperties), ("ext", S.RenderExt)]

        class OrCounters(C.Structure):
            _fields_ = [("paths", C.c_uint64), ("vertices", C.c_uint64), ("lookups", C.c_uint64),
                        ("svLookups", C.c_uint64), ("brickVisits", C.c_uint64)]

        class OrBeamScene(C.Structure):
            _fields_ = [("base", OrScene), ("beams", vp), ("data", vp), ("nodes", vp), ("n", C.c_uint32),
                        ("refCompat", C.c_uint32), ("jitter", C.c_uint32), ("albedo", C.c_float)]

        class OrPhotonMap(C.Structure):
            _fields_ = [("arrivals", vp), ("start", vp), ("photons", vp)]

        self.OrScene, self.OrCounters, self.OrBeamScene, self.OrPhotonMap = OrScene, OrCounters, OrBeamScene, OrPhotonMap
        sig = {
            "or_logf": (C.c_float, [C.c_float]),
            "or_expf": (C.c_float, [C.c_float]),
            "or_sincosf": (None, [C.c_float, P(C.c_float), P(C.c_float)]),
            "or_math_array": (None, [C.c_int, vp, C.c_uint32, vp]),
            "or_xorshift32": (C.c_uint32, [C.c_uint32]),
            "or_rng_stream": (None, [C.c_int32, C.c_uint32, C.c_uint32, P(C.c_uint32)]),
            "or_msvc_rand": (None, [C.c_uint32, C.c_uint32, P(C.c_int32)]),
            "or_trilinear_f32": (C.c_float, [vp, P(C.c_uint32), P(C.c_float)]),
            "or_phase_pdf": (C.c_float, [P(S.Parameters), P(C.c_float), P(C.c_float)]),
            "or_scatter_dir": (None, [P(S.Parameters), P(C.c_float), C.c_uint32, P(C.c_float), P(C.c_uint32)]),
            "or_background": (None, [P(S.RenderExt), P(C.c_float), P(C.c_float)]),
            "or_camera_ray": (None, [P(S.CameraProperties), C.c_int32, C.c_int32, P(C.c_float), P(C.c_float)]),
            "or_intersect_box": (C.c_int, [P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                           P(C.c_float)]),
            "or_sv_dims": (None, [P(OrScene), P(C.c_uint32)]),
            "or_sv_build": (None, [P(OrScene), vp]),
            "or_pt_sample": (None, [P(OrScene), C.c_int32, C.c_int32, C.c_int32, P(C.c_float), P(OrCounters)]),
            "or_pt_render": (None, [P(OrScene), P(S.PushConstants), C.c_uint32, vp, C.c_int32, P(OrCounters)]),
            "or_pt_render_region": (None, [P(OrScene), P(S.PushConstants), C.c_uint32, vp, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_int32, C.c_int32, P(OrCounters)]),
            "or_hash32": (C.c_uint32, [C.c_uint32]),
            "or_fbm_voxel": (C.c_float, [P(S.FbmParams), C.c_uint32, C.c_uint32, C.c_uint32]),
            "or_fbm_grid": (None, [P(S.FbmParams), vp, C.c_int32]),
            "or_f32_to_f16": (C.c_uint16, [C.c_float]),
            "or_f16_to_f32": (C.c_float, [C.c_uint16]),
            "or_grid_max": (C.c_float, [vp, C.c_uint64]),
            "or_brick_majorants": (None, [vp, P(C.c_uint32), C.c_uint32, vp, P(C.c_uint32)]),
            "or_expand_bits10": (C.c_uint32, [C.c_uint32]),
            "or_morton24_ref": (C.c_uint32, [P(C.c_float), P(C.c_float)]),
            "or_morton30": (C.c_uint32, [P(C.c_float), P(C.c_float), P(C.c_float)]),
            "or_beam_morton30": (None, [vp, C.c_uint32, P(C.c_float), P(C.c_float), vp]),
            "or_radix_sort_pairs": (None, [vp, C.c_uint32, vp, vp]),
            "or_karras_build": (None, [vp, C.c_uint32, vp]),
            "or_topdown_build": (None, [vp, C.c_uint32, vp]),
            "or_leaf_bounds": (None, [vp, P(C.c_float), P(C.c_float)]),
            "or_refit": (None, [vp, C.c_uint32, vp]),
            "or_lbvh_build": (None, [vp, C.c_uint32, P(C.c_float), P(C.c_float), vp, vp, vp]),
            "or_trace_beams_count": (None, [P(OrScene), P(S.BeamTraceParams), vp, vp]),
            "or_trace_beams_write": (None, [P(OrScene), P(S.BeamTraceParams), vp, vp, vp, vp, C.c_uint64,
                                            C.c_uint64]),
            "or_gather_pixel": (None, [P(OrBeamScene), C.c_int32, C.c_int32, C.c_int32, P(C.c_float),
                                       P(C.c_uint64), P(C.c_uint64)]),
            "or_gather_pixel_brute": (None, [P(OrBeamScene), C.c_int32, C.c_int32, C.c_int32, P(C.c_float)]),
            "or_gather_render": (None, [P(OrBeamScene), P(S.PushConstants), C.c_uint32, vp, C.c_int32]),
            "or_ppb_update_radius": (C.c_float, [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float]),
            "or_sqrt_correction_errors": (C.c_uint64, [C.c_uint32, C.c_uint32]),
            "or_ppm_trace_count": (None, [P(OrScene), P(S.PhotonMapProperties), P(S.PpmParams), C.c_int32, vp]),
            "or_ppm_trace_write": (None, [P(OrScene), P(S.PhotonMapProperties), P(S.PpmParams), C.c_int32, vp, vp,
                                          vp]),
            "or_ppm_build": (None, [vp, vp, C.c_uint32, C.c_uint32, vp, vp, vp]),
            "or_ppm_query": (None, [P(OrScene), P(S.PhotonMapProperties), P(OrPhotonMap), P(C.c_float),
                                    P(C.c_float), C.c_float, C.c_float, P(C.c_float)]),
            "or_ppm_estimate_render": (None, [P(OrScene), P(S.PhotonMapProperties), P(OrPhotonMap), P(S.PpmParams),
                                              P(S.PushConstants), C.c_uint32, vp, C.c_int32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    # --------------------------------------------------------------- scene
    def scene(self, grid, cloud, cam, params, svp, ext, sv=None, majorants=None):
        """grid: float32 or float16 array [z, y, x]; sv: float32 [z, y, x]."""
        S = self.S
        s = self.OrScene()
        s._keep = []
        g = np.ascontiguousarray(grid)
        s._keep.append(g)
        if g.dtype == np.float16:
            s.grid_f16 = g.ctypes.data
        else:
            g = np.ascontiguousarray(g, np.float32)
            s._keep.append(g)
            s.grid_f32 = g.ctypes.data
        nz, ny, nx = g.shape
        s.dims[:] = (nx, ny, nz)
        if majorants is None:
            majorants, md = self.brick_majorants(np.asarray(grid, np.float32), ext.majorantBrick or 16)
        else:
            md = majorants.shape[::-1]
        majorants = np.ascontiguousarray(majorants, np.float32)
        s._keep.append(majorants)
        s.majorants = majorants.ctypes.data
        s.mdims[:] = tuple(int(v) for v in md)
        if sv is not None:
            sv = np.ascontiguousarray(sv, np.float32)
            s._keep.append(sv)
            s.sv = sv.ctypes.data
            s.svdims[:] = sv.shape[::-1]
        s.cam, s.cloud, s.params, s.svp, s.ext = cam, cloud, params, svp, ext
        return s

    def with_sv(self, s, sv):
        sv = np.ascontiguousarray(sv, np.float32)
        s._keep.append(sv)
        s.sv = sv.ctypes.data
        s.svdims[:] = sv.shape[::-1]
        return s

    # ----------------------------------------------------------------- PT
    def sv_dims(self, s):
        d = (C.c_uint32 * 3)()
        self.lib.or_sv_dims(C.byref(s), d)
        return tuple(d)

    def sv_build(self, s):
        dx, dy, dz = self.sv_dims(s)
        out = np.empty((dz, dy, dx), np.float32)
        self.lib.or_sv_build(C.byref(s), out.ctypes.data)
        return out

    def pt_render(self, s, frames, accum=None, threads=0):
        H, W = 2 * s.cam.halfHeight, 2 * s.cam.halfWidth
        a = np.zeros((H, W, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        arr = (self.S.PushConstants * len(frames))(*frames)
        cnt = self.OrCounters()
        self.lib.or_pt_render(C.byref(s), arr, len(frames), a.ctypes.data, threads, C.byref(cnt))
        return a, cnt

    def pt_render_region(self, s, frames, x0, y0, w, h, threads=0):
        a = np.zeros((h, w, 4), np.float32)
        arr = (self.S.PushConstants * len(frames))(*frames)
        cnt = self.OrCounters()
        self.lib.or_pt_render_region(C.byref(s), arr, len(frames), a.ctypes.data, x0, y0, w, h, threads,
                                     C.byref(cnt))
        return a, cnt

    def pt_sample(self, s, x, y, seed):
        out = (C.c_float * 4)()
        cnt = self.OrCounters()
        self.lib.or_pt_sample(C.byref(s), x, y, seed, out, C.byref(cnt))
        return np.array(out, np.float32), cnt

    # --------------------------------------------------------------- grids
    def math_array(self, op, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty_like(x)
        self.lib.or_math_array(op, x.ctypes.data, x.size, out.ctypes.data)
        return out

    def fbm_grid(self, p, threads=0):
        out = np.empty((p.size[2], p.size[1], p.size[0]), np.float32)
        self.lib.or_fbm_grid(C.byref(p), out.ctypes.data, threads)
        return out

    def brick_majorants(self, grid, brick=16):
        g = np.ascontiguousarray(grid, np.float32)
        nz, ny, nx = g.shape
        b = brick
        md = ((nx + b - 1) // b, (ny + b - 1) // b, (nz + b - 1) // b)
        out = np.empty((md[2], md[1], md[0]), np.float32)
        dims = (C.c_uint32 * 3)(nx, ny, nz)
        mo = (C.c_uint32 * 3)()
        self.lib.or_brick_majorants(g.ctypes.data, dims, b, out.ctypes.data, mo)
        return out, tuple(mo)

    def f32_to_f16(self, a):
        a = np.asarray(a, np.float32).ravel()
        return np.array([self.lib.or_f32_to_f16(float(v)) for v in a], np.uint16)

    def ppb_radii(self, frames, beams_per_pass, alpha, initial):
        """Radius per frame 0..frames-1, each update fed the previous radius
        (RecordDrawCommands calls UpdateRadius once per frame)."""
        out = np.empty(frames, np.float32)
        r = np.float32(0.0)
        for f in range(frames):
            r = np.float32(self.lib.or_ppb_update_radius(f, beams_per_pass, alpha, initial, float(r)))
            out[f] = r
        return out

    # ------------------------------------------------------ photon mapping
    def ppm_trace(self, s, pm, params, seed):
        """Deposits in emission order: (photons [n] of cr_photon as a structured
        uint8 view, cells [n])."""
        P_ = params.numPaths
        counts = np.empty(P_, np.uint32)
        self.lib.or_ppm_trace_count(C.byref(s), C.byref(pm), C.byref(params), seed, counts.ctypes.data)
        offsets = np.zeros(P_, np.uint64)
        offsets[1:] = np.cumsum(counts[:-1], dtype=np.uint64)
        n = int(counts.sum())
        photons = (self.S.Photon * max(n, 1))()
        cells = np.empty(max(n, 1), np.uint32)
        self.lib.or_ppm_trace_write(C.byref(s), C.byref(pm), C.byref(params), seed, offsets.ctypes.data, photons,
                                    cells.ctypes.data)
        return photons, cells[:n], n

    def ppm_build(self, photons, cells, n, n_cells):
        arrivals = np.empty(n_cells, np.uint32)
        start = np.empty(n_cells, np.uint32)
        sorted_ = (self.S.Photon * max(n, 1))()
        cells = np.ascontiguousarray(cells, np.uint32)
        self.lib.or_ppm_build(photons, cells.ctypes.data, n, n_cells, arrivals.ctypes.data, start.ctypes.data, sorted_)
        return arrivals, start, sorted_

    def photon_map(self, arrivals, start, sorted_):
        m = self.OrPhotonMap()
        m._keep = [arrivals, start, sorted_]
        m.arrivals = arrivals.ctypes.data
        m.start = start.ctypes.data
        m.photons = C.addressof(sorted_)
        return m

    def ppm_query(self, s, pm, m, pos, dir_, scatter, radius):
        out = (C.c_float * 4)()
        f3 = C.c_float * 3
        self.lib.or_ppm_query(C.byref(s), C.byref(pm), C.byref(m), f3(*pos), f3(*dir_), scatter, radius, out)
        return np.array(out, np.float32)

    def ppm_estimate(self, s, pm, m, params, frames, accum=None, threads=0):
        H, W = 2 * s.cam.halfHeight, 2 * s.cam.halfWidth
        a = np.zeros((H, W, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        arr = (self.S.PushConstants * len(frames))(*frames)
        self.lib.or_ppm_estimate_render(C.byref(s), C.byref(pm), C.byref(m), C.byref(params), arr, len(frames),
                                        a.ctypes.data, threads)
        return a

    # ---------------------------------------------------------------- LBVH
    def lbvh_build(self, beams, b0, b1):
        n = len(beams)
        keys = np.empty(n, np.uint32)
        perm = np.empty(n, np.uint32)
        nodes = (self.S.TreeNode * max(2 * n - 1, 1))()
        f3 = C.c_float * 3
        self.lib.or_lbvh_build(beams, n, f3(*b0[:3]), f3(*b1[:3]), keys.ctypes.data, perm.ctypes.data, nodes)
        return keys, perm, nodes

    def radix_sort(self, keys):
        keys = np.ascontiguousarray(keys, np.uint32)
        ko = np.empty_like(keys)
        po = np.empty_like(keys)
        self.lib.or_radix_sort_pairs(keys.ctypes.data, len(keys), ko.ctypes.data, po.ctypes.data)
        return ko, po

    def karras(self, sorted_keys, topdown=False):
        k = np.ascontiguousarray(sorted_keys, np.uint32)
        n = len(k)
        nodes = (self.S.TreeNode * max(2 * n - 1, 1))()
        (self.lib.or_topdown_build if topdown else self.lib.or_karras_build)(k.ctypes.data, n, nodes)
        return nodes

    def trace_beams(self, s, p, max_beams):
        P_ = p.numPaths
        segs = np.empty(P_, np.uint32)
        datas = np.empty(P_, np.uint32)
        self.lib.or_trace_beams_count(C.byref(s), C.byref(p), segs.ctypes.data, datas.ctypes.data)
        so = np.zeros(P_, np.uint64)
        do = np.zeros(P_, np.uint64)
        sb = sd = 0
        for i in range(P_):
            so[i] = sb
            do[i] = sd
            if sb < max_beams:
                sd += int(datas[i])
            sb += int(segs[i])
        nb = min(sb, max_beams)
        beams = (self.S.PhotonBeam * max(nb, 1))()
        data = (self.S.PhotonBeamData * max(sd, 1))()
        self.lib.or_trace_beams_write(C.byref(s), C.byref(p), so.ctypes.data, do.ctypes.data, beams, data, nb, sd)
        return beams, data, nb, sd

    def beam_scene(self, s, sorted_beams, data, nodes, n, ref_compat=0, jitter=0, albedo=1.0):
        b = self.OrBeamScene()
        b.base = s
        b._keep = [s, sorted_beams, data, nodes]
        b.beams = C.addressof(sorted_beams)
        b.data = C.addressof(data)
        b.nodes = C.addressof(nodes)
        b.n = n
        b.refCompat = ref_compat
        b.jitter = jitter
        b.albedo = albedo
        return b

    def gather_pixel(self, b, x, y, seed, brute=False, stats=False):
        """Radiance of one pixel's ray; with stats, also (nodes visited, beams tested)."""
        out = (C.c_float * 4)()
        nv, nt = C.c_uint64(), C.c_uint64()
        if brute:
            self.lib.or_gather_pixel_brute(C.byref(b), x, y, seed, out)
        else:
            self.lib.or_gather_pixel(C.byref(b), x, y, seed, out, C.byref(nv), C.byref(nt))
        r = np.array(out, np.float32)
        return (r, nv.value, nt.value) if stats else r

    def gather_render(self, b, frames, accum=None, threads=0):
        H, W = 2 * b.base.cam.halfHeight, 2 * b.base.cam.halfWidth
        a = np.zeros((H, W, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        arr = (self.S.PushConstants * len(frames))(*frames)
        self.lib.or_gather_render(C.byref(b), arr, len(frames), a.ctypes.data, threads)
        return a
