#!/usr/bin/env python3
"""Benchmark of the volumetric hot path on one MI355X.

Default workload (the production-size plain path tracer): a 1024x512x1024
FP16 density grid (1 GiB, fBm anvil, 8 octaves) in 4^3-voxel bricks with
16^3 brick majorants, ratio-tracked shadow rays, albedo 0.999, HG g = 0.85,
1920x1080, up to 256 bounces, 256 spp.  One step = one persistent launch
that renders the workload's frames (1 spp each; `--spp-per-step` overrides
the count) of the full image; inputs are resident in HBM before timing
starts.

Prints one JSON line: value = path vertices per second (camera-ray origins
plus real collisions) over the timed steps, with
  roofline     : algorithmic bytes of the path-tracing kernel per launch
                 (16 B per FP16 trilinear lookup, 4 B per brick-majorant
                 visit, 16 B per shadow-volume lookup, 32 B per pixel) over
                 its average launch time, measured with HIP events on the
                 context stream, against 8 TB/s;
  cpu_baseline : the CPU oracle on the host cores, on a bounded sample of the
                 same render (a centred pixel rectangle, same streams).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to relMSE≤1e-3 at 1080p (learned vs plain PT); Mpath-vertices/s; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "mlp", "ppm"],
                    help="c1/c2: plain path tracer throughput; c4: 64 jobs batched vs one by one; "
                         "c5: learned vs plain time to relMSE <= 1e-3")
    ap.add_argument("--c5-clouds", type=int, default=400)
    ap.add_argument("--c5-points", type=int, default=80000, help="training points per cloud")
    ap.add_argument("--c5-target-spp", type=int, default=1024, help="continuations per training point")
    ap.add_argument("--c5-train-steps", type=int, default=80000)
    ap.add_argument("--c5-collect-frames", type=int, default=8, help="record-pass frames per training cloud")
    ap.add_argument("--c5-tracking", default="dda", choices=["dda", "global"],
                    help="dda: brick-majorant tracking + ratio-tracked shadow rays (the production tracker); "
                         "global: the reference's global majorant + shadow-volume NEE (c1's settings)")
    ap.add_argument("--c5-batch", type=int, default=8192)
    ap.add_argument("--c5-ema", type=float, default=0.0,
                    help="decay of the parameters' moving average used for inference (0: off)")
    ap.add_argument("--c5-ref-spp", type=int, default=16384, help="frames of the converged plain reference")
    ap.add_argument("--c5-width", type=int, default=1920)
    ap.add_argument("--c5-height", type=int, default=1080)
    ap.add_argument("--spp-per-step", type=int, default=0, help="frames per step; 0 = the workload's full spp")
    ap.add_argument("--c3-gather-spp", type=int, default=16, help="frames of the c3 beam gather (0: no gather)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c5", action="store_true",
                    help="leave the learned-vs-plain time-to-relMSE block (C5_QUICK) out of the default line")
    ap.add_argument("--via-batch", action="store_true",
                    help="render through the batched-launch API (one job, HIP-graph replay) instead of cr_pt_render")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "pt_c2_traffic.json"))
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# the sources that define the path-tracing kernel's code
PT_SOURCES = ("pt_kernel.h", "cr_device.h", "cr_internal.h", "kern_pt.hip")
LBVH_SOURCES = ("kern_lbvh.hip", "cr_internal_ppb.h", "cr_device.h")


def sources_sha16(names):
    import hashlib
    h = hashlib.sha256()
    for name in names:
        with open(os.path.join(ROOT, "cloudrendering-vulkan_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def pt_sources_sha16():
    return sources_sha16(PT_SOURCES)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the renderer runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by image tiles with one sum-reduce of the RGBA32F image)", file=sys.stderr)
        sys.exit(2)
    import cr_import
    pkg = cr_import.package()
    cfg = cr_import.configs()
    S = pkg.structs
    if args.workload == "c5":
        return bench_c5(args, pkg, cfg)
    if args.workload == "mlp":
        return bench_mlp(args, pkg)
    if args.workload == "ppm":
        return bench_ppm(args, pkg, cfg)
    if args.workload == "c4":
        return bench_c4(args, pkg, cfg)
    if args.workload == "c3":
        return bench_c3(args, pkg, cfg)

    sc = cfg.c2() if args.workload == "c2" else cfg.c1()
    ctx = pkg.Context(0)
    t0 = time.time()
    cloud, svp = sc.setup(ctx)
    log(f"{sc.name}: grid {tuple(sc.fbm.size)} generated, shadow volume {ctx.sv_dims()} built in "
        f"{time.time() - t0:.1f}s")
    spp = args.spp_per_step or sc.spp
    total_frames = (args.warmup + args.steps) * spp
    frames = cfg.frames(total_frames)
    W, H = sc.width, sc.height

    if args.via_batch:
        # the grid moves into the job; the baseline below needs it back
        ctx.batch_clear()
        ctx.batch_add(frames[:spp])

        def render(i):
            ctx.batch_set_frames(0, frames[i:i + spp])
            ctx.batch_render()
        stats, reset = ctx.batch_stats, ctx.batch_stats_reset
    else:
        def render(i):
            ctx.pt_render(frames[i:i + spp])
        stats, reset = ctx.pt_stats, ctx.pt_stats_reset

    for w in range(args.warmup):
        render(w * spp)
    ctx.synchronize()
    log(f"warmup {args.warmup} step(s) done")
    reset()
    ctx.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        render((args.warmup + k) * spp)
    ctx.synchronize()
    t_total = time.perf_counter() - t_start
    st = stats()
    log(f"timed {args.steps} step(s): {t_total:.3f}s wall, kernel {st.totalKernelMs:.1f} ms over {st.launches} launches")

    vertices = st.vertices
    value = vertices / t_total / 1e6
    kern_ms = st.totalKernelMs / max(st.launches, 1)
    bytes_per_launch = (16.0 * st.lookups + 4.0 * st.brickVisits + 16.0 * st.svLookups) / max(st.launches, 1) \
        + 32.0 * W * H
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    traffic, traffic_note = None, "no traffic file"
    if os.path.exists(args.traffic_file):
        try:
            tf = json.load(open(args.traffic_file))
            src = pt_sources_sha16()
            if tf.get("workload") != sc.name or tf.get("spp_per_step") != spp:
                traffic_note = "traffic file measured on another workload"
            elif tf.get("kernel_sources_sha16") != src:
                traffic_note = (f"stale: traffic measured on kernel sources {tf.get('kernel_sources_sha16')}, "
                                f"built sources {src}")
            else:
                traffic = tf.get("hbm_bytes_per_launch")
                traffic_note = f"rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel (sources {src})"
        except Exception as e:  # noqa: BLE001
            traffic_note = f"unreadable traffic file: {e}"

    cpu = None
    if args.via_batch:
        ctx.batch_clear()
        cloud, svp = sc.setup(ctx)
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, sc, svp, cfg, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpath-vertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{sc.name}: plain path tracer, {sc.fbm.size[0]}x{sc.fbm.size[1]}x{sc.fbm.size[2]} "
                         f"{'FP16' if sc.storage == S.CR_DTYPE_F16 else 'FP32'} fBm grid, {W}x{H}, "
                         f"{spp} spp per step, <= {sc.ext.bounceCap} bounces, albedo {sc.ext.albedo:.3f}, "
                         f"g {sc.params.phaseG:.2f}, "
                         f"{'ratio-tracked shadow rays' if sc.ext.neeMode == S.CR_NEE_RATIO_TRACKING else 'shadow-volume NEE'}, "
                         f"{'brick-majorant DDA' if sc.ext.trackingMode == S.CR_TRACK_BRICK_DDA else 'global majorant'}"),
            "spp_per_step": spp,
            "launch": "batched API, 1 job, HIP graph" if args.via_batch else "cr_pt_render",
            "image": [W, H],
            "grid": list(sc.fbm.size),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": "pt_kernel",
            "kernel_ms": round(kern_ms, 3),
            "algorithmic_bytes_per_launch": int(bytes_per_launch),
        },
        "cpu_baseline": cpu,
        "extra": {
            "paths": st.paths,
            "vertices": st.vertices,
            "lookups": st.lookups,
            "shadow_volume_lookups": st.svLookups,
            "brick_visits": st.brickVisits,
            "Mlookups_per_s": round(st.lookups / t_total / 1e6, 2),
            "Mpaths_per_s": round(st.paths / t_total / 1e6, 3),
            "spp_per_s_1080p": round(st.paths / t_total / (W * H), 3),
        },
    }
    ctx.close()
    if args.workload == "c2" and not args.no_c5:
        log(f"c5 (quick): {C5_QUICK}")
        t = time.perf_counter()
        out["extra"]["c5"] = run_c5(pkg, cfg, **C5_QUICK)
        out["extra"]["c5"]["wall_s"] = round(time.perf_counter() - t, 1)
    print(json.dumps(out), flush=True)


def rel_mse(img, ref, eps=1e-2):
    """mean over pixels and RGB of (I - R)^2 / (R^2 + eps)."""
    import numpy as np
    a = img[..., :3].astype(np.float64)
    r = ref[..., :3].astype(np.float64)
    return float(np.mean((a - r) ** 2 / (r * r + eps)))


def bench_mlp(args, pkg):
    """The learned network alone: --steps x 100 Adam steps of 8192 samples
    and --steps inference calls on 1920 x 1080 rows, random features and
    targets handed over from the host (cr_mlp_train_step / cr_mlp_forward).
    The wall times include the host copies; the kernels' own times come from
    a kernel trace of this command."""
    import ctypes as C
    import numpy as np
    ctx = pkg.Context(0)
    lib = ctx.lib
    ctx.check(lib.cr_mlp_init(ctx.h, 3), "init")
    B, rows = 8192, 1920 * 1080
    rng = np.random.default_rng(1)
    X = (rng.random((B, 224), dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)
    T = rng.random((B, 3), dtype=np.float32)
    loss = C.c_float()
    for _ in range(max(args.warmup, 1) * 20):
        ctx.check(lib.cr_mlp_train_step(ctx.h, X.ctypes.data, T.ctypes.data, B, 1e-3, C.byref(loss), None), "step")
    n_steps = max(args.steps, 1) * 100
    t = time.time()
    for _ in range(n_steps):
        ctx.check(lib.cr_mlp_train_step(ctx.h, X.ctypes.data, T.ctypes.data, B, 1e-3, C.byref(loss), None), "step")
    train_s = time.time() - t
    XI = np.ascontiguousarray(np.tile(X, (rows // B + 1, 1))[:rows])
    Y = np.zeros((rows, 3), np.float32)
    ctx.check(lib.cr_mlp_forward(ctx.h, XI.ctypes.data, rows, Y.ctypes.data), "forward")
    t = time.time()
    for _ in range(max(args.steps, 1)):
        ctx.check(lib.cr_mlp_forward(ctx.h, XI.ctypes.data, rows, Y.ctypes.data), "forward")
    infer_s = (time.time() - t) / max(args.steps, 1)
    flops = 2 * (224 * 128 + 3 * 128 * 128 + 128 * 16)
    out = {
        "metric": METRIC,
        "value": round(n_steps / train_s, 1),
        "unit": "Adam steps/s (8192 samples, host-driven: copies and loss readback included)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 MFMA, fp32 accumulate",
        "data": "synthetic",
        "config": {"workload": "mlp: 224-128-128-128-128-3 network, 8192-sample Adam steps, 1920x1080-row inference"},
        "steps": n_steps,
        "final_loss": round(float(loss.value), 6),
        "inference_rows": rows,
        "inference_call_s_with_copies": round(infer_s, 4),
        "inference_flop_per_row": flops,
        "outputs_finite": bool(np.isfinite(Y).all()),
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_ppm(args, pkg, cfg):
    """RenderTechniquePPM's frame loop (RenderTechniquePPM.cpp:229-312) on the
    c1 cloud at 1920x1080: per frame the radius update, 12 800 photon paths
    traced into a 100^3 photon grid (stable sort by cell, 32 kept per cell),
    and the ray-marched estimate.  The reference's photon-map scales are for
    its 1000-unit scene; here the march step and the initial radius are 1/100
    of the cloud box and the sky ray's step 1/10.  Wall time per stage with a
    synchronisation after each; --warmup frames, then --steps x 8 timed."""
    import ctypes as C
    import numpy as np
    S = pkg.structs
    ctx = pkg.Context(0)
    lib = ctx.lib
    sc = cfg.c1(width=1920, height=1080)
    sc.ext = cfg.render_ext(sky=cfg.SKY)
    cloud, svp = sc.setup(ctx)
    pm = S.PhotonMapProperties()
    lib.cr_photon_map_default(C.byref(pm))
    pm.voxelCount[:] = (100, 100, 100)
    lib.cr_photon_map_set_bounds(C.byref(pm), cloud.bounds[0], cloud.bounds[1])
    pm.stepSize = 0.01
    ctx.set_photon_map(pm)
    p = S.PpmParams()
    lib.cr_ppm_params_default(C.byref(p))
    p.ambientStep = 0.1
    n_timed = max(args.steps, 1) * 8
    frames = cfg.frames(max(args.warmup, 1) + n_timed, ref_frame_count=True)
    ctx.clear_image()
    radius = 0.01
    t_trace = t_est = 0.0
    deposits = 0
    for i, f in enumerate(frames):
        radius = float(lib.cr_ppm_update_radius(f.frameCount, S.CR_PPM_ALPHA, 0.01, radius))
        fr = S.PushConstants(0.0, f.seed, f.frameCount, radius, 0)
        t = time.perf_counter()
        n = ctx.ppm_trace(f.seed, p)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.ppm_estimate([fr], p)
        ctx.synchronize()
        t2 = time.perf_counter()
        if i >= len(frames) - n_timed:
            t_trace += t1 - t
            t_est += t2 - t1
            deposits += n
    img = ctx.download_image(1080, 1920)
    ms = 1e3 * (t_trace + t_est) / n_timed
    out = {
        "metric": METRIC,
        "value": round(1e3 / ms, 2),
        "unit": "frames/s (photon mapping at 1920x1080: trace + photon grid + estimate)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ppm: c1 cloud (128^3 FP32), 1920x1080, 12800 photon paths per frame into a 100^3 photon "
                               "grid (32 photons per cell), march step 0.01 and initial radius 0.01 of the unit box"},
        "frames": n_timed,
        "ms_per_frame": round(ms, 3),
        "trace_ms_per_frame": round(1e3 * t_trace / n_timed, 3),
        "estimate_ms_per_frame": round(1e3 * t_est / n_timed, 3),
        "deposits_per_frame": round(deposits / n_timed, 1),
        "final_radius": radius,
        "image_finite": bool(np.isfinite(img).all()),
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_c4(args, pkg, cfg):
    """64 independent jobs (uneven grids, images, frame counts, phase functions
    and suns): each rendered alone with cr_pt_render, then all of them in one
    batched persistent launch (HIP graph), default and autotuned shapes.
    Every batched image is compared with its job's single render."""
    import hashlib
    import numpy as np
    ctx = pkg.Context(0)
    jobs = cfg.c4()
    seq_s, digests, paths = 0.0, [], 0
    for k, sc in enumerate(jobs):
        sc.setup(ctx, build_sv=False)
        fr = cfg.frames(sc.spp)
        ctx.synchronize()
        t = time.perf_counter()
        ctx.pt_render(fr)
        ctx.synchronize()
        seq_s += time.perf_counter() - t
        digests.append(hashlib.sha256(ctx.download_image(sc.height, sc.width).tobytes()).hexdigest())
        paths += sc.width * sc.height * sc.spp
        ctx.batch_add(fr)
        if (k + 1) % 16 == 0:
            log(f"c4: {k + 1} jobs rendered one by one ({seq_s:.2f}s)")

    def batched():
        ctx.batch_clear_images()
        ctx.synchronize()
        t = time.perf_counter()
        ctx.batch_render()
        ctx.synchronize()
        dt = time.perf_counter() - t
        same = sum(hashlib.sha256(ctx.batch_download(k).tobytes()).hexdigest() == digests[k] for k in range(len(jobs)))
        return dt, same

    batched()  # records the graph
    t_default, same_default = batched()
    log(f"c4: batched (default shape) {t_default:.2f}s, {same_default}/{len(jobs)} images identical")
    t0 = time.perf_counter()
    tun, rep = ctx.batch_autotune(False)
    t_tune = time.perf_counter() - t0
    batched()
    t_tuned, same_tuned = batched()
    log(f"c4: batched (tuned {tun.blockThreads}x{tun.blocksPerCU}/CU, {tun.tileWidth}-wide tiles) {t_tuned:.2f}s")
    out = {
        "metric": METRIC,
        "value": round(paths / t_tuned / 1e6, 2),
        "unit": "Mpaths/s (batched, tuned)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"c4: {len(jobs)} jobs, FP16 grids {min(j.fbm.size[0] for j in jobs)}^3.."
                               f"{max(j.fbm.size[0] for j in jobs)}^3, images {min(j.width for j in jobs)}^2.."
                               f"{max(j.width for j in jobs)}^2, {min(j.spp for j in jobs)}..{max(j.spp for j in jobs)} spp, "
                               "brick-majorant DDA, ratio-tracked shadow rays"},
        "paths": paths,
        "one_by_one_s": round(seq_s, 3),
        "batched_default_s": round(t_default, 3),
        "batched_tuned_s": round(t_tuned, 3),
        "speedup_vs_one_by_one": round(seq_s / t_tuned, 3),
        "tuning": {"blockThreads": tun.blockThreads, "blocksPerCU": tun.blocksPerCU, "tileWidth": tun.tileWidth,
                   "candidates": rep.candidates, "autotune_s": round(t_tune, 2)},
        "images_identical_to_single_renders": {"default": same_default, "tuned": same_tuned, "jobs": len(jobs)},
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_c3(args, pkg, cfg):
    """Photon beams on the c1 volume: 2^22 sun-emitted beams traced once and
    then frozen, the LBVH build over them (30-bit Morton codes, 4 x 8-bit LSD
    radix sort, Karras hierarchy, AABB refit) timed over --steps builds after
    --warmup and checked bit-exact against the oracle at full size, then the
    1080p beam gather at 16 spp.  value = beams per second through the build."""
    import numpy as np
    S = pkg.structs
    O = None if args.no_cpu else __import__("cr_import").oracle()
    ctx = pkg.Context(0)
    sc = cfg.c1(width=1920, height=1080)
    cloud, _ = sc.setup(ctx, build_sv=False)
    # camera transmittance for the gather: a shadow volume along the camera
    # forward axis (main.cpp:226-229)
    ctx.set_shadow_volume(cfg.shadow_volume(cloud, tuple(sc.cam.forward), 256))
    ctx.sv_build()
    side = max(cloud.bounds[1][k] - cloud.bounds[0][k] for k in range(3))
    target = 1 << 22
    paths = 1 << 19
    t0 = time.perf_counter()
    while True:
        p = cfg.beam_params(paths, seed=38, radius=0.002 * side, albedo=0.9, light_dir=sc.light_dir,
                            max_interactions=64)
        nb, nd = ctx.trace_beams(p, target)
        ctx.synchronize()
        if nb >= target or paths >= (1 << 24):
            break
        paths *= 2
    t_trace = time.perf_counter() - t0
    log(f"c3: {nb} beams ({nd} beam data) from {paths} photon paths traced in {t_trace:.2f}s")
    stage_ms, wall = [], []
    for k in range(args.warmup + args.steps):
        ctx.synchronize()
        t = time.perf_counter()
        ctx.lbvh_build()
        ctx.synchronize()
        if k >= args.warmup:
            wall.append(time.perf_counter() - t)
            stage_ms.append(ctx.lbvh_timings())
    build_ms = float(np.mean([sum(v for key, v in st.items() if key != "pack") for st in stage_ms]))
    stages = {key: round(float(np.mean([st[key] for st in stage_ms])), 4) for key in stage_ms[0]}
    keys, perm, nodes = ctx.lbvh_download()
    parity = None
    cpu = None
    if O is not None:
        beams, _ = ctx.download_beams()
        t = time.perf_counter()
        rk, rp, rn = O.lbvh_build(beams, cloud.bounds[0], cloud.bounds[1])
        t_cpu = time.perf_counter() - t
        same_nodes = bytes(nodes)[: (2 * nb - 1) * 64] == bytes(rn)[: (2 * nb - 1) * 64]
        parity = {"keys": bool(np.array_equal(keys, rk)), "perm": bool(np.array_equal(perm, rp)),
                  "nodes": bool(same_nodes)}
        cpu = {"value": round(nb / t_cpu / 1e6, 3), "unit": "Mbeams/s (LBVH build)", "cores": 1, "kind": "port",
               "sample": f"oracle (oracle/liboracle.so, C++ -O3) building the LBVH of the same {nb} beams "
                         f"(Morton codes, radix sort, Karras hierarchy, refit): {t_cpu:.2f}s"}
        log(f"c3: oracle build {t_cpu:.2f}s; parity {parity}")
    # gather: 1080p x --c3-gather-spp frames in one call
    frames = cfg.frames(max(1, args.c3_gather_spp))
    ext = S.GatherExt(0, 0, 0.9, 0)
    gather = None
    if args.c3_gather_spp > 0:
        ctx.clear_image()
        ctx.gather(frames[:1], ext)      # warm-up frame
        ctx.synchronize()
        ctx.clear_image()
        g0 = ctx.gather_stats()
        ctx.synchronize()
        t = time.perf_counter()
        ctx.gather(frames, ext)
        ctx.synchronize()
        t_gather = time.perf_counter() - t
        g1 = ctx.gather_stats()
        rays, nodes_v, tested = g1.rays - g0.rays, g1.nodesVisited - g0.nodesVisited, g1.beamsTested - g0.beamsTested
        img = ctx.download_image(sc.height, sc.width)
        gather = {"spp": args.c3_gather_spp, "seconds": round(t_gather, 4), "kernel_ms": round(g1.lastKernelMs, 3),
                  "Mrays_per_s": round(rays / (g1.lastKernelMs * 1e-3) / 1e6, 2),
                  "nodes_per_ray": round(nodes_v / max(rays, 1), 2),
                  "beams_tested_per_ray": round(tested / max(rays, 1), 2),
                  "finite": bool(np.isfinite(img).all())}
    # algorithmic bytes of the build per beam: Morton 24 B read + 8 B written, sort 4 x (4 + 8 + 8) B,
    # hierarchy ~20 B, refit 28 B read + 48 B written
    alg = 208.0 * nb
    achieved = alg / (build_ms * 1e-3) / 1e9
    # HBM bytes of one build, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this workload
    # (keyed to the kernel sources it was measured on: a stale file is not used)
    traffic = None
    tsrc = "absent"
    tfile = os.path.join(ROOT, "profiles", "lbvh_c3_traffic.json")
    if os.path.exists(tfile):
        try:
            tf = json.load(open(tfile))
            built = sources_sha16(LBVH_SOURCES)
            if tf.get("workload") == "c3" and tf.get("beams") == nb and tf.get("kernel_sources_sha16") == built:
                traffic = tf.get("hbm_bytes_per_build")
                tsrc = f"profiles/lbvh_c3_traffic.json (kernel sources {built})"
            else:
                tsrc = f"stale: traffic measured on kernel sources {tf.get('kernel_sources_sha16')}, built sources {built}"
        except Exception:
            traffic = None
    out = {
        "metric": METRIC,
        "value": round(nb / (build_ms * 1e-3) / 1e6, 2),
        "unit": "Mbeams/s (LBVH build: Morton, sort, hierarchy, refit; HIP events)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32 beams, u32 keys",
        "data": "synthetic",
        "config": {"workload": f"c3: {nb} sun-emitted beams traced on the c1 volume (seed 38, radius "
                               f"0.002 x box side), LBVH build, 1920x1080 beam gather at "
                               f"{args.c3_gather_spp} spp"},
        "build_ms": round(build_ms, 4),
        "build_wall_ms": round(1e3 * float(np.mean(wall)), 3),
        "stages_ms": stages,
        "trace_s": round(t_trace, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": traffic, "traffic_source": tsrc,
                     "kernel": "lbvh build (5 stages)",
                     "algorithmic_bytes_per_build": alg},
        "parity_vs_oracle": parity,
        "cpu_baseline": cpu,
        "gather": gather,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


# The learned-vs-plain measurement carried by the default line: 400 training
# clouds with 20 000 points of 256 continuations each, so the block takes about
# two minutes (`--workload c5` runs the full-size set: 80 000 points of 1024
# continuations).  Cloud count matters more than points per cloud: the
# network's error on the held-out cloud is a generalisation error.  120 000
# Adam steps take about 5 s; with 40 000 the learned render's error floor sat
# near 1e-3 and the crossing swung with the training draw (27 or 79 spp).
C5_QUICK = dict(clouds=400, points=20000, target_spp=256, train_steps=120000, collect_frames=8, tracking="dda",
                batch=8192, ema=0.0, ref_spp=16384, W=1920, H=1080)
C5_REF_SEED0 = 40000  # reference frame seeds 40000, 40003, ...: above MSVC rand()'s 0..32767
# Frames per cr_learn_render call: one call renders up to this many frames
# (record pass, compaction, features, network, composition once for all of
# them).  A 1080p frame holds at most ~1 GB of records, features and outputs.
# Measured on 32 spp: 29.8 ms in one call, 36.5 ms in two, 44.4 ms in four.
C5_LEARN_CALL_FRAMES = 64


def run_c5(pkg, cfg, clouds, points, target_spp, train_steps, collect_frames, tracking, batch, ema, ref_spp,
           W, H, plain_max_spp=4096):
    """Learned in-scattering vs the plain tracer: train on samples the plain
    tracer generates over seeded 128^3 clouds of c1 statistics, then time both
    renderers at W x H on a held-out cloud until relMSE <= 1e-3 against a
    converged plain render.  The converged render's frame seeds are disjoint
    from the evaluated renders' (the seed times the pixel id starts a pixel's
    stream, so a shared seed would repeat a reference sample exactly)."""
    import math
    import numpy as np
    ctx = pkg.Context(0)
    rng = np.random.default_rng(5)
    S = pkg.structs

    def cloud_scene(seed, light, w=256, h=256):
        sc = cfg.c1(width=w, height=h)
        sc.fbm.seed = seed
        sc.light_dir = light
        if tracking == "dda":
            sc.ext.trackingMode = S.CR_TRACK_BRICK_DDA
            sc.ext.neeMode = S.CR_NEE_RATIO_TRACKING
        return sc

    def sun():
        ct = float(rng.uniform(0.2, 1.0))
        ph = float(rng.uniform(0.0, 2.0 * math.pi))
        st = math.sqrt(1.0 - ct * ct)
        return (st * math.cos(ph), -ct, st * math.sin(ph))

    # ---- training data from the plain tracer
    t0 = time.perf_counter()
    ctx.learn_dataset_clear()
    for k in range(clouds):
        sc = cloud_scene(1000 + k, sun())
        sc.setup(ctx)
        ctx.learn_collect(cfg.frames(collect_frames, srand_seed=1 + k), target_spp, points)
        if (k + 1) % 20 == 0:
            log(f"c5: {k + 1} clouds collected, {ctx.learn_dataset_size()} points, {time.perf_counter() - t0:.1f}s")
    ctx.synchronize()
    t_collect = time.perf_counter() - t0
    n_points = ctx.learn_dataset_size()
    t0 = time.perf_counter()
    ctx.mlp_init(11)
    ctx.mlp_set_ema(ema)
    loss = None
    for s in range(0, train_steps, 1000):
        lr = 1e-3 if s < train_steps // 3 else (3e-4 if s < 2 * train_steps // 3 else 1e-4)
        loss = ctx.learn_train(min(1000, train_steps - s), batch, lr, seed=3)
        if (s // 1000) % 10 == 9:
            log(f"c5: {s + 1000} training steps, loss {loss:.4g}")
    ctx.synchronize()
    t_train = time.perf_counter() - t0
    log(f"c5: {n_points} points collected in {t_collect:.1f}s, {train_steps} steps trained in {t_train:.1f}s")

    # ---- held-out cloud
    sc = cloud_scene(7, (1.0, -1.0, 0.0), W, H)
    sc.setup(ctx)
    t0 = time.perf_counter()
    ref_frames = [S.PushConstants(0.0, C5_REF_SEED0 + 3 * i, i, 0.0, 0) for i in range(ref_spp)]
    for i in range(0, ref_spp, 256):
        ctx.pt_render(ref_frames[i:i + 256])
        ctx.synchronize()
        if (i // 256) % 16 == 15:
            log(f"c5: reference {i + 256} spp, {time.perf_counter() - t0:.1f}s")
    ref = ctx.download_image(H, W)
    t_ref = time.perf_counter() - t0
    log(f"c5: reference {ref_spp} spp in {t_ref:.1f}s")

    def progressive(render, max_spp, chunk):
        """(spp, seconds, relMSE) after each step of a 1, 2, 3, 4, 5, 6, 7, 9, ... (x1.2) schedule;
        stops at relMSE <= 1e-3, or when the error has stopped falling."""
        ctx.clear_image()
        ctx.synchronize()
        out, done, secs = [], 0, 0.0
        frames = cfg.frames(max_spp)
        sched = sorted({max(1, int(round(1.2 ** i))) for i in range(0, 64)})
        for target in sched:
            if target > max_spp:
                break
            t = time.perf_counter()
            for i in range(done, target, chunk):
                render(frames[i:min(target, i + chunk)])
            ctx.synchronize()
            secs += time.perf_counter() - t
            done = target
            e = rel_mse(ctx.download_image(H, W), ref)
            out.append((done, secs, e))
            if e <= 1e-3:
                break
            if len(out) >= 8 and e > 0.97 * out[-5][2]:
                break  # a floor: the estimator's bias, not its noise
        return out

    plain = progressive(ctx.pt_render, plain_max_spp, 256)
    learned = progressive(ctx.learn_render, 1024, C5_LEARN_CALL_FRAMES)

    def first_below(curve):
        for spp, secs, e in curve:
            if e <= 1e-3:
                return spp
        return None

    def timed(render, spp, chunk):
        """Wall time of one render of exactly spp frames (the quality the
        progressive pass found sufficient), as a renderer that knows its
        target would run it; buffers are warm."""
        if spp is None:
            return None, None
        frames = cfg.frames(spp)
        best = None
        for _ in range(2):
            ctx.clear_image()
            ctx.synchronize()
            t = time.perf_counter()
            for i in range(0, spp, chunk):
                render(frames[i:i + chunk])
            ctx.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return best, rel_mse(ctx.download_image(H, W), ref)

    p_spp, l_spp = first_below(plain), first_below(learned)
    p_s, p_e = timed(ctx.pt_render, p_spp, 256)
    l_s, l_e = timed(ctx.learn_render, l_spp, C5_LEARN_CALL_FRAMES)
    log(f"c5: plain {p_spp} spp in {p_s}s (relMSE {p_e}), learned {l_spp} spp in {l_s}s (relMSE {l_e})")
    ctx.close()
    return {
        "speedup": round(p_s / l_s, 3) if (p_s and l_s) else None,
        "unit": "x (plain time / learned time to relMSE <= 1e-3)",
        "workload": f"c5: learned in-scattering, {clouds} seeded 128^3 clouds of c1 statistics, {points} points and "
                    f"{target_spp} continuations per point, evaluated at {W}x{H} on a held-out cloud against a "
                    f"{ref_spp}-spp plain render (reference seeds disjoint from the evaluated renders')",
        "tracking": "brick-majorant DDA + ratio-tracked shadow rays" if tracking == "dda"
        else "global majorant + shadow-volume NEE (the reference's)",
        "plain": {"spp": p_spp, "seconds": p_s, "relMSE": p_e,
                  "progressive_curve_spp_s_relmse": [[a, round(b, 4), round(c, 6)] for a, b, c in plain]},
        "learned": {"spp": l_spp, "seconds": l_s, "relMSE": l_e,
                    "progressive_curve_spp_s_relmse": [[a, round(b, 4), round(c, 6)] for a, b, c in learned]},
        "training": {"clouds": clouds, "points": n_points, "collect_s": round(t_collect, 2),
                     "train_s": round(t_train, 2), "steps": train_steps, "batch": batch, "ema_decay": ema,
                     "final_loss": loss},
        "reference_s": round(t_ref, 2),
    }


def bench_c5(args, pkg, cfg):
    r = run_c5(pkg, cfg, args.c5_clouds, args.c5_points, args.c5_target_spp, args.c5_train_steps,
               args.c5_collect_frames, args.c5_tracking, args.c5_batch, args.c5_ema, args.c5_ref_spp,
               args.c5_width, args.c5_height)
    out = {
        "metric": METRIC,
        "value": r["speedup"],
        "unit": r["unit"],
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "bf16 network, f32 tracer",
        "data": "synthetic",
        "config": {"workload": r["workload"], "tracking": r["tracking"]},
    }
    out.update({k: r[k] for k in ("plain", "learned", "training", "reference_s")})
    print(json.dumps(out), flush=True)


# Host cores the GPU pool grants one GPU's job (its CPU share); the process may
# see more in nproc or its affinity mask, but those belong to other jobs.
HOST_CORE_SHARE = 16


def host_cpu():
    """The host's CPU model, core counts and the cores a CPU baseline uses."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except Exception:
        info["affinity"] = None
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(period)
    except Exception:
        pass
    info["cgroup_cpus"] = quota
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    info["model"] = model
    n = info["affinity"] or info["nproc"] or 1
    if quota:
        n = min(n, max(1, int(quota)))
    info["threads"] = max(1, min(n, HOST_CORE_SHARE))
    return info


def native_oracle():
    """Builds the oracle's sources with -march=native for this host (the
    baseline's build; the shipped oracle is portable x86-64-v2)."""
    import subprocess
    d = os.path.join(ROOT, "oracle")
    try:
        r = subprocess.run(["make", "-B", "-C", d, "native"], capture_output=True, timeout=300)
        if r.returncode == 0:
            return os.path.join(d, "_native", "liboracle.so"), "-O3 -march=native -ffp-contract=off -fopenmp"
        log(f"native oracle build failed: {r.stderr.decode()[-400:]}")
    except Exception as e:  # noqa: BLE001
        log(f"native oracle build failed: {e}")
    return None, "-O3 -march=x86-64-v2 -ffp-contract=off -fopenmp (the native build failed)"


def cpu_baseline(ctx, sc, svp, cfg, seconds):
    """The oracle, built -march=native on this host and run on its cores, over
    a centred pixel band of the same render (same pixel ids and RNG streams),
    sized to about `seconds`."""
    import numpy as np
    import cr_import
    lib, flags = native_oracle()
    O = cr_import.oracle(lib)
    cpu = host_cpu()
    threads = cpu["threads"]
    grid = ctx.download_grid()
    if sc.storage == 1:
        grid = grid.astype(np.float16)
    s = O.scene(grid, ctx.cloud(), sc.cam, sc.params, svp, sc.ext, sv=ctx.sv_download(), majorants=ctx.majorants())
    W, H = sc.width, sc.height
    # calibrate on one frame of a centred band, then time a block and frame
    # count sized to the target
    w, h = W, max(1, min(H, (1 << 21) // W))
    t = time.perf_counter()
    _, cnt = O.pt_render_region(s, cfg.frames(1), 0, (H - h) // 2, w, h, threads)
    dt = time.perf_counter() - t
    per_px = dt / (w * h)
    npx = int(max(256, seconds / max(per_px, 1e-9)))
    nfr = int(max(1, min(sc.spp, npx // (W * H))))
    npx = min(npx, W * H)
    h = max(1, min(H, npx // W))
    w = W if h > 1 else min(W, npx)
    x0, y0 = (W - w) // 2, (H - h) // 2
    fr = cfg.frames(nfr)
    t = time.perf_counter()
    _, cnt = O.pt_render_region(s, fr, x0, y0, w, h, threads)
    dt = time.perf_counter() - t
    log(f"cpu baseline: {w}x{h} px x {nfr} spp in {dt:.1f}s on {threads} threads ({cpu['model']})")
    return {
        "value": round(cnt.vertices / dt / 1e6, 4),
        "unit": "Mpath-vertices/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle (oracle/*.cpp, C++ {flags}, OpenMP dynamic rows, {threads} threads) rendering the "
                  f"centred {w}x{h} pixels of the same {W}x{H} image, {nfr} spp, same pixel ids and RNG streams; "
                  f"{cnt.vertices} vertices in {dt:.1f}s",
        "host": {"model": cpu["model"], "nproc": cpu["nproc"], "affinity": cpu["affinity"],
                 "cgroup_cpus": cpu["cgroup_cpus"], "threads_used": threads,
                 "note": f"threads = min(affinity, cgroup quota, {HOST_CORE_SHARE}): the GPU pool's host-core share "
                         "for one GPU"},
    }


if __name__ == "__main__":
    main()
