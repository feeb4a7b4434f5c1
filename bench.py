#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Default workload (the metric's single-GPU configuration): the ~0.92 M-triangle
procedural city at 1920x1080, PT_MIS estimator (GGX bounded-VNDF + Lambert, NEE to
emissive triangles, MIS), 1 spp, 8 bounces, FP32, then reprojection and the bilateral
denoiser. Scene and BVH are resident in HBM before the timed region; a step renders one
frame. ``value`` is the whole-job ray throughput (primary + extension + shadow rays that
the trace kernels processed, divided by wall time). The timed frames run as two path
pipelines on two streams whose launches overlap; the roofline object describes the
traversal kernel k_trace, timed with HIP events around every launch of a few more frames
rendered as one pipeline right after the timed region; ``cpu_baseline`` is the reference's
tinybvh BVH::Intersect on host cores.

Other configurations (``--config``): ``dynamic`` (4096 moving instanced bricks,
TLAS refit every frame, 4 bounces, reprojection + bilateral, 120 frames) and ``cornell``
(256x256, 16 frames x 1 spp, 4 bounces), ``diff`` (path-replay adjoint
at 512x512, 4 spp, with its finite-difference and 4096-spp reference-gradient checks) and
``denoise`` (learned vs bilateral denoiser on 1-spp 1080p frames of the city and
the Cornell box), ``probes`` (the ambient-cube probe grid: voxelize and bake over the
city), ``output`` (SuperResolution 1 render, HDR compose, linear and FSR EASU resolves), ``copy`` (a
256 MiB device copy: the measured HBM ceiling and the counter calibration kernel). ``--relmse N``
adds the relMSE of the measured frame against an N-spp FP64-accumulated render of the
same frame (``denoise`` defaults to N = 4096).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_RAY = 52  # 32 B ray in (o, tmin, d, tmax) + 20 B hit out (t, u, v, prim, inst)
METRIC = "Mrays/s & ms/frame @1080p; relMSE vs 4096-spp, bilateral vs learned denoise"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gt_tonemap(x):
    """GranTurismoTonemapper (Const_Func.slang:100-121), numpy."""
    import numpy as np
    P, a, m, l, c, b = 1.0, 0.7, 0.22, 0.4, 1.33, 0.0
    l0 = (P - m) * l / a
    S0, S1 = m + l0, m + a * l0
    C2 = a * P / (P - S1)
    x = np.maximum(np.asarray(x, np.float64), 0.0)
    T = m * (x / m) ** c + b
    L = m + a * (x - m)
    S = P - (P - S1) * 2.71828 ** (-(C2 * (x - S0) / P))
    t = np.clip(x / m, 0.0, 1.0)
    w0 = 1.0 - t * t * (3.0 - 2.0 * t)
    w2 = (x >= m + l0).astype(np.float64)
    w1 = 1.0 - w0 - w2
    return T * w0 + L * w1 + S * w2


def relmse(x, ref):
    import numpy as np
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def relmse_stats(x, ref):
    """relMSE with non-finite pixels shown as black, relMSE over the finite pixels only, and
    the number of pixels with a non-finite channel (the bilateral filter divides by a weight
    sum that is zero when every tap fails the luminance test, as the reference's does)."""
    import numpy as np
    x = np.asarray(x, np.float64)
    bad = ~np.all(np.isfinite(x), axis=-1)
    xs = np.where(np.isfinite(x), x, 0.0)
    ok = ~bad
    return relmse(xs, ref), relmse(x[ok], np.asarray(ref)[ok]), int(bad.sum())


def host_cpu():
    """Host CPU model (lscpu's "Model name", else /proc/cpuinfo), the CPUs this process may
    run on, and the threads the CPU baselines use: the CPU share of one GPU on the GPU hosts
    (OMP_NUM_THREADS, 16 there), at most the CPUs available."""
    import subprocess
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), "")
    except (OSError, subprocess.SubprocessError):
        pass
    if not model and Path("/proc/cpuinfo").exists():
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    visible = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return {"cpu_model": model, "host_cpus_available": visible, "threads": max(1, min(visible, share))}


def cpu_baseline(scene, params, W, H):
    """tinybvh BVH::Intersect (the reference's CPU ray cast) on the camera rays and the
    first-bounce diffuse rays of the same frame, on the CPU share of one GPU, 64K-ray tasks."""
    import numpy as np

    from oracle import pyoracle as po
    if not po.ref_available():
        return None
    hc = host_cpu()
    threads = hc["threads"]
    tris = po.scene_triangles(scene.meshes[0])
    t0 = time.time()
    bvh = po.TinyBVH(tris)
    build_ms = (time.time() - t0) * 1e3
    rays = po.primary_rays(params, W, H)
    hits, ms_primary = bvh.intersect(rays, threads=threads)
    ok = hits[:, 3].view(np.uint32) != 0xFFFFFFFF
    t = np.where(ok, hits[:, 0], 0.0)[:, None]
    rng = np.random.default_rng(1)
    d = rng.normal(size=(rays.shape[0], 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = np.zeros_like(rays)
    b[:, :3] = rays[:, :3] + rays[:, 4:7] * t * np.float32(0.999)
    b[:, 4:7] = d
    b[:, 7] = 1000.0
    b = b[ok]
    _, ms_bounce = bvh.intersect(b, threads=threads)
    n = rays.shape[0] + b.shape[0]
    value = n / ((ms_primary + ms_bounce) * 1e-3) / 1e6
    return {"value": round(value, 2), "unit": "Mrays/s", "cores": threads,
            "kind": "reference",
            "sample": f"tinybvh 1.3.8 BVH::Intersect (binned SAH, single ray) on {rays.shape[0]} camera + "
                      f"{b.shape[0]} first-bounce rays of the same 1080p city frame; BLAS build {build_ms:.0f} ms "
                      f"single-threaded",
            "cpu_model": hc["cpu_model"], "host_cpus_available": hc["host_cpus_available"],
            "mrays_per_thread": round(value / threads, 3),
            "all_available_cpus_extrapolated": round(value / threads * hc["host_cpus_available"], 1)}


BYTES_PER_VERTEX = 244  # path state r+w 2 x 48 B, material 64 B, 3 indices 12 B, 3 vertices 3 x 24 B
BYTES_PER_PIXEL = 32    # output targets written per pixel and frame


def frame_bytes_roofline(s1, s0, steps, wall, pixels, bvh):
    """The frame's compulsory bytes (52 B per traced ray, 244 B per shaded path vertex =
    closest-hit ray, 32 B of targets per pixel, one read of the BVH and triangles) over the
    frame's wall time, against the HBM peak."""
    rays = (s1.rays_traced - s0.rays_traced) / steps
    vertices = (s1.extension_rays - s0.extension_rays) / steps
    bvh_bytes = bvh.blas_bytes + bvh.tlas_bytes
    total = BYTES_PER_RAY * rays + BYTES_PER_VERTEX * vertices + BYTES_PER_PIXEL * pixels + bvh_bytes
    t = wall / steps
    achieved = total / t / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "bytes_per_frame": int(total),
            "terms_per_frame": {"rays": int(rays), "path_vertices": int(vertices), "pixels": int(pixels),
                                "bvh_and_triangle_bytes": int(bvh_bytes)},
            "definition": "(52 B x rays + 244 B x path vertices + 32 B x pixels + BVH bytes) / frame wall time"}


def reference_frame(ctx, p, spp, first_frame=100000):
    """FP64-accumulated spp-frame PT_MIS render of the current scene state (no motion)."""
    import numpy as np

    import gknextengine_amd as gk
    q = gk.FrameParams.from_buffer_copy(p)
    q.estimator = gk.ESTIMATOR_PT_MIS
    q.flags = gk.FRAME_ACCUM_FP64 | gk.FRAME_SKIP_POST
    q.prev_view_projection[:] = q.view_projection[:]
    q.prev_view_projection_unjit[:] = q.view_projection_unjit[:]
    ctx.reset()
    for f in range(spp):
        q.total_frames = first_frame + f
        ctx.frame(q)
        ctx.render()
        if f % 256 == 255:
            ctx.sync()
    r = ctx.read("RADIANCE64")
    return r[..., :3] / np.maximum(r[..., 3:4], 1.0)


def run_city(args):
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    W, H = args.width, args.height
    log("building city scene")
    scene = gk.procgen_city(64, 4, 256, 64, 2024)
    ntris = scene.meshes[0]["indices"].shape[0]
    if args.mixed_materials:
        # same geometry, materials spread over four models (dielectric shading differs)
        import numpy as np
        rng = np.random.default_rng(5)
        models = [gk.MAT_LAMBERTIAN, gk.MAT_METALLIC, gk.MAT_DIELECTRIC, gk.MAT_MIXTURE]
        for m in scene.materials:
            if m.material_model != gk.MAT_DIFFUSE_LIGHT:
                m.material_model = int(models[rng.integers(4)])
    models_used = {m.material_model for m in scene.materials if m.material_model != gk.MAT_DIFFUSE_LIGHT}
    ctx = gk.Context(W, H)
    ctx.upload(scene)
    bvh = ctx.bvh_stats()
    cam = scenes.CITY_CAM
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = p.max_bounces = args.bounces
    p.samples = 1
    p.global_seed = 1
    p.denoiser = gk.DENOISER_BILATERAL if args.denoiser == "bilateral" else gk.DENOISER_NONE
    if args.sort == "off":
        p.flags |= gk.FRAME_NO_MATERIAL_SORT
    elif args.sort == "on":
        p.flags |= gk.FRAME_MATERIAL_SORT
    sorted_shading = args.sort == "on"
    if args.pipelines == 1:
        p.flags |= gk.FRAME_SINGLE_PIPELINE
    frame = 0

    def step():
        nonlocal frame
        p.total_frames = frame  # new per-pixel pcg4d seeds every frame
        ctx.frame(p)
        ctx.render()
        frame += 1

    for _ in range(args.warmup):
        step()
    ctx.sync()
    s0 = ctx.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s1 = ctx.stats()
    wall = t1 - t0
    rays = s1.rays_traced - s0.rays_traced
    ms_trace_split = s1.ms_trace - s0.ms_trace
    value = rays / wall / 1e6
    frame_roofline = frame_bytes_roofline(s1, s0, args.steps, wall, W * H, bvh)
    # The timed frames run as two path pipelines whose trace launches overlap, so a launch's
    # HIP-event duration there is not the kernel's own. The traversal roofline is taken from
    # the same frames rendered as one pipeline (GK_FRAME_SINGLE_PIPELINE) after the timed
    # region, also bracketed by HIP events around every launch.
    kframes = max(2, min(args.steps, 5))
    p.flags |= gk.FRAME_SINGLE_PIPELINE
    k0 = ctx.stats()
    for _ in range(kframes):
        step()
    ctx.sync()
    k1 = ctx.stats()
    if args.pipelines != 1:
        p.flags &= ~gk.FRAME_SINGLE_PIPELINE
    krays = k1.rays_traced - k0.rays_traced
    ms_trace = (k1.ms_trace - k0.ms_trace) / kframes * args.steps
    achieved = BYTES_PER_RAY * krays / ((k1.ms_trace - k0.ms_trace) * 1e-3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "trace_city1080p_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
    extra = {}
    if args.relmse:
        import numpy as np
        # one 1-spp frame denoised on its own (TemporalFrames = 1: the reprojection keeps
        # only the current sample) against an N-spp FP64 render
        ctx.reset()
        p.total_frames = 7
        p.temporal_frames = 1
        p.flags |= gk.FRAME_LINEAR_OUTPUT
        ctx.frame(p)
        ctx.render()
        raw = ctx.read("RADIANCE")[..., :3]
        den = ctx.read("DENOISED")[..., :3].astype(np.float64)
        ref = reference_frame(ctx, p, args.relmse)
        sc = p.paper_white_nit / 40000.0
        b, b_finite, nbad = relmse_stats(den, ref)
        extra = {"relmse_1spp": relmse(raw, ref), "relmse_bilateral": b, "relmse_bilateral_finite_only": b_finite,
                 "nonfinite_pixels_bilateral": nbad, "reference_spp": args.relmse,
                 "relmse_tonemapped_1spp": relmse(gt_tonemap(raw * sc), gt_tonemap(ref * sc))}
    cpu = None if args.no_cpu_baseline else cpu_baseline(scene, p, W, H)
    ctx.close()
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "procedural city, PT_MIS 1 spp 8 bounces + reproject + bilateral",
                   "resolution": [W, H], "triangles": int(ntris), "bounces": args.bounces, "spp": 1,
                   "estimator": "PT_MIS", "material_sort": sorted_shading,
                   "material_models": len(models_used),
                   "rays_per_frame": round(rays / args.steps), "trace_ms_per_frame": round(ms_trace / args.steps, 3),
                   "path_pipelines": args.pipelines,
                   "trace_launch_ms_per_frame_timed": round(ms_trace_split / args.steps, 3),
                   "reproject_ms": round(s1.ms_reproject, 3), "denoise_ms": round(s1.ms_denoise, 3),
                   "blas_build_ms": round(bvh.build_ms, 2), "bvh_sah_cost": round(bvh.sah_cost, 2),
                   "blas_nodes": bvh.blas_nodes, **extra},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_trace (primary rays; per-bounce extension + shadow rays)",
                     "bytes_per_ray": BYTES_PER_RAY,
                     "timed_on": f"{kframes} frames after the timed region rendered as one pipeline"},
        "frame_roofline": frame_roofline,
        "cpu_baseline": cpu,
    }


def dynamic_cpu_baseline(ctx, b, p, W, H, worlds):
    """The reference's CPU side of the dynamic scene on host cores: tinybvh's TLAS rebuild
    per frame over the 4096 brick instances (single thread, as UpdateBVH), its TLAS
    Intersect on the frame's camera rays (all threads), and the restated reprojection +
    bilateral filter on the last frame's own 1080p inputs (one thread)."""
    import numpy as np

    from oracle import pyoracle as po
    if not po.ref_available():
        return None
    hc = host_cpu()
    threads = hc["threads"]
    bb = po.TinyBVH(po.scene_triangles(b.scene.meshes[0]))
    tl = po.TinyTLAS([bb], worlds[0].reshape(-1, 16), np.zeros(b.n, np.uint32))
    rebuild_ms = [tl.rebuild(worlds[f].reshape(-1, 16)) for f in range(1, 17)]
    rays = po.primary_rays(p, W, H)
    _, _, ms_trace = tl.intersect(rays, threads=threads)
    h2f = lambda a: np.asarray(a).astype(np.float32)
    src = [h2f(ctx.read(k)) for k in ("SINGLE_DIFFUSE", "SINGLE_SPECULAR", "ALBEDO", "NORMAL")]
    acc = [h2f(ctx.read(k)) for k in ("ACCUM_DIFFUSE", "ACCUM_SPECULAR", "ACCUM_ALBEDO")]
    t0 = time.perf_counter()
    out = po.reproject(*src, *acc, ctx.read("OBJECTID_0"), ctx.read("OBJECTID_1"), ctx.read("MOTIONVECTOR"),
                       ctx.read("MOTIONMOMENT"), acc, temporal_frames=16)
    t1 = time.perf_counter()
    po.jbf(*out)
    t2 = time.perf_counter()
    return {"value": round(rays.shape[0] / (ms_trace * 1e-3) / 1e6, 2), "unit": "Mrays/s", "cores": threads,
            "kind": "reference",
            "sample": f"tinybvh 1.3.8 TLAS (4096 instances) Intersect on the {rays.shape[0]} camera rays of the last "
                      f"frame; TLAS rebuild per frame (1 thread, 16 frames) and the restated reprojection + bilateral "
                      f"of that frame (1 thread)",
            "tlas_rebuild_ms": round(float(np.median(rebuild_ms)), 3),
            "reproject_ms": round((t1 - t0) * 1e3, 1), "bilateral_ms": round((t2 - t1) * 1e3, 1),
            "cpu_model": hc["cpu_model"], "host_cpus_available": hc["host_cpus_available"]}


def run_dynamic(args):
    """120 frames of 4096 moving bricks at 1080p, 4 bounces, TLAS refit per
    frame from device-resident matrices, reprojection + bilateral."""
    import numpy as np
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    W, H = args.width, args.height
    b = scenes.Bricks()
    nframes = args.frames
    worlds = np.stack([b.worlds(f / 60.0) for f in range(nframes + 1)])           # (F, n, 4, 4)
    d_worlds = torch.from_numpy(np.ascontiguousarray(np.transpose(worlds, (0, 1, 3, 2)))).cuda()  # column-major
    ctx = gk.Context(W, H)
    ctx.upload(b.scene)
    cam = scenes.BRICKS_CAM
    prev = None
    times = []
    stats0 = None
    p = None
    rebuild = args.tlas == "rebuild"
    trace_ms, sah = [], []
    for f in range(nframes + 1):
        p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"], frame=f, prev=prev)
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.bounces = p.max_bounces = 4
        p.samples = 1
        p.has_sky = 1
        p.sky_color[:] = [0.6, 0.7, 0.9, 1.0]
        p.temporal_frames = 16
        p.denoiser = gk.DENOISER_BILATERAL
        if rebuild:
            p.flags |= gk.FRAME_TLAS_REBUILD
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if f > 0:
            gk.check(gk.lib().gk_instances_update_device(ctx.h, d_worlds[f].data_ptr(), 0, b.n), "update")
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        times.append(time.perf_counter() - t0)
        if f == 8:
            stats0 = ctx.stats()
        if args.per_frame:  # outside the frame's timed interval
            st = ctx.stats()
            trace_ms.append(st.ms_trace)
            sah.append(ctx.bvh_stats().tlas_sah_cost)
        prev = p
    s1 = ctx.stats()
    rays = s1.rays_traced - stats0.rays_traced
    ms = np.array(times[9:]) * 1e3
    if args.per_frame:
        tr = np.diff(np.array(trace_ms))
        extra_pf = {"per_frame_trace_ms": [round(float(x), 3) for x in tr],
                    "per_frame_tlas_sah": [round(float(x), 3) for x in sah],
                    "per_frame_ms": [round(float(x) * 1e3, 3) for x in times]}
    else:
        extra_pf = {}
    cpu = None if args.no_cpu_baseline else dynamic_cpu_baseline(ctx, b, p, W, H, worlds)
    extra = {}
    if args.relmse:
        den = ctx.read("DENOISED")[..., :3].astype(np.float64)
        ref = reference_frame(ctx, p, args.relmse)
        sc = p.paper_white_nit / 40000.0
        b, b_finite, nbad = relmse_stats(den, gt_tonemap(ref * sc))
        extra = {"relmse_tonemapped_bilateral_frame120": b, "relmse_tonemapped_bilateral_frame120_finite_only": b_finite,
                 "nonfinite_pixels_bilateral": nbad, "reference_spp": args.relmse}
    ctx.close()
    return {"metric": METRIC, "value": round(rays / (ms.sum() * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
            "steps": len(ms), "warmup": 9, "ms_per_step": round(float(np.median(ms)), 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"dynamic bricks: 4096 instances x 24 tris, TLAS "
                                   f"{'rebuilt' if rebuild else 'refitted'} per frame, "
                                   "PT_MIS 1 spp 4 bounces, reproject + bilateral", "resolution": [W, H],
                       "tlas_update": args.tlas,
                       "ms_frame_mean": round(float(ms.mean()), 3), "ms_frame_p95": round(float(np.percentile(ms, 95)), 3),
                       "ms_frame_max": round(float(ms.max()), 3),
                       "tlas_ms": round(s1.ms_tlas, 3), "reproject_ms": round(s1.ms_reproject, 3),
                       "denoise_ms": round(s1.ms_denoise, 3), **extra, **extra_pf},
            "cpu_baseline": cpu}


def run_cornell(args):
    """The 256x256 Cornell box, 16 frames x 1 spp, 4 bounces."""
    import numpy as np
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    W = H = 256
    s = gk.procgen_cornell()
    out = {}
    for est, name in ((gk.ESTIMATOR_REF_V2, "REF_V2"), (gk.ESTIMATOR_PT_MIS, "PT_MIS")):
        ctx = gk.Context(W, H)
        ctx.upload(s)
        cam = scenes.CORNELL_CAM
        p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
        p.estimator = est
        p.bounces = p.max_bounces = 4
        p.denoiser = gk.DENOISER_NONE
        for f in range(3):
            p.total_frames = f
            ctx.frame(p)
            ctx.render()
        ctx.reset()
        s0 = ctx.stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(16):
            p.total_frames = f
            ctx.frame(p)
            ctx.render()
        ctx.sync()
        dt = time.perf_counter() - t0
        s1 = ctx.stats()
        rad = ctx.read("RADIANCE")
        img = rad[..., :3] / np.maximum(rad[..., 3:4], 1)
        res = {"ms_per_frame": round(dt / 16 * 1e3, 3),
               "mrays_per_s": round((s1.rays_traced - s0.rays_traced) / dt / 1e6, 2)}
        if est == gk.ESTIMATOR_PT_MIS:
            ref_spp = args.relmse or 4096
            ref = reference_frame(ctx, p, ref_spp)
            res["relmse_16spp_vs_fp64_ref"] = relmse(img, ref)
            res["reference_spp"] = ref_spp
        if est == gk.ESTIMATOR_REF_V2:
            refv2_params = gk.FrameParams.from_buffer_copy(p)
        out[name] = res
        ctx.close()
    cpu = None
    if not args.no_cpu_baseline:
        # the REF_V2 restatement (FPathTracingRendererV2) on all host threads, 4 frames
        from oracle import pyoracle as po
        hc = host_cpu()
        threads = hc["threads"]
        osc = po.OracleScene(s)
        nrays = 0
        t0 = time.perf_counter()
        for f in range(4):
            refv2_params.total_frames = f
            nrays += int(osc.render_refv2(refv2_params, W, H, threads=threads)["rays"])
        dt = (time.perf_counter() - t0) / 4
        cpu = {"value": round(nrays / 4 / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"oracle REF_V2 frames 0-3 of the Cornell box at {W}x{H}, 4 bounces "
                         f"({dt * 1e3:.1f} ms/frame against {out['REF_V2']['ms_per_frame']} ms on the GPU)",
               "ms_per_frame": round(dt * 1e3, 2), "cpu_model": hc["cpu_model"],
               "host_cpus_available": hc["host_cpus_available"]}
    return {"metric": METRIC, "value": out["PT_MIS"]["mrays_per_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": 16,
            "warmup": 3, "ms_per_step": out["PT_MIS"]["ms_per_frame"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "Cornell box, 256x256, 16 frames x 1 spp, 4 bounces", **out},
            "cpu_baseline": cpu}


def denoise_compare(ctx, net, p, ref_spp, reps=20):
    """One 1-spp frame of the current view (TemporalFrames = 1) through the bilateral and the
    learned denoisers, both writing linear radiance; relMSE of the raw frame and of each
    denoised image against an FP64 ref_spp render; post-pass milliseconds of each. Also
    both denoisers after 16 frames of temporal accumulation of the static view (keys
    *_temporal16), the renderer's usual mode."""
    import numpy as np

    import gknextengine_amd as gk
    q = gk.FrameParams.from_buffer_copy(p)
    q.temporal_frames = 1
    q.flags |= gk.FRAME_LINEAR_OUTPUT
    out, ms = {}, {}
    for name, den in (("bilateral", gk.DENOISER_BILATERAL), ("unet", gk.DENOISER_UNET)):
        q.denoiser = den
        t = []
        for r in range(reps):  # the same frame (same seeds) every time; HIP events around the denoiser
            ctx.reset()
            ctx.frame(q)
            ctx.render()
            ctx.sync()
            t.append(ctx.stats().ms_denoise)
        ms[name] = float(np.median(t[2:]))
        out[name] = ctx.read("DENOISED")[..., :3].astype(np.float64)
        rad = ctx.read("RADIANCE")
        out["noisy"] = rad[..., :3] / np.maximum(rad[..., 3:4], 1.0)
    # the renderer's own temporal path: 16 frames of the static view (other seeds each frame)
    # accumulated by reprojection with TemporalFrames = 16, then each denoiser
    for name, den in (("bilateral", gk.DENOISER_BILATERAL), ("unet", gk.DENOISER_UNET)):
        q16 = gk.FrameParams.from_buffer_copy(p)
        q16.temporal_frames = 16
        q16.flags |= gk.FRAME_LINEAR_OUTPUT
        q16.denoiser = den
        ctx.reset()
        for f in range(16):
            q16.total_frames = p.total_frames + f
            ctx.frame(q16)
            ctx.render()
        ctx.sync()
        out[f"{name}_temporal16"] = ctx.read("DENOISED")[..., :3].astype(np.float64)
    ref = reference_frame(ctx, q, ref_spp)
    res = {}
    for k, v in out.items():
        a, b, nbad = relmse_stats(v, ref)
        res[f"relmse_{k}"] = a
        if nbad:
            res[f"relmse_{k}_finite_only"] = b
            res[f"nonfinite_pixels_{k}"] = nbad
    res.update({"reference_spp": ref_spp, "denoise_ms_bilateral": round(ms["bilateral"], 3),
                "denoise_ms_unet": round(ms["unet"], 3),
                "unet_extra_ms_per_frame": round(ms["unet"] - ms["bilateral"], 3),
                "unet_params": net.num_params()})
    return res


def run_diff(args):
    """Gradient of an L2 image loss with respect to 256 materials x (albedo rgb,
    roughness, metallic) of the tiled Cornell box at 512x512, 4 spp, 4 bounces, by the
    path-replay adjoint. The target is rendered (256 spp, other seeds) with the parameters
    perturbed by N(0, 0.05). Checks: central differences (h = 1e-3) of the same-seed 4-spp
    loss for albedo (its sampling does not depend on albedo, so the fixed-seed loss is
    smooth in it); a 4096-spp reference gradient (adjoint of the 4096-spp residual); and a
    brute-force gradient for all five parameter kinds, central differences of the 4096-spp
    render at 128x128 with common seeds (it sees how roughness and metallic move the
    sampling), against the adjoint of the same linearised loss."""
    import numpy as np
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import diff, scenes
    scene, nmat = scenes.cornell_tiles()
    base = list(scene.materials)
    cam = scenes.CORNELL_CAM

    def params(W, H, spp, frame, seed=3):
        p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.samples = spp
        p.bounces = p.max_bounces = 4
        p.total_frames = frame
        p.global_seed = seed
        return p

    def render_big(ctx, theta, W, H, spp, frame0, chunk=64):
        """Mean of spp paths per pixel, chunk spp per launch."""
        acc = torch.zeros((H, W, 3), device="cuda")
        for k in range(0, spp, chunk):
            acc += diff.render(theta, ctx, base, params(W, H, chunk, frame0 + k)).detach() * chunk
        return acc / spp

    def adjoint(ctx, theta, W, H, spp, frame0, dL, chunk=64):
        """Gradient of <dL, image> with the image at spp paths per pixel (the adjoint replays
        the seeds of frames frame0 .. frame0 + spp - 1)."""
        g = torch.zeros_like(theta)
        for k in range(0, spp, chunk):
            t = theta.clone().requires_grad_(True)
            (diff.render(t, ctx, base, params(W, H, chunk, frame0 + k), decorrelate=False) * dL).sum().backward()
            g += t.grad * (chunk / spp)
        return g

    W = H = 512
    spp = 4
    ctx = gk.Context(W, H)
    ctx.upload(scene)
    theta = diff.params_of(base).cuda()
    gen = torch.Generator().manual_seed(5)
    pert = theta.clone()
    pert[:nmat] = (pert[:nmat] + 0.05 * torch.randn(nmat, 5, generator=gen).cuda()).clamp(0.02, 0.98)
    log("rendering target")
    target = render_big(ctx, pert, W, H, 256, 50_000_000)
    p4 = params(W, H, spp, 1)

    def loss_grad(t, decorrelate=True):
        t = t.clone().requires_grad_(True)
        img = diff.render(t, ctx, base, p4, decorrelate=decorrelate)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        return ((img.detach().double() - target.double()) ** 2).mean().item(), t.grad

    def loss_only(t):
        with torch.no_grad():
            img = diff.render(t, ctx, base, p4)
        return ((img.double() - target.double()) ** 2).mean().item()

    for _ in range(2):
        loss_grad(theta)
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, g4 = loss_grad(theta)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = float(np.median(times) * 1e3)
    # plain loss.backward() replays independent seeds (the default); the same-seed gradient
    # is the exact derivative of the 4-spp loss, which central differences measure
    _, g4_dec = loss_grad(theta)
    _, g4 = loss_grad(theta, decorrelate=False)
    # central differences of the same-seed 4-spp loss in albedo
    rng = np.random.default_rng(9)
    lit = np.nonzero(g4[:nmat, :3].abs().sum(dim=1).cpu().numpy() > 0)[0]
    h = 1e-3
    fd_err = []
    for m in rng.choice(lit, 32, replace=False):
        k = int(rng.integers(3))
        tp, tm = theta.clone(), theta.clone()
        tp[m, k] += h
        tm[m, k] -= h
        fd = (loss_only(tp) - loss_only(tm)) / (2 * h)
        fd_err.append(abs(fd - g4[m, k].item()) / max(abs(fd), 1e-12))
    # 4096-spp reference gradient of the same loss: independent seeds for image and adjoint
    log("4096-spp reference gradient")
    t0 = time.perf_counter()
    img_ref = render_big(ctx, theta, W, H, 4096, 60_000_000)
    dL = 2.0 * (img_ref - target) / target.numel()
    g_ref = adjoint(ctx, theta, W, H, 4096, 70_000_000, dL)
    t_ref = time.perf_counter() - t0

    def cos(x, y):
        x, y = x[:nmat].flatten(), y[:nmat].flatten()
        return float((x @ y) / (x.norm() * y.norm()))

    # The 4-spp gradient two ways. Correlated (decorrelate=False): residual and adjoint
    # from the same paths; biased by the gradient of the image variance. Decorrelated (the
    # default of loss.backward()): residual from one 4-spp render, adjoint with independent
    # seeds; an unbiased estimate of the reference gradient (its average over draws
    # converges to it).
    cos_corr = cos(g4, g_ref)
    cos_default = cos(g4_dec, g_ref)
    # (pcg4d seeds are (x, y, frame + sample): a call with `spp` samples at frame f uses
    # frames f .. f+spp-1, so the two renders of a pair start spp frames apart)
    g_dec = []
    for d in range(16):
        f = 2_000_000 + 2 * spp * d
        img_a = diff.render(theta, ctx, base, params(W, H, spp, f)).detach()
        g_dec.append(adjoint(ctx, theta, W, H, spp, f + spp, 2.0 * (img_a - target) / target.numel(), chunk=spp))
    cos_dec1 = cos(g_dec[0], g_ref)
    cos_dec16 = cos(sum(g_dec) / 16, g_ref)
    # Brute force at 128x128 for every parameter kind: central differences (h = 0.02) of the
    # mean radiance rendered with common seeds, 8 chunks of 512 spp giving a standard error,
    # against the adjoint of the same functional (also in 8 chunks). The finite difference
    # includes how roughness and metallic move the sampled directions and lobe choices.
    log("brute-force gradient at 128x128")
    ctx.close()
    Ws = 128
    cs = gk.Context(Ws, Ws)
    cs.upload(scene)
    dLs = torch.full((Ws, Ws, 3), 1.0 / (Ws * Ws * 3), device="cuda")
    chunks, cspp, f0 = 8, 512, 80_000_000
    g_ch = [adjoint(cs, theta, Ws, Ws, cspp, f0 + 100_000 + c * cspp, dLs) for c in range(chunks)]
    g_adj = sum(g_ch) / chunks
    g_adj_se = torch.stack(g_ch).std(dim=0) / np.sqrt(chunks)
    hb = 0.02
    bf = []
    order = torch.argsort(g_adj[:nmat].abs(), dim=0, descending=True).cpu().numpy()
    for k in range(5):
        for m in order[:8, k]:
            tp, tm = theta.clone(), theta.clone()
            tp[m, k] += hb
            tm[m, k] -= hb
            fds = []
            for c in range(chunks):
                fr = f0 + c * cspp
                fds.append(float(((render_big(cs, tp, Ws, Ws, cspp, fr) - render_big(cs, tm, Ws, Ws, cspp, fr))
                                  * dLs).sum().item() / (2 * hb)))
            fd, se = float(np.mean(fds)), float(np.std(fds, ddof=1) / np.sqrt(chunks))
            a, sa = g_adj[m, k].item(), g_adj_se[m, k].item()
            bf.append((k, fd, a, (a - fd) / max(np.hypot(se, sa), 1e-12)))
    cs.close()
    kinds = ["albedo_r", "albedo_g", "albedo_b", "roughness", "metallic"]
    by_kind = {}
    for k in range(5):
        rows = [x for x in bf if x[0] == k]
        fdv = np.array([x[1] for x in rows])
        adv = np.array([x[2] for x in rows])
        z = np.array([x[3] for x in rows])
        by_kind[kinds[k]] = {"rel_err": round(float(np.linalg.norm(fdv - adv) / np.linalg.norm(fdv)), 4),
                             "mean_abs_z": round(float(np.abs(z).mean()), 2), "max_abs_z": round(float(np.abs(z).max()), 2)}
    return {"metric": METRIC, "value": round(ms, 3), "unit": "ms per forward+adjoint (512x512, 4 spp)",
            "n_gpus": 1, "steps": args.steps, "warmup": 2, "ms_per_step": round(ms, 3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "differentiable render: tiled Cornell box, 256 materials x 5 params, "
                                   "512x512, 4 spp, 4 bounces, L2 loss, path-replay adjoint",
                       "loss": loss, "paths_per_pass": W * H * spp,
                       "fd_albedo_h1e-3_median_rel_err": float(np.median(fd_err)),
                       "fd_albedo_h1e-3_max_rel_err": float(np.max(fd_err)), "fd_params": len(fd_err),
                       "reference_gradient_seconds": round(t_ref, 1),
                       "cosine_vs_4096spp_reference": {"4spp_loss_backward_default": round(cos_default, 4),
                                                       "4spp_correlated": round(cos_corr, 4),
                                                       "4spp_decorrelated": round(cos_dec1, 4),
                                                       "mean_of_16_decorrelated_4spp": round(cos_dec16, 4)},
                       "bruteforce_128px_4096spp_h0.02": by_kind, "bruteforce_params": len(bf)}}


MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 matrix-core peak of one MI355X


def unet_flops(net, W, H, real_in=10):
    """Algorithmic FLOPs of one network evaluation at W x H: sum over convolutions of
    2 * 9 * cin * cout * pixels at the layer's level, with the real channel counts (10 input
    channels, 3 output channels), not the padded ones the kernels compute."""
    total = 0
    per_layer = []
    for i in range(net.LAYERS):
        L = net.layer(i)
        cin = (real_in if i == 0 else L.cin_a) + L.cin_b
        px = ((W + (1 << L.level) - 1) >> L.level) * ((H + (1 << L.level) - 1) >> L.level)
        f = 2 * 9 * cin * L.cout_real * px
        per_layer.append(f)
        total += f
    return total, per_layer


def run_denoise(args):
    """The learned denoiser against the bilateral filter on 1-spp frames at 1080p
    (the city at the benchmark camera, the Cornell box), relMSE against FP64 renders."""
    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    W, H = args.width, args.height
    res = {}
    for name, scene, cam, bounces in (("city", lambda: gk.procgen_city(64, 4, 256, 64, 2024), scenes.CITY_CAM, 8),
                                      ("cornell", gk.procgen_cornell, scenes.CORNELL_CAM, 4)):
        ctx = gk.Context(W, H)
        ctx.upload(scene())
        net = gk.UNet.pretrained(ctx)
        p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.bounces = p.max_bounces = bounces
        p.global_seed = 11
        p.total_frames = 5
        res[name] = denoise_compare(ctx, net, p, args.relmse or 4096)
        flops, _ = unet_flops(net, W, H)
        res[name]["unet_gflop_per_frame"] = round(flops / 1e9, 1)
        res[name]["unet_tflops"] = round(flops / (res[name]["denoise_ms_unet"] * 1e-3) / 1e12, 1)
        log(name, res[name])
        ctx.close()
    c = res["city"]
    # the network's time includes its feature gather and output conversion (HIP events
    # around the whole denoiser pass): a lower bound of the convolutions' own MFMA rate
    tflops = c["unet_tflops"]
    return {"metric": METRIC, "value": c["relmse_unet"], "unit": "relMSE (linear, vs FP64 render)", "n_gpus": 1,
            "steps": 20, "warmup": 0, "ms_per_step": c["denoise_ms_unet"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
            "config": {"workload": "learned denoiser vs bilateral on 1-spp PT_MIS frames, 1920x1080",
                       "resolution": [W, H], **{f"{k}_{m}": v for k, d in res.items() for m, v in d.items()}},
            "roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None,
                         "kernel": "the denoiser pass (12 MFMA convolutions: feature gather, pooling and display conversion fused)",
                         "flop_per_frame": c["unet_gflop_per_frame"] * 1e9}}


def run_probes(args):
    """Ambient-cube probe grid on the city: the reference's 192 x 48 x 192 grid (0.25 m),
    device voxelization of every probe (14 rays each), full bake passes (96 hemisphere rays
    and one sun ray per face of every probe within 2 m of a surface) and the renderer's
    per-frame slice (1/120 of the grid, BakeSpeedLevel 1)."""
    import numpy as np

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    log("building city scene")
    scene = gk.procgen_city(64, 4, 256, 64, 2024)
    ctx = gk.Context(64, 64)
    ctx.upload(scene)
    cam = scenes.CITY_CAM
    p = gk.make_frame_params(cam["eye"], cam["target"], 64, 64, fov=cam["fov"])
    p.estimator = gk.ESTIMATOR_REF_V2
    p.has_sky = 1
    p.sky_color[:] = [60.0, 80.0, 120.0, 1.0]
    p.has_sun = 1
    p.sun_direction[:] = [0.36, 0.86, -0.36, 0.0]
    p.sun_color[:] = [400.0, 380.0, 340.0, 1.0]
    ctx.frame(p)
    ctx.probes_init()
    ctx.probes_voxelize()
    ctx.probes_bake()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.probes_voxelize()
    ctx.sync()
    t_vox = time.perf_counter() - t0
    _, vox = ctx.probes_read()
    n = len(vox)
    active = int(((vox["dist_gg_z01"] & 0xFF) < 8).sum())
    rays_pass = active * (6 * 16 + 6 * 1)      # hemisphere rays + one sun ray per face
    for _ in range(args.warmup):
        ctx.probes_bake()
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.probes_bake()
    ctx.sync()
    t_bake = (time.perf_counter() - t0) / args.steps
    s1 = ctx.stats()
    trace_ms = (s1.ms_trace - s0.ms_trace) / args.steps
    rays_counted = (s1.rays_traced - s0.rays_traced) // args.steps
    t0 = time.perf_counter()
    for f in range(120):
        ctx.probes_bake_frame(f, 120)
    ctx.sync()
    t_slice = (time.perf_counter() - t0) / 120
    ctx.close()
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po
        osc = po.OracleScene(scene)
        G = po.OracleProbes(*gk.REFERENCE_PROBE_GRID, gk.CUBE_DTYPE, gk.VOXEL_DTYPE)
        G.vox[:] = vox
        G.vox["age"] = 0
        act_idx = np.nonzero((vox["dist_gg_z01"] & 0xFF) < 8)[0]
        first = int(act_idx[len(act_idx) // 2])
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        G.bake(osc, p, None, first=first, count=16384, threads=threads)
        dt = time.perf_counter() - t0
        a = int(((vox["dist_gg_z01"][first:first + 16384] & 0xFF) < 8).sum())
        cpu = {"value": round(a * 102 / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"oracle bake of probes [{first}, {first + 16384}) ({a} baked), city, reference grid"}
    mrays = rays_pass / t_bake / 1e6
    return {"metric": METRIC, "value": round(mrays, 1), "unit": "Mrays/s (probe bake)", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_bake * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "ambient-cube probe grid 192x48x192 @0.25 m over the 0.92M-triangle city",
                       "probes": n, "baked_probes": active, "rays_per_pass": rays_pass,
                       "rays_counted_per_pass": int(rays_counted),
                       "voxelize_ms": round(t_vox * 1e3, 3), "voxelize_mrays_per_s": round(14 * n / t_vox / 1e6, 1),
                       "bake_pass_ms": round(t_bake * 1e3, 3), "bake_trace_ms": round(trace_ms, 3),
                       "frame_slice_ms": round(t_slice * 1e3, 3)},
            "cpu_baseline": cpu}


def run_output(args):
    """The output stage on the city at SuperResolution 1 (render 1280x720, display
    1920x1080): HDR compose, then the linear-blit and FSR EASU resolves, per frame."""
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    ow, oh = args.width, args.height
    W, H = gk.render_extent(ow, oh, 1)
    ctx = gk.Context(W, H)
    ctx.upload(gk.procgen_city(64, 4, 256, 64, 2024))
    cam = scenes.CITY_CAM
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = p.max_bounces = args.bounces
    p.denoiser = gk.DENOISER_BILATERAL
    p.hdr = 1
    p.flags |= gk.FRAME_NO_MATERIAL_SORT
    out = torch.empty((oh, ow, 4), dtype=torch.float16, device="cuda")
    torch.cuda.synchronize()
    res = {}
    for name, mode in (("linear", gk.UPSCALE_LINEAR), ("fsr_easu", gk.UPSCALE_FSR_EASU)):
        for f in range(args.warmup):
            p.total_frames = f
            ctx.frame(p)
            ctx.render()
            check = gk.lib().gk_upscale(ctx.h, mode, ow, oh, out.data_ptr())
            assert check == 0
        ctx.sync()
        t0 = time.perf_counter()
        for f in range(args.steps):
            p.total_frames = args.warmup + f
            ctx.frame(p)
            ctx.render()
        ctx.sync()
        t_frame = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            gk.lib().gk_upscale(ctx.h, mode, ow, oh, out.data_ptr())
        ctx.sync()
        t_up = (time.perf_counter() - t0) / args.steps
        res[name] = {"frame_ms": round(t_frame * 1e3, 3), "upscale_ms": round(t_up * 1e3, 4)}
    ctx.close()
    px = ow * oh
    bytes_px = 8 + 8  # RGBA16F out + at least one RGBA16F source texel read per output pixel
    return {"metric": METRIC, "value": res["fsr_easu"]["upscale_ms"], "unit": "ms (FSR EASU 1280x720 -> 1920x1080)",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["fsr_easu"]["frame_ms"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "city PT_MIS at SuperResolution 1 (1280x720) + HDR compose + resolve to 1920x1080",
                       "render": [W, H], "display": [ow, oh], **{f"{k}_{m}": v for k, d in res.items() for m, v in d.items()},
                       "resolve_min_bytes_per_launch": px * bytes_px}}


def run_copy(args):
    """A plain 256 MiB device-to-device copy (torch ``copy_``, kernel __amd_rocclr_copyBuffer):
    the measured HBM ceiling beside the 8 TB/s peak, and the known-size kernel whose
    FETCH_SIZE / WRITE_SIZE counters calibrate the traffic figures of the other kernels."""
    import torch
    n = 256 * 1024 * 1024 // 4
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    assert bool((y == 1).all())
    gbs = 2 * 4 * n / dt / 1e9
    return {"metric": "HBM copy bandwidth", "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "256 MiB device-to-device copy (read + write)", "bytes_per_step": 2 * 4 * n},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None}}


def run_diff_timing(args):
    """The differentiable render's step alone: forward + adjoint of the L2 loss at 512x512, 4 spp, 4 bounces,
    256 materials x 5 parameters (run_diff adds the gradient checks)."""
    import numpy as np
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import diff, scenes
    scene, nmat = scenes.cornell_tiles()
    base = list(scene.materials)
    cam = scenes.CORNELL_CAM
    W = H = 512
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.samples = 4
    p.bounces = p.max_bounces = 4
    p.total_frames = 1
    p.global_seed = 3
    ctx = gk.Context(W, H)
    ctx.upload(scene)
    theta = diff.params_of(base).cuda()
    target = torch.zeros((H, W, 3), device="cuda")

    def step():
        t = theta.clone().requires_grad_(True)
        img = diff.render(t, ctx, base, p)
        ((img - target) ** 2).mean().backward()
        return t.grad

    for _ in range(2):
        step()
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ctx.close()
    ms = float(np.median(times) * 1e3)
    return {"metric": "forward + adjoint ms", "value": round(ms, 3), "unit": "ms", "steps": args.steps,
            "ms_per_step": round(ms, 3), "higher_is_better": False,
            "config": {"workload": "differentiable PT_MIS step (512x512, 4 spp): render + L2 loss + path-replay adjoint",
                       "resolution": [W, H], "spp": 4, "bounces": 4, "materials": nmat, "params_per_material": 5,
                       "paths_per_step": W * H * 4}}


def run_city_ref_v2(args):
    """The reference's own estimator (REF_V2: FPathTracingRendererV2's Render loop, early exit
    0.5, no next-event estimation, probes off) on the 1080p procedural city, 1 spp, 8
    bounces, with reprojection and the bilateral filter: 3 warm-up frames, then 20 timed."""
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    W, H = 1920, 1080
    scene = gk.procgen_city(64, 4, 256, 64, 2024)
    ctx = gk.Context(W, H)
    ctx.upload(scene)
    cam = scenes.CITY_CAM
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    p.estimator = gk.ESTIMATOR_REF_V2
    p.bounces = p.max_bounces = 8
    p.samples = 1
    p.global_seed = 1
    p.denoiser = gk.DENOISER_BILATERAL
    frames = 20
    for f in range(3):
        p.total_frames = f
        ctx.frame(p)
        ctx.render()
    ctx.sync()
    s0 = ctx.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(frames):
        p.total_frames = 3 + f
        ctx.frame(p)
        ctx.render()
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s1 = ctx.stats()
    ctx.close()
    rays = s1.rays_traced - s0.rays_traced
    return {"metric": METRIC, "value": round(rays / wall / 1e6, 2), "unit": "Mrays/s", "steps": frames,
            "ms_per_step": round(wall / frames * 1e3, 3), "higher_is_better": True,
            "config": {"workload": "procedural city, REF_V2 estimator 1 spp 8 bounces + reproject + bilateral",
                       "resolution": [W, H], "rays_per_frame": round(rays / frames),
                       "trace_ms_per_frame": round((s1.ms_trace - s0.ms_trace) / frames, 3)}}


def run_side_configs(args):
    """The other configurations, measured in the same process after the headline one (a
    few seconds each; reduced reference spp for the relMSE figures)."""
    import copy

    import torch
    side = {}
    specs = (("city_ref_v2", run_city_ref_v2, {}), ("cornell", run_cornell, {"relmse": 1024}), ("dynamic", run_dynamic, {}),
             ("denoise", run_denoise, {"relmse": 256}), ("diff", run_diff_timing, {"steps": 10}))
    for name, fn, over in specs:
        a = copy.copy(args)
        a.config = name
        a.no_cpu_baseline = True
        for k, v in over.items():
            setattr(a, k, v)
        t0 = time.perf_counter()
        try:
            r = fn(a)
            r.pop("cpu_baseline", None)
            r["wall_s"] = round(time.perf_counter() - t0, 1)
            side[name] = r
        except Exception as e:  # a side measurement never hides the headline line
            side[name] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
        log("side config", name, "done")
    return side


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bounces", type=int, default=8)
    ap.add_argument("--config", default="city", choices=["city", "dynamic", "cornell", "denoise", "diff", "probes", "output", "copy"])
    ap.add_argument("--relmse", type=int, default=0, help="reference spp for relMSE (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sort", default="off", choices=["on", "off"], help="material-ordered shading")
    ap.add_argument("--mixed-materials", action="store_true",
                    help="city materials spread over Lambertian/Metallic/Dielectric/Mixture")
    ap.add_argument("--denoiser", default="bilateral", choices=["none", "bilateral"])
    ap.add_argument("--tlas", default="rebuild", choices=["rebuild", "refit"],
                    help="dynamic config: per-frame TLAS update")
    ap.add_argument("--per-frame", action="store_true", help="dynamic config: per-frame trace ms and TLAS SAH")
    ap.add_argument("--frames", type=int, default=120, help="dynamic config: frames after the first (profiling runs)")
    ap.add_argument("--pipelines", type=int, default=2, choices=[1, 2],
                    help="city: 1 renders the timed frames as one path pipeline (GK_FRAME_SINGLE_PIPELINE)")
    ap.add_argument("--no-side", action="store_true",
                    help="city config: skip the other configurations' figures (side_configs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the path tracer has no multi-GPU implementation in this library.",
              file=sys.stderr)
        sys.exit(2)
    run = {"city": run_city, "dynamic": run_dynamic, "cornell": run_cornell, "denoise": run_denoise,
           "diff": run_diff, "probes": run_probes, "output": run_output,
           "copy": run_copy}[args.config]
    res = run(args)
    if args.config == "city" and not args.no_side:
        res["side_configs"] = run_side_configs(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
