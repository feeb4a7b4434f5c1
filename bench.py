#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Default workload (the metric's single-GPU configuration, config 2): the ~0.92 M-triangle
procedural city at 1920x1080, PT_MIS estimator (GGX bounded-VNDF + Lambert, NEE to
emissive triangles, MIS), 1 spp, 8 bounces, FP32, then reprojection and the bilateral
denoiser. Scene and BVH are resident in HBM before the timed region; a step renders one
frame. ``value`` is the whole-job ray throughput (primary + extension + shadow rays that
the trace kernels processed, divided by wall time). The roofline object describes the
traversal kernel k_trace, timed with HIP events around every launch inside the timed
region; ``cpu_baseline`` is the reference's tinybvh BVH::Intersect on host cores.

Other configurations (``--config``): ``dynamic`` (config 3: 4096 moving instanced bricks,
TLAS refit every frame, 4 bounces, reprojection + bilateral, 120 frames) and ``cornell``
(config 1: 256x256, 16 frames x 1 spp, 4 bounces). ``--relmse N`` adds the relMSE of the
measured frame against an N-spp FP64-accumulated render of the same frame.
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


def cpu_baseline(scene, params, W, H):
    """tinybvh BVH::Intersect (the reference's CPU ray cast) on the camera rays and the
    first-bounce diffuse rays of the same frame, all host threads, 64K-ray tasks."""
    import numpy as np

    from oracle import pyoracle as po
    if not po.ref_available():
        return None
    threads = min(os.cpu_count() or 1, 16)
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
    return {"value": round(n / ((ms_primary + ms_bounce) * 1e-3) / 1e6, 2), "unit": "Mrays/s", "cores": threads,
            "kind": "reference",
            "sample": f"tinybvh 1.3.8 BVH::Intersect (binned SAH, single ray) on {rays.shape[0]} camera + "
                      f"{b.shape[0]} first-bounce rays of the same 1080p city frame; BLAS build {build_ms:.0f} ms "
                      f"single-threaded"}


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
    if args.no_sort:
        p.flags |= gk.FRAME_NO_MATERIAL_SORT
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
    ms_trace = s1.ms_trace - s0.ms_trace
    value = rays / wall / 1e6
    achieved = BYTES_PER_RAY * rays / (ms_trace * 1e-3) / 1e9
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
        ctx.frame(p)
        ctx.render()
        raw = ctx.read("RADIANCE")[..., :3]
        den = ctx.read("DENOISED")[..., :3].astype(np.float64)
        ref = reference_frame(ctx, p, args.relmse)
        sc = p.paper_white_nit / 40000.0
        extra = {"relmse_1spp": relmse(raw, ref), "reference_spp": args.relmse,
                 "relmse_tonemapped_1spp": relmse(gt_tonemap(raw * sc), gt_tonemap(ref * sc)),
                 "relmse_tonemapped_bilateral": relmse(den, gt_tonemap(ref * sc))}
    cpu = None if args.no_cpu_baseline else cpu_baseline(scene, p, W, H)
    ctx.close()
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "procedural city (config 2), PT_MIS 1 spp 8 bounces + reproject + bilateral",
                   "resolution": [W, H], "triangles": int(ntris), "bounces": args.bounces, "spp": 1,
                   "estimator": "PT_MIS", "material_sort": not args.no_sort,
                   "rays_per_frame": round(rays / args.steps), "trace_ms_per_frame": round(ms_trace / args.steps, 3),
                   "reproject_ms": round(s1.ms_reproject, 3), "denoise_ms": round(s1.ms_denoise, 3),
                   "blas_build_ms": round(bvh.build_ms, 2), "bvh_sah_cost": round(bvh.sah_cost, 2),
                   "blas_nodes": bvh.blas_nodes, **extra},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_trace (closest + any hit)", "bytes_per_ray": BYTES_PER_RAY},
        "cpu_baseline": cpu,
    }


def run_dynamic(args):
    """Config 3: 120 frames of 4096 moving bricks at 1080p, 4 bounces, TLAS refit per
    frame from device-resident matrices, reprojection + bilateral."""
    import numpy as np
    import torch

    import gknextengine_amd as gk
    from gknextengine_amd import scenes
    W, H = args.width, args.height
    b = scenes.Bricks()
    nframes = 120
    worlds = np.stack([b.worlds(f / 60.0) for f in range(nframes + 1)])           # (F, n, 4, 4)
    d_worlds = torch.from_numpy(np.ascontiguousarray(np.transpose(worlds, (0, 1, 3, 2)))).cuda()  # column-major
    ctx = gk.Context(W, H)
    ctx.upload(b.scene)
    cam = scenes.BRICKS_CAM
    prev = None
    times = []
    stats0 = None
    p = None
    for f in range(nframes + 1):
        p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"], frame=f, prev=prev)
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.bounces = p.max_bounces = 4
        p.samples = 1
        p.has_sky = 1
        p.sky_color[:] = [0.6, 0.7, 0.9, 1.0]
        p.temporal_frames = 16
        p.denoiser = gk.DENOISER_BILATERAL
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
        prev = p
    s1 = ctx.stats()
    rays = s1.rays_traced - stats0.rays_traced
    ms = np.array(times[9:]) * 1e3
    extra = {}
    if args.relmse:
        den = ctx.read("DENOISED")[..., :3].astype(np.float64)
        ref = reference_frame(ctx, p, args.relmse)
        sc = p.paper_white_nit / 40000.0
        extra = {"relmse_tonemapped_bilateral_frame120": relmse(den, gt_tonemap(ref * sc)),
                 "reference_spp": args.relmse}
    ctx.close()
    return {"metric": METRIC, "value": round(rays / (ms.sum() * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
            "steps": len(ms), "warmup": 9, "ms_per_step": round(float(np.median(ms)), 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "dynamic bricks (config 3): 4096 instances x 24 tris, TLAS refit per frame, "
                                   "PT_MIS 1 spp 4 bounces, reproject + bilateral", "resolution": [W, H],
                       "ms_frame_mean": round(float(ms.mean()), 3), "ms_frame_p95": round(float(np.percentile(ms, 95)), 3),
                       "tlas_ms": round(s1.ms_tlas, 3), "reproject_ms": round(s1.ms_reproject, 3),
                       "denoise_ms": round(s1.ms_denoise, 3), **extra}}


def run_cornell(args):
    """Config 1: 256x256 Cornell box, 16 frames x 1 spp, 4 bounces."""
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
        if args.relmse and est == gk.ESTIMATOR_PT_MIS:
            ref = reference_frame(ctx, p, args.relmse)
            res["relmse_16spp_vs_ref"] = relmse(img, ref)
        out[name] = res
        ctx.close()
    return {"metric": METRIC, "value": out["PT_MIS"]["mrays_per_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": 16,
            "warmup": 3, "ms_per_step": out["PT_MIS"]["ms_per_frame"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "Cornell box (config 1), 256x256, 16 frames x 1 spp, 4 bounces", **out}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bounces", type=int, default=8)
    ap.add_argument("--config", default="city", choices=["city", "dynamic", "cornell"])
    ap.add_argument("--relmse", type=int, default=0, help="reference spp for relMSE (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sort", action="store_true", help="shade in path order (no material sort)")
    ap.add_argument("--denoiser", default="bilateral", choices=["none", "bilateral"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (DESIGN.md describes how it would "
              "shard).", file=sys.stderr)
        sys.exit(2)
    run = {"city": run_city, "dynamic": run_dynamic, "cornell": run_cornell}[args.config]
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
