#!/usr/bin/env python3
"""Benchmark: path-traced frames of the procedural city at 1920x1080.

Workload (the metric's single-GPU configuration): ~0.92 M-triangle procedural city
(64x64 lognormal boxes, faces split 4x4, 256x256 ground, 64 emissive roof quads),
PT_MIS estimator (GGX VNDF + Lambert, NEE to emissive triangles, MIS), 1 spp,
8 bounces, FP32, followed by reprojection and the bilateral denoiser. The scene and
its BVH are resident in HBM before the timed region; one step renders one frame.

Prints one JSON line (metric, value = whole-job Mrays/s, ms_per_step, roofline of the
traversal kernel measured with HIP events inside the timed region, and a CPU baseline:
the reference's tinybvh Intersect on host cores).
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
CITY_CAM = dict(eye=(-20.0, 28.0, -44.0), target=(2.0, 0.0, 4.0), fov=50.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, params, W, H, max_rays=4_147_200):
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
            "sample": f"tinybvh 1.3.8 BVH::Intersect (binned SAH, single-ray) on {rays.shape[0]} camera + "
                      f"{b.shape[0]} first-bounce rays of the same 1080p city frame; BLAS build {build_ms:.0f} ms "
                      f"single-threaded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bounces", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--denoiser", default="bilateral", choices=["none", "bilateral"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import gknextengine_amd as gk
    W, H = args.width, args.height
    log("building city scene")
    scene = gk.procgen_city(64, 4, 256, 64, 2024)
    ntris = scene.meshes[0]["indices"].shape[0]
    ctx = gk.Context(W, H)
    ctx.upload(scene)
    bvh = ctx.bvh_stats()
    p = gk.make_frame_params(CITY_CAM["eye"], CITY_CAM["target"], W, H, fov=CITY_CAM["fov"])
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = args.bounces
    p.max_bounces = args.bounces
    p.samples = 1
    p.global_seed = 1
    p.denoiser = gk.DENOISER_BILATERAL if args.denoiser == "bilateral" else gk.DENOISER_NONE
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
    cpu = None if args.no_cpu_baseline else cpu_baseline(scene, p, W, H)
    out = {
        "metric": "Mrays/s & ms/frame @1080p; relMSE vs 4096-spp, bilateral vs learned denoise",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "procedural city (config 2), PT_MIS 1 spp 8 bounces + reproject + bilateral",
                   "resolution": [W, H], "triangles": int(ntris), "bounces": args.bounces, "spp": 1,
                   "estimator": "PT_MIS", "rays_per_frame": round(rays / args.steps),
                   "trace_ms_per_frame": round(ms_trace / args.steps, 3),
                   "reproject_ms": round(s1.ms_reproject, 3), "denoise_ms": round(s1.ms_denoise, 3),
                   "blas_build_ms": round(bvh.build_ms, 2), "bvh_sah_cost": round(bvh.sah_cost, 2),
                   "blas_nodes": bvh.blas_nodes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_trace (closest + any hit)", "bytes_per_ray": BYTES_PER_RAY},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
