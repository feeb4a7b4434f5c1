"""Training data and training loop of the learned denoiser (gk_unet_*).

Pairs are rendered by the path tracer itself. For each pair a random view of the Cornell
box or of the procedural city is drawn; a 1-spp PT_MIS frame (reprojection with
TemporalFrames = 1, so the accumulation targets hold that frame alone) gives the 10
input features through gk_denoise_input, and an N-spp FP32 accumulation of the same
view with independent seeds gives the target radiance. The network is trained with
gk_unet_train_step (this library's forward, backward and Adam kernels) on batches of
256 x 256 tiles, with random flips and transposes of each tile.

    python -m gknextengine_amd.denoise_train --pairs 1024 --spp 256 --steps 3000 --out unet.bin

Device memory holds every pair: 3.4 MB per 256 x 256 pair.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

import gknextengine_amd as gk
from gknextengine_amd import scenes

TILE = 256


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gt_tonemap(x):
    """The display transform of GK_RT_DENOISED (GT tonemapper), numpy."""
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
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def cornell_view(rng):
    eye = (0.5 + rng.uniform(-0.3, 0.3), 0.5 + rng.uniform(-0.3, 0.3), rng.uniform(-1.7, -0.9))
    tgt = (0.5 + rng.uniform(-0.15, 0.15), 0.5 + rng.uniform(-0.15, 0.15), 0.0)
    return eye, tgt, rng.uniform(30.0, 50.0)


def city_view(rng):
    if rng.uniform() < 0.2:  # the benchmark's camera, other crops of it through the fov
        cam = scenes.CITY_CAM
        return cam["eye"], cam["target"], rng.uniform(20.0, 50.0)
    ex, ez = rng.uniform(-28.0, 28.0, 2)
    ey = rng.uniform(4.0, 30.0)
    ang, dist = rng.uniform(0.0, 2 * math.pi), rng.uniform(8.0, 30.0)
    tgt = (ex + dist * math.cos(ang), rng.uniform(0.0, 3.0), ez + dist * math.sin(ang))
    return (ex, ey, ez), tgt, rng.uniform(35.0, 60.0)


class PairRenderer:
    """One context per scene at TILE x TILE."""

    def __init__(self, size=TILE):
        self.size = size
        self.ctx = {}
        for name, scene in (("cornell", gk.procgen_cornell()), ("city", gk.procgen_city(64, 4, 256, 64, 2024))):
            c = gk.Context(size, size)
            c.upload(scene)
            self.ctx[name] = c

    def close(self):
        for c in self.ctx.values():
            c.close()

    def params(self, which, view, seed):
        eye, tgt, fov = view
        p = gk.make_frame_params(eye, tgt, self.size, self.size, fov=fov)
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.bounces = p.max_bounces = 4 if which == "cornell" else 8
        p.samples = 1
        p.global_seed = seed
        p.temporal_frames = 1
        p.denoiser = gk.DENOISER_BILATERAL
        p.flags |= gk.FRAME_LINEAR_OUTPUT
        return p

    def pair(self, which, view, seed, frame, spp, d_feat: int):
        """Renders one pair; writes the features to d_feat and returns (target rgb (h, w, 3) host,
        bilateral output (h, w, 3) linear radiance, host)."""
        ctx = self.ctx[which]
        p = self.params(which, view, seed)
        ctx.reset()
        p.total_frames = frame
        ctx.frame(p)
        ctx.render()
        ctx.denoise_input(d_feat)
        ctx.sync()
        bil = ctx.read("DENOISED")[..., :3].astype(np.float32)
        q = gk.FrameParams.from_buffer_copy(p)
        q.flags = gk.FRAME_SKIP_POST  # the target is the tracer's own FP32 accumulation
        q.samples = min(spp, 64)
        ctx.reset()
        for f in range((spp + q.samples - 1) // q.samples):
            q.total_frames = 1_000_000 + frame * 4096 + f
            ctx.frame(q)
            ctx.render()
        r = ctx.read("RADIANCE")
        return (r[..., :3] / np.maximum(r[..., 3:4], 1.0)).astype(np.float32), bil


def render_pairs(pr: PairRenderer, n, spp, seed, d_feat, d_tgt, bil_out=None, t_budget=None):
    """Fills d_feat[:n], d_tgt[:n] (torch device tensors); returns the number rendered."""
    import torch
    rng = np.random.default_rng(seed)
    t0 = time.time()
    for i in range(n):
        which = "cornell" if i % 4 == 0 else "city"
        view = cornell_view(rng) if which == "cornell" else city_view(rng)
        tgt, bil = pr.pair(which, view, int(rng.integers(1 << 30)), int(rng.integers(1 << 20)), spp,
                           d_feat[i].data_ptr())
        d_tgt[i].copy_(torch.from_numpy(tgt))
        if bil_out is not None:
            bil_out.append(bil)
        if i % 64 == 63:
            log(f"pairs {i + 1}/{n} in {time.time() - t0:.1f}s")
        if t_budget is not None and time.time() - t0 > t_budget:
            log(f"pair budget reached at {i + 1}")
            return i + 1
    return n


def augment(x, k):
    """k in 0..7: optional transpose, then flips of rows / columns (device tensors (n, h, w, c))."""
    if k & 4:
        x = x.transpose(1, 2)
    if k & 1:
        x = x.flip(1)
    if k & 2:
        x = x.flip(2)
    return x.contiguous()


def train(net: gk.UNet, d_feat, d_tgt, n, steps, batch=8, lr=1e-3, seed=0, lr_final=1e-4, t_budget=None):
    import torch
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.time()
    decay_from = int(0.75 * steps)
    done = 0
    for s in range(steps):
        idx = torch.from_numpy(rng.integers(0, n, batch)).cuda()
        k = int(rng.integers(8))
        bf = augment(d_feat[idx], k)
        bt = augment(d_tgt[idx], k)
        rate = lr if s < decay_from else lr * (lr_final / lr) ** ((s - decay_from) / max(1, steps - decay_from))
        loss = net.train_step(bf.data_ptr(), bt.data_ptr(), batch, TILE, TILE, rate)  # returns when done
        done = s + 1
        if s % 50 == 0 or s == steps - 1:
            hist.append((s, loss))
            if s % 250 == 0 or s == steps - 1:
                log(f"step {s} loss {loss:.5f} lr {rate:.2e} {time.time() - t0:.1f}s")
        if t_budget is not None and time.time() - t0 > t_budget:
            log(f"train budget reached at step {s + 1}")
            break
    torch.cuda.synchronize()
    return hist, done, time.time() - t0


def evaluate(net: gk.UNet, d_feat, d_tgt, idx, bil, paper_white=600.0):
    """relMSE of noisy / bilateral / U-Net against the targets, on linear radiance and on
    the tonemapped display values."""
    import torch
    sc = paper_white / 40000.0
    out = torch.empty(TILE * TILE * 3, device="cuda")
    res = {}
    for j, i in enumerate(idx):
        net.infer(d_feat[i].data_ptr(), out.data_ptr(), TILE, TILE)
        net.ctx.sync()
        img = {"noisy": d_feat[i].cpu().numpy()[..., :3], "bilateral": bil[j],
               "unet": out.cpu().numpy().reshape(TILE, TILE, 3)}
        t = d_tgt[i].cpu().numpy()
        for k, v in img.items():
            bad = ~np.all(np.isfinite(v), axis=-1)  # bilateral: zero weight sum at isolated outliers
            v = np.where(np.isfinite(v), v, 0.0)
            res.setdefault(k, []).append(relmse(v, t))
            res.setdefault(k + "_tonemapped", []).append(relmse(gt_tonemap(v * sc), gt_tonemap(t * sc)))
            res.setdefault(k + "_nonfinite_pixels", []).append(float(bad.sum()))
    return {k: float(np.mean(v)) for k, v in res.items()}


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=1024)
    ap.add_argument("--val", type=int, default=32)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--pair-budget", type=float, default=None, help="seconds for rendering training pairs")
    ap.add_argument("--train-budget", type=float, default=None, help="seconds for training")
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default=None)
    a = ap.parse_args()
    pr = PairRenderer()
    ntot = a.pairs + a.val
    d_feat = torch.empty((ntot, TILE, TILE, gk.UNet.IN_CHANNELS), device="cuda")
    d_tgt = torch.empty((ntot, TILE, TILE, 3), device="cuda")
    t0 = time.time()
    n = render_pairs(pr, a.pairs, a.spp, a.seed, d_feat, d_tgt, t_budget=a.pair_budget)
    t_pairs = time.time() - t0
    bil = []
    render_pairs(pr, a.val, a.spp, a.seed + 1000, d_feat[ntot - a.val:], d_tgt[ntot - a.val:], bil_out=bil)
    net = gk.UNet(pr.ctx["city"], a.base, seed=a.seed)
    if a.init:
        net.load(a.init)
    val_idx = list(range(ntot - a.val, ntot))
    before = evaluate(net, d_feat, d_tgt, val_idx, bil)
    log("validation before", before)
    hist, steps, t_train = train(net, d_feat, d_tgt, n, a.steps, a.batch, a.lr, a.seed, t_budget=a.train_budget)
    net.save(a.out)
    after = evaluate(net, d_feat, d_tgt, val_idx, bil)
    log("validation after", after)
    rep = {"pairs": n, "val_pairs": a.val, "target_spp": a.spp, "tile": TILE, "steps": steps, "batch": a.batch,
           "lr": a.lr, "seconds_pairs": round(t_pairs, 1), "seconds_train": round(t_train, 1),
           "ms_per_step": round(t_train / max(steps, 1) * 1e3, 2), "val_before": before, "val_after": after,
           "loss": hist, "params": net.num_params()}
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    print(json.dumps({k: v for k, v in rep.items() if k != "loss"}), flush=True)
    pr.close()


if __name__ == "__main__":
    main()
