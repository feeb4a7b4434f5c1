"""Path-replay adjoint of the differentiable PT_MIS (512x512 Cornell tiles, 256 materials).

Checks:
  * the GPU primal equals the oracle's double-precision detached render;
  * every adjoint gradient component tested equals the central finite difference of the
    oracle's detached render (BSDF values perturbed, sampling held), which pins the
    derivative formulas without restating them;
  * albedo gradients equal finite differences of the GPU primal itself (sampling does not
    depend on albedo, so the fixed-seed primal is differentiable in it);
  * the gradient drives an optimisation towards a perturbed target (loss decreases).
"""
from __future__ import annotations

import copy

import numpy as np
import pytest

import gknextengine_amd as gk
from gknextengine_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def setup(W=48, H=48, spp=2, bounces=4):
    scene, nmat = scenes.cornell_tiles()
    cam = scenes.CORNELL_CAM
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.samples = spp
    p.bounces = p.max_bounces = bounces
    p.total_frames = 11
    p.global_seed = 5
    return scene, nmat, p


def test_primal_matches_oracle_and_adjoint_matches_detached_fd():
    import torch
    from gknextengine_amd import diff
    W = H = 48
    scene, nmat, p = setup(W, H)
    base = list(scene.materials)
    osc = po.OracleScene(scene)
    rng = np.random.default_rng(0)
    dL = rng.normal(size=(H, W, 3)).astype(np.float32)
    cpu_img = osc.diff_render(p, W, H, spp=p.samples, frame0=p.total_frames)
    with gk.Context(W, H) as ctx:
        ctx.upload(scene)
        params = diff.params_of(base).cuda().requires_grad_(True)
        img = diff.render(params, ctx, base, p, decorrelate=False)
        gpu_img = img.detach().cpu().numpy()
        close = np.all(np.abs(gpu_img - cpu_img) <= 1e-3 * (1 + np.abs(cpu_img)), axis=2)
        assert close.mean() > 0.98, close.mean()
        # A float path and a double path can part ways at a grazing hit or a shadow-ray
        # end point; the derivative check weighs only pixels whose paths agree.
        dL *= close[..., None]
        loss = (img * torch.from_numpy(dL).cuda()).sum()
        loss.backward()
        g = params.grad.cpu().numpy()
    # detached central differences on the oracle for a sample of parameters
    lit = np.nonzero(np.abs(g).sum(axis=1) > 0)[0]
    assert len(lit) > 20
    picks = [(int(m), k) for m in rng.choice(lit, 12, replace=False) for k in range(5)]
    h = 1e-4
    errs = []
    for m, k in picks:
        plus, minus = copy.deepcopy(base), copy.deepcopy(base)
        for mm, sign in ((plus, 1.0), (minus, -1.0)):
            x = mm[m]
            if k < 3:
                x.diffuse[k] += sign * h
            elif k == 3:
                x.fuzziness += sign * h
            else:
                x.metalness += sign * h
        ip = osc.diff_render(p, W, H, eval_materials=plus, spp=p.samples, frame0=p.total_frames)
        im = osc.diff_render(p, W, H, eval_materials=minus, spp=p.samples, frame0=p.total_frames)
        fd = float(((ip - im) / (2 * h) * dL).sum())
        errs.append((m, k, g[m, k], fd))
    gg = np.array([e[2] for e in errs])
    ff = np.array([e[3] for e in errs])
    scale = np.abs(ff).max()
    worst = int(np.argmax(np.abs(gg - ff)))
    assert np.abs(gg - ff).max() <= 2e-2 * scale + 1e-6, (errs[worst], list(zip(gg, ff)))
    assert np.corrcoef(gg, ff)[0, 1] > 0.999


def test_albedo_gradient_matches_gpu_finite_differences():
    import torch
    from gknextengine_amd import diff
    W = H = 48
    scene, nmat, p = setup(W, H, spp=2)
    base = list(scene.materials)
    rng = np.random.default_rng(1)
    dL = torch.from_numpy(rng.normal(size=(H, W, 3)).astype(np.float32)).cuda()
    with gk.Context(W, H) as ctx:
        ctx.upload(scene)
        params = diff.params_of(base).cuda().requires_grad_(True)
        (diff.render(params, ctx, base, p, decorrelate=False) * dL).sum().backward()
        g = params.grad.cpu().numpy()
        lit = np.nonzero(np.abs(g[:, :3]).sum(axis=1) > 0)[0]
        h = 1e-2
        for m in rng.choice(lit, 8, replace=False):
            for k in range(3):
                q = params.detach().clone()
                q[m, k] += h
                ip = diff.render(q, ctx, base, p)
                q[m, k] -= 2 * h
                im = diff.render(q, ctx, base, p)
                fd = float(((ip - im) / (2 * h) * dL).sum())
                assert fd == pytest.approx(float(g[m, k]), rel=2e-2, abs=2e-3 * abs(g[:, :3]).max())


def test_gradient_descent_reduces_image_loss():
    import torch
    from gknextengine_amd import diff
    W = H = 48
    scene, nmat, p = setup(W, H, spp=4)
    base = list(scene.materials)
    with gk.Context(W, H) as ctx:
        ctx.upload(scene)
        truth = diff.params_of(base).cuda()
        target = diff.render(truth, ctx, base, p).detach()
        gen = torch.Generator().manual_seed(3)
        start = truth.clone()
        start[:nmat, :3] = (start[:nmat, :3] + 0.15 * torch.randn(nmat, 3, generator=gen).cuda()).clamp(0.05, 0.95)
        params = start.clone().requires_grad_(True)
        opt = torch.optim.Adam([params], lr=0.02)
        losses = []
        for it in range(30):
            opt.zero_grad()
            loss = ((diff.render(params, ctx, base, p) - target) ** 2).mean()
            loss.backward()
            params.grad[nmat:] = 0  # the light is not a parameter
            params.grad[:, 3:] = 0  # fit albedo only
            opt.step()
            with torch.no_grad():
                params[:nmat, :3].clamp_(0.01, 0.99)
            losses.append(loss.item())
    assert losses[-1] < 0.3 * losses[0], losses


def test_backward_replays_its_own_forward_after_interleaved_renders():
    """Two renders (other materials, other seeds) before one backward: each backward
    replays the parameters and frame of its own forward, so the gradients equal those of
    the renders run one at a time."""
    import torch
    from gknextengine_amd import diff
    W = H = 32
    scene, nmat, p = setup(W, H, spp=2)
    q = gk.FrameParams.from_buffer_copy(p)
    q.total_frames, q.global_seed = 77, 9
    base = list(scene.materials)
    rng = np.random.default_rng(4)
    dL = torch.from_numpy(rng.normal(size=(H, W, 3)).astype(np.float32)).cuda()
    with gk.Context(W, H) as ctx:
        ctx.upload(scene)
        pa = diff.params_of(base).cuda()
        pb = pa.clone()
        pb[:nmat, :3] = (pb[:nmat, :3] * 0.7).clamp(0.01, 0.99)
        ga_ref, gb_ref = [], []
        for t, fr, out in ((pa, p, ga_ref), (pb, q, gb_ref)):
            x = t.clone().requires_grad_(True)
            (diff.render(x, ctx, base, fr) * dL).sum().backward()
            out.append(x.grad.clone())
        xa, xb = pa.clone().requires_grad_(True), pb.clone().requires_grad_(True)
        la = (diff.render(xa, ctx, base, p) * dL).sum()
        lb = (diff.render(xb, ctx, base, q) * dL).sum()
        lb.backward()
        la.backward()
        # (a work-group sums its gradient rows with LDS float atomics: last-bit differences)
        for g, r in ((xa.grad, ga_ref[0]), (xb.grad, gb_ref[0])):
            assert torch.allclose(g, r, rtol=1e-4, atol=1e-6 * float(r.abs().max())), (g - r).abs().max()
        assert not torch.allclose(ga_ref[0], gb_ref[0], rtol=1e-2)


def test_descent_recovers_perturbed_parameters_from_an_independent_target():
    """The differentiable render's set-up at reduced size: the target is a 2048-spp render with parameters
    perturbed by N(0, 0.05), drawn from seeds the optimisation never uses; plain
    loss.backward() (adjoint replayed with independent seeds) drives Adam back towards the
    target's albedos.

    Sizing: a material's albedo is recovered to within the noise of the pixels it covers.
    At 128x128 the well-lit tiles cover about 60 pixels each; the target's own noise at
    2048 spp moves their optimum by well under the 0.04 mean perturbation, and 16 spp per
    step keeps Adam's steps dominated by the signal until the error is a fraction of it."""
    import torch
    from gknextengine_amd import diff
    W = H = 128
    scene, nmat, p = setup(W, H, spp=16)
    base = list(scene.materials)
    with gk.Context(W, H) as ctx:
        ctx.upload(scene)
        theta = diff.params_of(base).cuda()
        gen = torch.Generator().manual_seed(12)
        truth = theta.clone()
        truth[:nmat, :3] = (truth[:nmat, :3] + 0.05 * torch.randn(nmat, 3, generator=gen).cuda()).clamp(0.02, 0.98)
        acc = torch.zeros((H, W, 3), device="cuda")
        for k in range(8):
            t = gk.FrameParams.from_buffer_copy(p)
            t.samples, t.total_frames, t.global_seed = 256, 1_000_000 + 256 * k, 99
            with torch.no_grad():
                acc += diff.render(truth, ctx, base, t)
        target = acc / 8
        # coverage: how strongly each material's albedo moves the image (64 spp)
        x = theta.clone().requires_grad_(True)
        c = gk.FrameParams.from_buffer_copy(p)
        c.samples, c.total_frames = 64, 500_000
        diff.render(x, ctx, base, c).sum().backward()
        cover = x.grad[:nmat, :3].abs().sum(dim=1)
        seen = cover > cover.max() * 0.2
        params = theta.clone().requires_grad_(True)
        opt = torch.optim.Adam([params], lr=0.005)
        err0 = (params[:nmat, :3] - truth[:nmat, :3]).abs()[seen].mean().item()
        losses = []
        for it in range(100):
            opt.zero_grad()
            f = gk.FrameParams.from_buffer_copy(p)
            f.total_frames = 100 + 16 * it
            loss = ((diff.render(params, ctx, base, f) - target) ** 2).mean()
            loss.backward()
            params.grad[nmat:] = 0
            params.grad[:, 3:] = 0
            opt.step()
            with torch.no_grad():
                params[:nmat, :3].clamp_(0.01, 0.99)
            losses.append(loss.item())
        err1 = (params[:nmat, :3] - truth[:nmat, :3]).abs()[seen].mean().item()
    print(f"albedo error over {int(seen.sum())} materials: {err0:.4f} -> {err1:.4f}; "
          f"loss {np.mean(losses[:5]):.3e} -> {np.mean(losses[-5:]):.3e}")
    assert err1 < 0.5 * err0, (err0, err1)
