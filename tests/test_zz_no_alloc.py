"""gk_trace_static, the allocation-free variant of gk_trace.

Its hits equal gk_trace's bit for bit (closest hit with instance, and occlusion) on a
single-level and a two-level scene, over two calls in a row of which the second is smaller
(the shared scratch array keeps the first call's contents); a pending instance update is
refitted in place; device memory in use does not change across a call of the maximum size;
larger batches are rejected with GK_ERR_LIMIT.
"""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes

pytestmark = pytest.mark.gpu


def cast(ctx, fn, rays, device, any_hit=False):
    import torch
    n = rays.shape[0]
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(device)
    h = torch.full((n, 4), float("nan"), dtype=torch.float32, device=device)
    inst = torch.full((n,), -7, dtype=torch.int32, device=device)
    torch.cuda.synchronize()
    print(f"{fn.__name__}: {n} rays, any_hit={any_hit}", flush=True)
    fn(r.data_ptr(), n, h.data_ptr(), inst.data_ptr(), any_hit=any_hit)
    ctx.sync()
    return h.cpu().numpy().view(np.uint32), inst.cpu().numpy().view(np.uint32)


def same(ctx, rays, device):
    for any_hit in (False, True):
        h0, i0 = cast(ctx, ctx.trace, rays, device, any_hit)
        h1, i1 = cast(ctx, ctx.trace_static, rays, device, any_hit)
        if any_hit:
            assert np.array_equal(h0[:, 3], h1[:, 3])
        else:
            assert np.array_equal(h0, h1)
            assert np.array_equal(i0, i1)


def test_trace_static_equals_trace_single_level(device):
    s = gk.procgen_cornell()
    big = scenes.random_rays(200_000, 0.02, 0.98, seed=31, tmin=1e-3, tmax=1000.0)
    small = scenes.random_rays(3_000, 0.02, 0.98, seed=32, tmin=1e-3, tmax=0.3)
    with gk.Context(64, 64) as ctx:
        ctx.upload(s)
        same(ctx, big, device)
        same(ctx, small, device)
        ctx.stats()  # no traversal stack overflowed


def test_trace_static_equals_trace_two_level_after_an_instance_update(device):
    b = scenes.Bricks()
    W, H = 480, 400
    c = scenes.BRICKS_CAM
    p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"])
    cam = po.primary_rays(p, W, H)
    with gk.Context(64, 64) as ctx:
        ctx.upload(b.scene)
        same(ctx, cam, device)
        ctx.update_instances(b.worlds(1.0 / 60.0))
        # the static call refits the TLAS first (in place); gk_trace then finds it clean
        h1, i1 = cast(ctx, ctx.trace_static, cam[::64], device)
        h0, i0 = cast(ctx, ctx.trace, cam[::64], device)
        assert np.array_equal(h0, h1) and np.array_equal(i0, i1)
        same(ctx, cam, device)
        same(ctx, cam[::97], device)
        ctx.stats()


def test_trace_static_allocates_nothing_and_rejects_oversized_batches(device):
    import torch
    s = gk.procgen_cornell()
    n = gk.TRACE_STATIC_MAX_RAYS
    rays = torch.from_numpy(scenes.random_rays(n + 1, 0.02, 0.98, seed=33, tmin=1e-3, tmax=1000.0)).to(device)
    hits = torch.zeros((n + 1, 4), dtype=torch.float32, device=device)
    inst = torch.zeros(n + 1, dtype=torch.int32, device=device)
    with gk.Context(64, 64) as ctx:
        ctx.upload(s)
        # load the kernels' code first (closest and any hit)
        for any_hit in (False, True):
            ctx.trace_static(rays.data_ptr(), 1000, hits.data_ptr(), inst.data_ptr(), any_hit=any_hit)
        ctx.sync()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for any_hit in (False, True):
            ctx.trace_static(rays.data_ptr(), n, hits.data_ptr(), inst.data_ptr(), any_hit=any_hit)
        ctx.sync()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)
        st = gk.lib().gk_trace_static(ctx.h, rays.data_ptr(), n + 1, hits.data_ptr(), inst.data_ptr(), 0)
        assert st == gk.GK_ERR_LIMIT
        assert gk.lib().gk_trace_static(ctx.h, rays.data_ptr(), 0, hits.data_ptr(), None, 0) == gk.GK_OK
        assert gk.lib().gk_trace_static(ctx.h, None, 5, hits.data_ptr(), None, 0) == gk.GK_ERR_ARG
        assert gk.lib().gk_trace_static(ctx.h, rays.data_ptr(), 5, hits.data_ptr(), None, 7) == gk.GK_ERR_ARG
