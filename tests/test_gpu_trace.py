"""Parity of the HIP traversal (gk_trace) with the oracle and the reference's tinybvh."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes
from tests.test_oracle import _compare_hits, _u32

GOLDEN = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu


def gpu_trace(ctx, rays, device, any_hit=False):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(device)
    h = torch.zeros((rays.shape[0], 4), dtype=torch.float32, device=device)
    inst = torch.zeros(rays.shape[0], dtype=torch.int32, device=device)
    torch.cuda.synchronize()
    ctx.trace(r.data_ptr(), rays.shape[0], h.data_ptr(), inst.data_ptr(), any_hit=any_hit)
    ctx.sync()
    return h.cpu().numpy(), inst.cpu().numpy().view(np.uint32)


def test_cornell_primary_bit_exact(device):
    s = gk.procgen_cornell()
    W = H = 256
    p = gk.make_frame_params(scenes.CORNELL_CAM["eye"], scenes.CORNELL_CAM["target"], W, H, fov=40.0)
    rays = po.primary_rays(p, W, H)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        h, inst = gpu_trace(ctx, rays, device)
    o, _ = po.OracleScene(s).trace(rays)
    # identical to the oracle bit for bit (same canonical tie-break)
    assert np.array_equal(_u32(h), _u32(o))
    assert np.all(inst[_u32(h[:, 3]) != 0xFFFFFFFF] == 0)
    # and to the reference's tinybvh up to ties
    g = np.load(GOLDEN / "cornell_tinybvh.npz")
    sub = h.reshape(H, W, 4)[::4, ::4].reshape(-1, 4)
    _compare_hits(sub, g["hits"], "gpu vs tinybvh")


def test_cornell_random_and_occlusion(device):
    s = gk.procgen_cornell()
    rays = scenes.random_rays(100000, 0.02, 0.98, seed=21, tmin=1e-3, tmax=1000.0)
    osc = po.OracleScene(s)
    with gk.Context(64, 64) as ctx:
        ctx.upload(s)
        h, _ = gpu_trace(ctx, rays, device)
        occ_rays = scenes.random_rays(100000, 0.02, 0.98, seed=22, tmax=0.4)
        ho, _ = gpu_trace(ctx, occ_rays, device, any_hit=True)
    o, _ = osc.trace(rays)
    assert np.array_equal(_u32(h), _u32(o))
    oo, _ = osc.trace(occ_rays, any_hit=True)
    assert np.array_equal(_u32(ho[:, 3]), _u32(oo[:, 3]))
    g = np.load(GOLDEN / "cornell_tinybvh.npz")
    with gk.Context(64, 64) as ctx:
        ctx.upload(s)
        hg, _ = gpu_trace(ctx, g["occ_rays"], device, any_hit=True)
    assert np.array_equal(_u32(hg[:, 3]), g["occ"])


def test_city_camera_and_random_rays(device):
    c = gk.procgen_city()
    g = np.load(GOLDEN / "city_tinybvh.npz")
    p = gk.make_frame_params(scenes.CITY_CAM["eye"], scenes.CITY_CAM["target"], 1920, 1080,
                             fov=scenes.CITY_CAM["fov"])
    cam = po.primary_rays(p, 1920, 1080)[::16]
    rr = scenes.random_rays(60000, -30.0, 30.0, seed=8, tmin=1e-3, tmax=1000.0)
    rr[:, 1] = np.abs(rr[:, 1]) * 0.3
    rays = np.concatenate([g["rays"], cam, rr])
    with gk.Context(64, 64) as ctx:
        ctx.upload(c)
        h, _ = gpu_trace(ctx, rays, device)
        st = ctx.bvh_stats()
    assert st.blas_nodes > 0 and st.sah_cost > 0
    o, _ = po.OracleScene(c).trace(rays)
    assert np.array_equal(_u32(h), _u32(o))
    _compare_hits(h[: len(g["rays"])], g["hits"], "city gpu vs tinybvh")


def test_small_launches_match_large(device):
    """A launch of at most 8 rays per quad of the grid (524288 rays on 256 CUs) runs the
    four-lanes-per-ray traversal, a larger one the one-lane-per-ray traversal: the same rays
    give the same bits either way, closest and any hit, and the small launches equal the
    oracle."""
    c = gk.procgen_city()
    rr = scenes.random_rays(1200000, -30.0, 30.0, seed=31, tmin=1e-3, tmax=1000.0)
    rr[:, 1] = np.abs(rr[:, 1]) * 0.3
    occ = rr.copy()
    occ[:, 7] = 4.0
    with gk.Context(64, 64) as ctx:
        ctx.upload(c)
        big, ibig = gpu_trace(ctx, rr, device)
        small, ismall = gpu_trace(ctx, rr[:40000], device)
        one, _ = gpu_trace(ctx, rr[7:8], device)
        obig, _ = gpu_trace(ctx, occ, device, any_hit=True)
        osmall, _ = gpu_trace(ctx, occ[:40000], device, any_hit=True)
    assert np.array_equal(_u32(small), _u32(big[:40000]))
    assert np.array_equal(ismall, ibig[:40000])
    assert np.array_equal(_u32(one), _u32(big[7:8]))
    assert np.array_equal(_u32(osmall[:, 3]), _u32(obig[:40000, 3]))
    hit = _u32(small[:, 3]) != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.999 and 0.05 < (_u32(osmall[:, 3]) == 1).mean() < 0.95
    o, _ = po.OracleScene(c).trace(rr[:40000])
    assert np.array_equal(_u32(small), _u32(o))


def test_two_level_small_launches_match_large(device):
    """Two-level scene: a small launch (quad traversal, instances entered per quad) and a
    large one (one lane per ray) give the same bits, instance ids included, closest and
    any hit; the small launch equals the oracle."""
    b = scenes.Bricks()
    rr = scenes.random_rays(1200000, -14.0, 14.0, seed=37, tmin=1e-3, tmax=1000.0)
    occ = rr.copy()
    occ[:, 7] = 3.0
    with gk.Context(64, 64) as ctx:
        ctx.upload(b.scene)
        big, ibig = gpu_trace(ctx, rr, device)
        small, ismall = gpu_trace(ctx, rr[:30000], device)
        one, ione = gpu_trace(ctx, rr[5:6], device)
        obig, _ = gpu_trace(ctx, occ, device, any_hit=True)
        osmall, _ = gpu_trace(ctx, occ[:30000], device, any_hit=True)
    assert np.array_equal(_u32(small), _u32(big[:30000]))
    assert np.array_equal(ismall, ibig[:30000])
    assert np.array_equal(_u32(one), _u32(big[5:6])) and np.array_equal(ione, ibig[5:6])
    assert np.array_equal(_u32(osmall[:, 3]), _u32(obig[:30000, 3]))
    hit = _u32(small[:, 3]) != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.999 and 0.05 < (_u32(osmall[:, 3]) == 1).mean() < 0.95
    o, oi = po.OracleScene(b.scene).trace(rr[:30000])
    assert np.array_equal(_u32(small), _u32(o))
    assert np.array_equal(ismall[hit], oi[hit])


def test_tlas_bricks_bit_exact(device):
    b = scenes.Bricks()
    g = np.load(GOLDEN / "bricks_tinybvh.npz")
    osc = po.OracleScene(b.scene)
    rays = np.concatenate([g["rays"], scenes.random_rays(30000, -14.0, 14.0, seed=12, tmin=1e-3, tmax=1000.0)])
    with gk.Context(64, 64) as ctx:
        ctx.upload(b.scene)
        h, inst = gpu_trace(ctx, rays, device)
        o, oi = osc.trace(rays)
        hit = _u32(o[:, 3]) != 0xFFFFFFFF
        assert np.array_equal(_u32(h), _u32(o))
        assert np.array_equal(inst[hit], oi[hit])
        # reference TLAS (tinybvh IntersectTLAS) up to ties
        n = len(g["rays"])
        same = (_u32(h[:n, 3]) == _u32(g["hits"][:, 3])) & (inst[:n] == g["inst"])
        assert same.mean() > 0.99
        # after a refit with moved instances
        w1 = b.worlds(0.5)
        ctx.update_instances(w1)
        osc.update_instances(w1)
        h2, i2 = gpu_trace(ctx, rays, device)
    o2, oi2 = osc.trace(rays)
    assert np.array_equal(_u32(h2), _u32(o2))
    hit2 = _u32(o2[:, 3]) != 0xFFFFFFFF
    assert np.array_equal(i2[hit2], oi2[hit2])


def test_face_plane_rays_bit_exact(device):
    """Origins on a face plane, directions parallel to it: the kernel's box test must
    cull a box that ends on that plane exactly as tinybvh's (plane - O) * rD does."""
    s = gk.procgen_cornell()
    rays = scenes.face_plane_rays()
    with gk.Context(8, 8) as ctx:
        ctx.upload(s)
        h, _ = gpu_trace(ctx, rays, device)
    o, _ = po.OracleScene(s).trace(rays)
    assert np.array_equal(_u32(h), _u32(o))


def test_empty_and_edge_inputs(device):
    s = gk.procgen_cornell()
    with gk.Context(8, 8) as ctx:
        ctx.upload(s)
        ctx.trace(0, 0, 0)  # n = 0 is a no-op
        ctx.sync()
        # rays along the axes, through edges and from outside
        rays = np.zeros((6, 8), np.float32)
        rays[:, 3] = 0.0
        rays[:, 7] = 1e30
        rays[0, :3] = [0.5, 0.5, -1.0]; rays[0, 4:7] = [0, 0, 1]
        rays[1, :3] = [0.5, 0.5, 0.5]; rays[1, 4:7] = [1, 0, 0]
        rays[2, :3] = [0.5, 2.0, 0.5]; rays[2, 4:7] = [0, -1, 0]
        rays[3, :3] = [0.5, 0.5, 0.5]; rays[3, 4:7] = [0, 0, -1]   # leaves through the open face
        rays[4, :3] = [0.0, 0.0, -1.0]; rays[4, 4:7] = [0, 0, 1]   # along an edge of the box
        rays[5, :3] = [0.3, 0.3, 0.3]; rays[5, 4:7] = [0, 1, 0]; rays[5, 7] = 0.0  # empty interval
        h, _ = gpu_trace(ctx, rays, device)
    o, _ = po.OracleScene(s).trace(rays)
    assert np.array_equal(_u32(h), _u32(o))
    assert _u32(h[5, 3]) == 0xFFFFFFFF


def test_tlas_rebuild_and_refit_alternate_bit_exact(device):
    """GK_FRAME_TLAS_REBUILD builds a new TLAS on the device (LBVH + single-work-group 4-wide
    collapse that also writes the refit order); primary hits equal the oracle's (whose TLAS
    is rebuilt) bit for bit, whether frames rebuild or refit, in any order. After the bricks
    drift apart for two seconds a refitted tree's SAH cost is well above a rebuilt one's."""
    b = scenes.Bricks()
    osc = po.OracleScene(b.scene)
    W, H = 160, 96
    c = scenes.BRICKS_CAM
    sah = {}
    with gk.Context(W, H) as ctx:
        ctx.upload(b.scene)
        sah["built"] = ctx.bvh_stats().tlas_sah_cost
        for k, (t, rebuild) in enumerate([(0.25, True), (0.5, False), (1.0, True), (1.5, True), (2.0, False),
                                          (2.0, True)]):
            w = b.worlds(t)
            ctx.update_instances(w)
            osc.update_instances(w)
            p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"], frame=k)
            p.estimator = gk.ESTIMATOR_PT_MIS
            p.bounces = p.max_bounces = 1
            p.flags |= gk.FRAME_SKIP_POST | (gk.FRAME_TLAS_REBUILD if rebuild else 0)
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            hit = ctx.read("PRIMARY_HIT").reshape(-1, 4)
            inst = ctx.read("PRIMARY_INST").reshape(-1)
            oh, oi = osc.trace(po.primary_rays(p, W, H))
            assert np.array_equal(_u32(hit), _u32(oh)), (t, rebuild)
            ok = _u32(oh[:, 3]) != 0xFFFFFFFF
            assert np.array_equal(inst[ok], oi[ok])
            sah[(t, rebuild)] = ctx.bvh_stats().tlas_sah_cost
        # random rays through gk_trace after a rebuild use the same tree
        rays = scenes.random_rays(20000, -14.0, 14.0, seed=13, tmin=1e-3, tmax=1000.0)
        h, i = gpu_trace(ctx, rays, device)
    o, oi = osc.trace(rays)
    assert np.array_equal(_u32(h), _u32(o))
    # the tree of t = 0 refitted to t = 2 s against the tree rebuilt there
    with gk.Context(W, H) as ctx:
        ctx.upload(b.scene)
        ctx.update_instances(b.worlds(2.0))
        h, _ = gpu_trace(ctx, rays, device)  # refits
        sah["refit from t=0"] = ctx.bvh_stats().tlas_sah_cost
    assert np.array_equal(_u32(h), _u32(o))
    print({str(k): round(v, 3) for k, v in sah.items()})
    assert sah["refit from t=0"] > 1.3 * sah[(2.0, True)]


def test_closest_hit_does_not_depend_on_ray_order(device):
    """The city camera moved by (0.15, 0, 0.075): pixel (417, 551) hits the shared edge of
    two coplanar wall triangles, 168450 and 168459, at the same t, and the triangle test
    accepts both (float barycentrics overlap by about 7e-6). Triangle 168450's box entry lies
    16 ulps beyond that t, so a 2^-20 box widening culled its leaf whenever 168459 was found
    first, and the answer depended on which rays shared a wave. With 2^-16 every order
    returns the lower primitive: the frame's 2 M primary rays traced in order and in two
    permutations give the same hits bit for bit."""
    import torch
    s = gk.procgen_city(64, 4, 256, 64, 2024)
    W, H = 1920, 1080
    eye = tuple(e + 0.15 * d for e, d in zip(scenes.CITY_CAM["eye"], (1.0, 0.0, 0.5)))
    p = gk.make_frame_params(eye, scenes.CITY_CAM["target"], W, H, fov=scenes.CITY_CAM["fov"], frame=3)
    rays = po.primary_rays(p, W, H).astype(np.float32)
    n = rays.shape[0]
    rng = np.random.default_rng(5)
    res = []
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        for k in range(3):
            perm = np.arange(n) if k == 0 else rng.permutation(n)
            r = torch.from_numpy(np.ascontiguousarray(rays[perm])).to(device)
            h = torch.zeros((n, 4), dtype=torch.float32, device=device)
            inst = torch.zeros(n, dtype=torch.int32, device=device)
            torch.cuda.synchronize()
            ctx.trace(r.data_ptr(), n, h.data_ptr(), inst.data_ptr())
            ctx.sync()
            out = np.empty((n, 4), np.float32)
            out[perm] = h.cpu().numpy()
            res.append(out)
    i = 551 * W + 417
    assert _u32(res[0][i:i + 1, 3])[0] == 168450, res[0][i]
    for k in (1, 2):
        assert np.array_equal(_u32(res[k]), _u32(res[0])), int((_u32(res[k]) != _u32(res[0])).any(1).sum())
