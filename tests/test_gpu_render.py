"""Parity of the wavefront path tracer (gk_render) with the CPU oracle."""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes

pytestmark = pytest.mark.gpu


def cornell_params(W, H, frame=0, estimator=gk.ESTIMATOR_REF_V2, bounces=4, samples=1, prev=None):
    p = gk.make_frame_params(scenes.CORNELL_CAM["eye"], scenes.CORNELL_CAM["target"], W, H, fov=40.0,
                             frame=frame, prev=prev)
    p.estimator = estimator
    p.bounces = bounces
    p.max_bounces = bounces
    p.samples = samples
    p.global_seed = 0
    p.denoiser = gk.DENOISER_NONE
    return p


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def test_ref_v2_cornell_frames_match_oracle():
    """The Cornell box at 256x256, 16 frames x 1 spp, 4 bounces, REF_V2 estimator."""
    W = H = 256
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    moment = np.zeros((H, W), np.uint16)
    exact_frac = []
    acc_gpu = np.zeros((H, W, 3))
    acc_cpu = np.zeros((H, W, 3))
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        for f in range(16):
            p = cornell_params(W, H, frame=f)
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            o = osc.render_refv2(p, W, H, moment=moment)
            gd, gs = ctx.read("SINGLE_DIFFUSE"), ctx.read("SINGLE_SPECULAR")
            od, os_ = o["diffuse"].astype(np.float16), o["spec"].astype(np.float16)
            same = np.all(gd.view(np.uint16) == od.view(np.uint16), axis=2) & \
                np.all(gs.view(np.uint16) == os_.view(np.uint16), axis=2)
            exact_frac.append(same.mean())
            acc_gpu += gd[..., :3].astype(np.float64) + gs[..., :3]
            acc_cpu += o["diffuse"][..., :3] + o["spec"][..., :3]
            if f == 0:
                # G-buffer of the primary hit
                assert np.array_equal(ctx.read("OBJECTID_0"), o["objid"])
                assert np.array_equal(ctx.read("ALBEDO").view(np.uint16), o["albedo"].astype(np.float16).view(np.uint16))
                assert np.array_equal(ctx.read("NORMAL").view(np.uint16), o["normal"].astype(np.float16).view(np.uint16))
                assert np.array_equal(ctx.read("PREV_DEPTH").view(np.uint32), o["depth"].view(np.uint32))
                assert np.array_equal(ctx.read("MOTIONVECTOR").view(np.uint32), o["motion"].view(np.uint32))
        assert np.array_equal(ctx.read("MOTIONMOMENT"), moment)
    # per-pixel identical fp16 outputs for nearly every pixel; the rest differ where a
    # transcendental result differs in the last ulp between the device and libm
    assert min(exact_frac) > 0.98, exact_frac
    err = relmse(acc_gpu / 16, acc_cpu / 16)
    assert err < 1e-4, err


def test_ref_v2_deterministic():
    W = H = 128
    s = gk.procgen_cornell()
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        outs = []
        for _ in range(2):
            ctx.frame(cornell_params(W, H, frame=3))
            ctx.render()
            outs.append((ctx.read("SINGLE_DIFFUSE").copy(), ctx.read("SINGLE_SPECULAR").copy()))
    assert np.array_equal(outs[0][0].view(np.uint16), outs[1][0].view(np.uint16))
    assert np.array_equal(outs[0][1].view(np.uint16), outs[1][1].view(np.uint16))


def test_pt_mis_cornell_matches_oracle_same_seeds():
    W = H = 128
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    frames = 8
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        for f in range(frames):
            ctx.frame(cornell_params(W, H, frame=f, estimator=gk.ESTIMATOR_PT_MIS, bounces=4))
            ctx.render()
        ctx.sync()
        rad = ctx.read("RADIANCE")
    gpu = rad[..., :3] / rad[..., 3:4]
    o = osc.render_ptmis(cornell_params(W, H, estimator=gk.ESTIMATOR_PT_MIS, bounces=4), W, H, frame0=0,
                         frames=frames)
    cpu = o["rgb"]
    close = np.all(np.abs(gpu - cpu) <= 1e-4 * (1 + np.abs(cpu)), axis=2)
    assert close.mean() > 0.97, close.mean()
    assert relmse(gpu, cpu) < 1e-3


def test_pt_mis_white_furnace_closed_form():
    """Closed box, Lambertian albedo a on every face, emission Le on every face, B bounces:
    every path sees Le * sum_{k=0..B} a^k (emission is collected on emitters only at the
    first hit, so here every face is an emitter: L = Le)."""
    W = H = 32
    s = gk.procgen_cornell()
    # all faces emissive -> the primary hit sees Le exactly
    m = gk.Material()
    m.diffuse[:] = [2.0, 2.0, 2.0, 1.0]
    m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1
    m.material_model = gk.MAT_DIFFUSE_LIGHT
    s.materials = [m, m, m, m]
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.frame(cornell_params(W, H, estimator=gk.ESTIMATOR_PT_MIS))
        ctx.render()
        rad = ctx.read("RADIANCE")
        hit = ctx.read("PRIMARY_INST") != 0xFFFFFFFF
    assert np.allclose(rad[hit][:, :3], 2.0)


def test_material_sorted_shading_gives_the_same_image():
    """GK_FRAME_MATERIAL_SORT only reorders the shading of a queue; every path owns its
    state and adds its contributions in bounce order, so the frame is bit-identical."""
    W = H = 96
    s = gk.procgen_cornell()
    models = [gk.MAT_LAMBERTIAN, gk.MAT_METALLIC, gk.MAT_DIELECTRIC, gk.MAT_MIXTURE]
    for i, m in enumerate(s.materials):
        if m.material_model != gk.MAT_DIFFUSE_LIGHT:
            m.material_model = models[i % 4]
            m.metalness = 0.5 * (i % 3)
            m.fuzziness = 0.3
            m.refraction_index = m.refraction_index2 = 1.5
    out = []
    for flags in (0, gk.FRAME_MATERIAL_SORT):
        with gk.Context(W, H) as ctx:
            ctx.upload(s)
            for f in range(3):
                p = cornell_params(W, H, frame=f, estimator=gk.ESTIMATOR_PT_MIS, bounces=6)
                p.flags |= flags
                ctx.frame(p)
                ctx.render()
            ctx.sync()
            out.append((ctx.read("RADIANCE"), ctx.read("ALBEDO"), ctx.read("DENOISED")))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert out[0][0][..., :3].mean() > 0


@pytest.mark.parametrize("which", ["cornell", "bricks"])
def test_primary_hits_from_visibility_buffer(which):
    """GK_FRAME_PRIMARY_FROM_VISBUFFER reconstructs each primary hit from the triangle the
    visibility buffer names. Fed the buffer a traced frame wrote, it reproduces that frame's
    primary hits and radiance bit for bit; pixels cleared to 0 become misses."""
    W, H = 96, 64
    if which == "cornell":
        s = gk.procgen_cornell()
        eye, tgt, fov = scenes.CORNELL_CAM["eye"], scenes.CORNELL_CAM["target"], 40.0
    else:
        s = scenes.Bricks().scene
        from gknextengine_amd import scenes as gs
        eye, tgt, fov = gs.BRICKS_CAM["eye"], gs.BRICKS_CAM["target"], gs.BRICKS_CAM["fov"]
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        p = gk.make_frame_params(eye, tgt, W, H, fov=fov, frame=3)
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.bounces = p.max_bounces = 3
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        ref = {k: ctx.read(k) for k in ("PRIMARY_HIT", "PRIMARY_INST", "RADIANCE", "MINIGBUFFER", "ALBEDO")}
        assert (ref["MINIGBUFFER"] != 0).mean() > 0.3
        ctx.reset()
        q = gk.FrameParams.from_buffer_copy(p)
        q.flags |= gk.FRAME_PRIMARY_FROM_VISBUFFER
        ctx.frame(q)
        ctx.render()
        ctx.sync()
        for k in ("PRIMARY_HIT", "PRIMARY_INST", "RADIANCE", "MINIGBUFFER", "ALBEDO"):
            # bitwise: a miss stores prim 0xffffffff, a NaN bit pattern, in PRIMARY_HIT.w
            assert np.array_equal(ctx.read(k).view(np.uint8), ref[k].view(np.uint8)), k
        # the caller's buffer: clear a band to 0 -> misses there, hits elsewhere unchanged
        vis = ref["MINIGBUFFER"].copy()
        vis[10:20] = 0
        ctx.write("MINIGBUFFER", vis)
        ctx.reset()
        ctx.frame(q)
        ctx.render()
        ctx.sync()
        inst = ctx.read("PRIMARY_INST")
        assert np.all(inst[10:20] == 0xFFFFFFFF)
        assert np.array_equal(inst[20:], ref["PRIMARY_INST"][20:])
        assert np.array_equal(ctx.read("PRIMARY_HIT")[20:].view(np.uint8), ref["PRIMARY_HIT"][20:].view(np.uint8))


def test_ref_v2_moving_pixels_at_the_sample_limit_match_oracle():
    """samples = 1024, gk_frame's limit, on pixels whose object moved: Core.PathTracing
    multiplies their samples by 4 (Core.PathTracing.comp.slang:25-28), so each such pixel
    averages 4096 samples. The per-path sample counter holds that count (the test failed
    with Inf/NaN when the count wrapped to 0), and the frame equals the oracle's."""
    W = H = 16
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    moved = np.eye(4, dtype=np.float32)
    moved[0, 3] = 0.05
    moment = np.zeros((H, W), np.uint16)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        p0 = cornell_params(W, H, frame=0, bounces=2)
        ctx.frame(p0)
        ctx.render()
        osc.render_refv2(p0, W, H, moment=moment)
        ctx.update_instances(moved[None])
        osc.update_instances(moved[None])
        p1 = cornell_params(W, H, frame=1, bounces=2, samples=1024, prev=p0)
        ctx.frame(p1)
        ctx.render()
        ctx.sync()
        gd, gs = ctx.read("SINGLE_DIFFUSE"), ctx.read("SINGLE_SPECULAR")
        gm = ctx.read("MOTIONMOMENT")
    o = osc.render_refv2(p1, W, H, moment=moment)
    assert np.array_equal(gm, moment)
    assert (gm > 0).mean() > 0.5  # most pixels moved: 4096 samples each
    assert np.isfinite(gd.astype(np.float32)).all() and np.isfinite(gs.astype(np.float32)).all()
    od, os_ = o["diffuse"].astype(np.float16), o["spec"].astype(np.float16)
    same = np.all(gd.view(np.uint16) == od.view(np.uint16), axis=2) & np.all(gs.view(np.uint16) == os_.view(np.uint16), axis=2)
    assert same.mean() > 0.9, same.mean()
    err = relmse(gd[..., :3].astype(np.float64) + gs[..., :3], o["diffuse"][..., :3].astype(np.float64) + o["spec"][..., :3])
    assert err < 1e-4, err


@pytest.mark.parametrize("estimator", [gk.ESTIMATOR_REF_V2, gk.ESTIMATOR_PT_MIS])
def test_thin_lens_frames_match_oracle(estimator):
    """Aperture > 0 (FHardwarePrimaryRayCaster's thin lens, Shading.slang:283-300): the lens
    ray is each path's first draw. Primary hits equal the oracle's lens rays' hits bit for
    bit; REF_V2 outputs are fp16-identical on nearly every pixel and PT_MIS (2 spp: the
    second sample draws its own lens point) agrees with the oracle with the same seeds."""
    W = H = 96
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    p = cornell_params(W, H, frame=3, estimator=estimator, samples=2 if estimator == gk.ESTIMATOR_PT_MIS else 1)
    p.aperture, p.focus_distance = 0.08, 1.9
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        hit = ctx.read("PRIMARY_HIT").reshape(-1, 4)
        oh, _ = osc.trace(po.lens_rays(p, W, H))
        assert np.array_equal(hit.view(np.uint32), oh.view(np.uint32))
        pin, _ = osc.trace(po.primary_rays(p, W, H))
        assert not np.array_equal(oh.view(np.uint32), pin.view(np.uint32))  # the lens does move the rays
        if estimator == gk.ESTIMATOR_REF_V2:
            o = osc.render_refv2(p, W, H)
            gd = ctx.read("SINGLE_DIFFUSE")
            same = np.all(gd.view(np.uint16) == o["diffuse"].astype(np.float16).view(np.uint16), axis=2)
            assert same.mean() > 0.98, same.mean()
        else:
            rad = ctx.read("RADIANCE")
            gpu = rad[..., :3] / rad[..., 3:4]
            cpu = osc.render_ptmis(p, W, H, frame0=p.total_frames, frames=1)["rgb"]
            close = np.all(np.abs(gpu - cpu) <= 1e-4 * (1 + np.abs(cpu)), axis=2)
            assert close.mean() > 0.97, close.mean()


def test_city_visibility_buffer_ids_and_hits_from_them():
    """The 1080p city's 0.92 M triangles span 15 sections of 65535. A traced frame's
    GK_RT_MINIGBUFFER holds the reference raster pass's ids (proxy per instance and
    section, triangle within the section), equal to the oracle's packing of the frame's
    primary hits, including triangles above 2^17 that a plain 17-bit triangle field would
    alias. Fed back, the ids reproduce the traced primary hits bit for bit, and those hit
    points are get_material_dataV2's world-space positions (oracle) to float rounding."""
    city = gk.procgen_city()
    W, H = 1920, 1080
    c = scenes.CITY_CAM
    p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"], frame=2)
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = p.max_bounces = 1
    osc = po.OracleScene(city)
    with gk.Context(W, H) as ctx:
        ctx.upload(city)
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        vis = ctx.read("MINIGBUFFER")
        hit = ctx.read("PRIMARY_HIT").reshape(-1, 4)
        inst = ctx.read("PRIMARY_INST").reshape(-1)
        prim = hit[:, 3].view(np.uint32)
        found = prim != 0xFFFFFFFF
        assert prim[found].max() >= 1 << 17
        np.testing.assert_array_equal(vis.reshape(-1), osc.vis_pack(hit, np.where(found, inst, 0)))
        ctx.reset()
        q = gk.FrameParams.from_buffer_copy(p)
        q.flags |= gk.FRAME_PRIMARY_FROM_VISBUFFER
        ctx.frame(q)
        ctx.render()
        ctx.sync()
        assert np.array_equal(ctx.read("PRIMARY_HIT").reshape(-1, 4).view(np.uint32), hit.view(np.uint32))
    rays = po.primary_rays(p, W, H)
    sub = np.nonzero(found)[0][::11]
    ri, rt, _, pos = osc.vis_decode(vis.reshape(-1)[sub], rays[sub])
    np.testing.assert_array_equal(rt, prim[sub])
    x = rays[sub, :3] + hit[sub, :1] * rays[sub, 4:7]
    np.testing.assert_allclose(pos, x, atol=2e-3, rtol=1e-4)


def test_ref_v2_two_pipelines_give_the_single_pipeline_frame():
    """REF_V2 frames of 2^18 paths or more run as two path pipelines (even and odd 8x8 tiles),
    each ending its bounce loop when its own queue is empty. On 4096 moving bricks at
    640x480 under a sky, 2 samples, three frames with instance motion: every target bit-identical to GK_FRAME_SINGLE_PIPELINE's, same ray count."""
    from gknextengine_amd import scenes as gs
    b = scenes.Bricks()
    W, H = 640, 480
    names = [t for t in gk.TARGETS if t not in ("RADIANCE64",)]
    outs = []
    for flags in (gk.FRAME_SINGLE_PIPELINE, 0):
        with gk.Context(W, H) as ctx:
            ctx.upload(b.scene)
            for f in range(3):
                if f:
                    ctx.update_instances(b.worlds(f / 60.0))
                p = gk.make_frame_params(gs.BRICKS_CAM["eye"], gs.BRICKS_CAM["target"], W, H,
                                         fov=gs.BRICKS_CAM["fov"], frame=f)
                p.estimator = gk.ESTIMATOR_REF_V2
                p.bounces = p.max_bounces = 4
                p.samples = 2
                p.has_sky = 1
                p.sky_color[:] = [0.5, 0.6, 0.8, 1.0]
                p.denoiser = gk.DENOISER_BILATERAL
                p.flags |= flags | gk.FRAME_TLAS_REBUILD
                ctx.frame(p)
                ctx.render()
            ctx.sync()
            outs.append(({t: ctx.read(t).copy() for t in names}, ctx.stats().rays_traced))
    (a, ra), (c, rc) = outs
    assert ra == rc, (ra, rc)
    for t in names:
        assert np.array_equal(a[t].view(np.uint8), c[t].view(np.uint8)), t
    print(f"moving pixels {(a['MOTIONMOMENT'] != 0).mean():.3f}, rays {ra}")
    assert ra > 2 * W * H and a["RADIANCE"][..., :3].max() > 0.1
