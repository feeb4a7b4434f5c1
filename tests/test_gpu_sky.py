"""The sky image on the device (gk_sky_upload): every miss of REF_V2, PT_MIS and the
differentiable primal sees SampleIBLV2 (Shading.slang:194-199), the probe bake sees
SampleIBLRough of the image's ProjectHdrToSh harmonics (Texture.cpp:180-253,
AmbientCube.slang:515); all against the oracle on a floor with four bricks in the open."""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes

pytestmark = pytest.mark.gpu

ROT = 0.3
COL = (1.5, 1.5, 1.5, 1.5)   # SkyIntensity 1.5


def sky_params(W, H, estimator, frame=0, bounces=4, samples=1):
    c = scenes.SKY_CAM
    p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"], frame=frame)
    p.estimator = estimator
    p.bounces = p.max_bounces = bounces
    p.samples = samples
    p.has_sky = 1
    p.sky_color[:] = COL
    p.sky_rotation = ROT
    return p


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def test_ref_v2_under_a_sky_image_matches_oracle():
    W, H = 160, 96
    s = scenes.sky_scene()
    img = scenes.sky_image(256, 128, seed=2)
    osc = po.OracleScene(s)
    osc.set_sky(img)
    moment = np.zeros((H, W), np.uint16)
    exact, acc_g, acc_c = [], 0.0, 0.0
    with gk.Context(W, H) as ctx:
        ctx.sky_upload(img)
        ctx.upload(s)                       # the sky image survives a scene upload
        for f in range(4):
            p = sky_params(W, H, gk.ESTIMATOR_REF_V2, frame=f)
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            o = osc.render_refv2(p, W, H, moment=moment)
            gd, gs = ctx.read("SINGLE_DIFFUSE"), ctx.read("SINGLE_SPECULAR")
            od, os_ = o["diffuse"].astype(np.float16), o["spec"].astype(np.float16)
            same = np.all(gd.view(np.uint16) == od.view(np.uint16), axis=2) & \
                np.all(gs.view(np.uint16) == os_.view(np.uint16), axis=2)
            exact.append(same.mean())
            acc_g = acc_g + gd[..., :3].astype(np.float64) + gs[..., :3]
            acc_c = acc_c + o["diffuse"][..., :3] + o["spec"][..., :3]
            if f == 0:
                miss = o["objid"] == 65535
                assert 0.2 < miss.mean() < 0.8, miss.mean()
                # primary misses are the image's texels along the camera rays
                rays = po.primary_rays(p, W, H)
                sky = osc.sky_radiance(p, rays[:, 4:7]).reshape(H, W, 4)
                assert np.array_equal(gd[miss].view(np.uint16), sky[miss].astype(np.float16).view(np.uint16))
        # the same frame under a constant sky is a different image
        ctx.sky_upload(None)
        ctx.frame(sky_params(W, H, gk.ESTIMATOR_REF_V2, frame=3))
        ctx.render()
        ctx.sync()
        flat = ctx.read("SINGLE_DIFFUSE")
    assert min(exact) > 0.98, exact
    assert relmse(acc_g / 4, acc_c / 4) < 1e-4
    assert np.abs(flat[..., :3].astype(np.float32) - gd[..., :3]).max() > 0.5


def test_pt_mis_under_a_sky_image_matches_oracle():
    W, H = 128, 80
    s = scenes.sky_scene()
    img = scenes.sky_image(256, 128, seed=5)
    osc = po.OracleScene(s)
    osc.set_sky(img)
    frames = 4
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.sky_upload(img)
        for f in range(frames):
            ctx.frame(sky_params(W, H, gk.ESTIMATOR_PT_MIS, frame=f))
            ctx.render()
        ctx.sync()
        rad = ctx.read("RADIANCE")
    gpu = rad[..., :3] / rad[..., 3:4]
    cpu = osc.render_ptmis(sky_params(W, H, gk.ESTIMATOR_PT_MIS), W, H, frame0=0, frames=frames)["rgb"]
    close = np.all(np.abs(gpu - cpu) <= 1e-4 * (1 + np.abs(cpu)), axis=2)
    assert close.mean() > 0.97, close.mean()
    assert relmse(gpu, cpu) < 1e-3
    osc.set_sky(None)                # the image does light the scene
    flat = osc.render_ptmis(sky_params(W, H, gk.ESTIMATOR_PT_MIS), W, H, frame0=0, frames=frames)["rgb"]
    assert relmse(flat, cpu) > 1e-2


def test_probe_bake_sees_the_sky_harmonics():
    s = scenes.sky_scene()
    img = scenes.sky_image(128, 64, seed=6)
    osc = po.OracleScene(s)
    osc.set_sky(img)
    grid = (24, 8, 0.25, (-3.0, -0.125, -3.0))
    G = po.OracleProbes(*grid, gk.CUBE_DTYPE, gk.VOXEL_DTYPE)
    G.voxelize(osc)
    p = sky_params(32, 32, gk.ESTIMATOR_REF_V2)
    p.sky_color[:] = [200.0] * 4     # probe light is stored as RGB10 over [0, 512]
    with gk.Context(32, 32) as ctx:
        ctx.upload(s)
        ctx.sky_upload(img)
        ctx.frame(p)
        ctx.probes_init(*grid[:3], grid[3])
        ctx.probes_voxelize()
        for it in range(2):
            ctx.probes_bake()
            G.bake(osc, p, None)
            cubes, vox = ctx.probes_read()
            assert np.array_equal(vox["age"], G.vox["age"])
            diff = np.zeros(len(cubes), bool)
            for f in gk.CUBE_FIELDS:
                diff |= cubes[f] != G.cubes[f]
            print(f"pass {it}: {int(diff.sum())} of {len(cubes)} probes differ", flush=True)
            assert not diff.any()
            G.cubes[:] = cubes
        # against a constant sky of the image's mean the baked light differs
        ctx.sky_upload(None)
        p2 = sky_params(32, 32, gk.ESTIMATOR_REF_V2)
        p2.sky_color[:] = [200.0 * float(x) for x in img[..., :3].mean((0, 1))] + [200.0]
        ctx.frame(p2)
        ctx.probes_bake()
        flat, _ = ctx.probes_read()
    act = (vox["dist_gg_z01"] & 0xFF) < 8
    assert act.sum() > 500
    # sky light is direct light: the +y faces' direct term
    assert (cubes["pos_y_d"][act] != 0).mean() > 0.5
    assert (flat["pos_y_d"][act] != cubes["pos_y_d"][act]).mean() > 0.5


def test_differentiable_primal_sees_the_sky_image():
    from gknextengine_amd import diff
    from gknextengine_amd.scenes import material
    W, H = 48, 32
    s = scenes.sky_scene()
    # no glass: a float and a double path part ways at its Fresnel choices
    s.materials[3] = material((0.9, 0.9, 0.9), gk.MAT_METALLIC, rough=0.2)
    img = scenes.sky_image(128, 64, seed=7)
    osc = po.OracleScene(s)
    osc.set_sky(img)
    p = sky_params(W, H, gk.ESTIMATOR_PT_MIS, samples=2)
    p.total_frames = 4
    base = list(s.materials)
    cpu = osc.diff_render(p, W, H, spp=p.samples, frame0=p.total_frames)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.sky_upload(img)
        params = diff.params_of(base).cuda()
        gpu = diff.render(params, ctx, base, p, decorrelate=False).cpu().numpy()
    close = np.all(np.abs(gpu - cpu) <= 1e-3 * (1 + np.abs(cpu)), axis=2)
    assert close.mean() > 0.98, close.mean()
    assert relmse(gpu, cpu) < 1e-3


def test_sky_upload_rejects_bad_images_and_keeps_the_last_good_one():
    W = H = 32
    s = scenes.sky_scene()
    img = scenes.sky_image(64, 32, seed=8)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.sky_upload(img)
        p = sky_params(W, H, gk.ESTIMATOR_REF_V2)
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        ref = ctx.read("SINGLE_DIFFUSE").copy()
        bad = img.copy()
        bad[3, 4, 1] = np.nan
        with pytest.raises(gk.GkError) as e:
            ctx.sky_upload(bad)
        assert e.value.status == gk.GK_ERR_ARG
        assert gk.lib().gk_sky_upload(ctx.h, img.ctypes.data, 16385, 1) == gk.GK_ERR_LIMIT
        assert gk.lib().gk_sky_upload(ctx.h, None, 64, 32) == gk.GK_ERR_ARG
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        assert np.array_equal(ctx.read("SINGLE_DIFFUSE").view(np.uint16), ref.view(np.uint16))
