"""Config 3 (dynamic instanced bricks): TLAS refit, motion vectors, reprojection and the
bilateral denoiser against the oracle restatements."""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes
from tests.test_oracle import _u32

pytestmark = pytest.mark.gpu


def frame_params(W, H, f, prev=None, denoiser=gk.DENOISER_BILATERAL):
    c = scenes.BRICKS_CAM
    p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"], frame=f, prev=prev)
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = 4
    p.max_bounces = 4
    p.samples = 1
    p.has_sky = 1
    p.sky_color[:] = [0.6, 0.7, 0.9, 1.0]
    p.temporal_frames = 16
    p.denoiser = denoiser
    return p


def h2f(a):
    return a.astype(np.float32)


def test_dynamic_bricks_refit_motion_reproject_bilateral():
    W, H = 320, 176
    b = scenes.Bricks()
    osc = po.OracleScene(b.scene)
    prev = None
    hist = None
    with gk.Context(W, H) as ctx:
        ctx.upload(b.scene)
        for f in range(4):
            t = f / 60.0
            w = b.worlds(t)
            if f > 0:
                ctx.update_instances(w)
                osc.update_instances(w)
            p = frame_params(W, H, f, prev)
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            # primary hits after the device refit equal the oracle's rebuilt TLAS
            hit = ctx.read("PRIMARY_HIT").reshape(-1, 4)
            inst = ctx.read("PRIMARY_INST").reshape(-1)
            oh, oi = osc.trace(po.primary_rays(p, W, H))
            assert np.array_equal(_u32(hit), _u32(oh)), f
            ok = _u32(oh[:, 3]) != 0xFFFFFFFF
            assert np.array_equal(inst[ok], oi[ok])
            # G-buffer motion vectors / object ids / motion moment (same formulas as REF_V2)
            ref = osc.render_refv2(p, W, H, moment=None if f == 0 else moment_ref)
            moment_ref = ref["moment"]
            assert np.array_equal(ctx.read("OBJECTID_0"), ref["objid"])
            mv = ctx.read("MOTIONVECTOR")
            assert np.allclose(mv, ref["motion"], atol=1e-3, rtol=1e-4), np.abs(mv - ref["motion"]).max()
            assert np.array_equal(ctx.read("MOTIONMOMENT"), moment_ref)
            # reprojection and bilateral on the frame's own inputs
            src = {k: h2f(ctx.read(k)) for k in ("SINGLE_DIFFUSE", "SINGLE_SPECULAR", "ALBEDO", "NORMAL")}
            acc = [h2f(ctx.read(k)) for k in ("ACCUM_DIFFUSE", "ACCUM_SPECULAR", "ACCUM_ALBEDO")]
            if hist is not None:
                rd, rs, ra = po.reproject(src["SINGLE_DIFFUSE"], src["SINGLE_SPECULAR"], src["ALBEDO"],
                                          src["NORMAL"], *hist, ctx.read("OBJECTID_0"), ctx.read("OBJECTID_1"),
                                          mv, ctx.read("MOTIONMOMENT"), prev_acc, temporal_frames=16)
                for g, r in zip(acc, (rd, rs, ra)):
                    r16 = r.astype(np.float16).astype(np.float32)
                    close = np.isclose(g, r16, rtol=2e-3, atol=1e-4)
                    assert close.mean() > 0.999, close.mean()
                den = h2f(ctx.read("DENOISED"))
                od = po.jbf(*acc).astype(np.float16).astype(np.float32)
                assert np.isclose(den, od, rtol=4e-3, atol=2e-3).mean() > 0.999
            hist = acc
            prev_acc = acc
            prev = p
