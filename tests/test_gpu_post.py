"""Dynamic instanced bricks: TLAS refit, motion vectors, reprojection and the
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


# ---------------------------------------------------------------- post passes on given targets
def post_only_params(W, H, progressive=False, denoiser=gk.DENOISER_BILATERAL, linear=True):
    c = scenes.CORNELL_CAM
    p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"])
    p.temporal_frames = 8
    p.progressive = int(progressive)
    p.denoiser = denoiser
    p.flags |= gk.FRAME_POST_ONLY | (gk.FRAME_LINEAR_OUTPUT if linear else 0)
    return p


def write_inputs(ctx, d, s, a, normal, obj0, obj1, motion, moment):
    for name, v in (("SINGLE_DIFFUSE", d), ("SINGLE_SPECULAR", s), ("ALBEDO", a), ("NORMAL", normal),
                    ("OBJECTID_0", obj0), ("OBJECTID_1", obj1), ("MOTIONVECTOR", motion),
                    ("MOTIONMOMENT", moment)):
        ctx.write(name, v)


def test_post_passes_on_given_targets_match_oracle():
    """GK_FRAME_POST_ONLY runs reprojection, bilateral filter and history copy on targets the
    caller wrote. Frame 1 (every pixel moving) seeds the history with its own values; frame
    2 reprojects with fractional motion (2.25, -1.5), object ids that change in places and
    moving pixels; frame 3 is progressive. Each matches the oracle to fp16 rounding, rows
    whose source row is -1 keep the current frame (the reference's unsigned bounds test),
    and frame 3 is lerp(history, current, 1/512)."""
    W, H = 40, 36
    rng = np.random.default_rng(8)
    f16 = lambda a: np.asarray(a, np.float32).astype(np.float16)  # noqa: E731

    def rand_targets():
        d = f16(rng.uniform(0.0, 3.0, (H, W, 4)))
        s = f16(rng.uniform(0.0, 1.0, (H, W, 4)))
        a = f16(rng.uniform(0.05, 0.95, (H, W, 4)))
        return d, s, a

    normal = np.zeros((H, W, 4), np.float16)
    normal[..., 2] = 1.0
    normal[5:9, 10:30, 1] = 0.5  # a strip whose normal leaves the 0.98 cone
    normal[5:9, 10:30, 2] = 0.85
    obj0 = np.full((H, W), 3, np.uint32)
    obj0[20:, 25:] = 4
    obj0[0, 0] = 65535
    with gk.Context(W, H) as ctx:
        ctx.upload(gk.procgen_cornell())
        # frame 1: all pixels moving, so the accumulation is the current frame
        d1, s1, a1 = rand_targets()
        write_inputs(ctx, d1, s1, a1, normal, obj0, obj0, np.zeros((H, W, 2), np.float32),
                     np.ones((H, W), np.uint16))
        ctx.frame(post_only_params(W, H))
        ctx.render()
        ctx.sync()
        acc1 = [ctx.read(k) for k in ("ACCUM_DIFFUSE", "ACCUM_SPECULAR", "ACCUM_ALBEDO")]
        for g, cur in zip(acc1, (d1, s1, a1)):
            assert np.array_equal(g, cur)
        # frame 2: reprojection with motion
        d2, s2, a2 = rand_targets()
        motion = np.empty((H, W, 2), np.float32)
        motion[..., 0], motion[..., 1] = 2.25, -1.5
        obj1 = obj0.copy()
        obj1[10:14, 6:12] = 9
        moment = np.zeros((H, W), np.uint16)
        moment[30:, :5] = 2
        write_inputs(ctx, d2, s2, a2, normal, obj0, obj1, motion, moment)
        ctx.frame(post_only_params(W, H))
        ctx.render()
        ctx.sync()
        acc2 = [ctx.read(k).astype(np.float32) for k in ("ACCUM_DIFFUSE", "ACCUM_SPECULAR", "ACCUM_ALBEDO")]
        den2 = ctx.read("DENOISED").astype(np.float32)
        ref = po.reproject(d2, s2, a2, normal, *(x.astype(np.float32) for x in acc1), obj0, obj1, motion, moment,
                           [x.astype(np.float32) for x in acc1], temporal_frames=8)
        for g, r in zip(acc2, ref):
            r16 = r.astype(np.float16).astype(np.float32)
            assert np.isclose(g, r16, rtol=2e-3, atol=1e-4).mean() > 0.999
        # rows 0 and 1 have source rows -2 and -1: no history, the current frame
        for g, cur in zip(acc2, (d2, s2, a2)):
            assert np.array_equal(g[:2], cur[:2].astype(np.float32))
        jbf = po.jbf(*acc2, sigma=2.0, sigma_lum=3.0, linear=True)
        assert np.isclose(den2[..., :3], jbf[..., :3].astype(np.float16).astype(np.float32), rtol=4e-3,
                          atol=2e-3).mean() > 0.999
        # frame 3: progressive
        d3, s3, a3 = rand_targets()
        write_inputs(ctx, d3, s3, a3, normal, obj0, obj1, motion, moment)
        ctx.frame(post_only_params(W, H, progressive=True))
        ctx.render()
        ctx.sync()
        acc3 = [ctx.read(k).astype(np.float32) for k in ("ACCUM_DIFFUSE", "ACCUM_SPECULAR", "ACCUM_ALBEDO")]
        for g, prev, cur in zip(acc3, acc2, (d3, s3, a3)):
            want = (prev.astype(np.float64) * (1 - 1 / 512) + cur.astype(np.float64) / 512).astype(np.float16)
            assert np.isclose(g, want.astype(np.float32), rtol=1e-3, atol=1e-4).all()


def test_taa_jittered_primary_hits_bit_exact():
    """TAA frames: the Halton-jittered projection of gk_make_frame_params drives the traced
    primary rays; their hits equal the oracle's bit for bit."""
    W, H = 96, 64
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    c = scenes.CORNELL_CAM
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        prev = None
        for f in range(4):
            p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"], frame=f, taa=True, prev=prev)
            p.estimator = gk.ESTIMATOR_PT_MIS
            p.bounces = p.max_bounces = 2
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            hit = ctx.read("PRIMARY_HIT").reshape(-1, 4)
            oh, _ = osc.trace(po.primary_rays(p, W, H))
            assert np.array_equal(_u32(hit), _u32(oh)), f
            prev = p
