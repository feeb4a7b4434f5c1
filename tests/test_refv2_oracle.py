"""Pins the oracle's camera rays and REF_V2 estimator (oracle/oracle_render.cpp) by what
the geometry and the estimator's structure fix, independently of the formulas they are
written with.

Camera (FHardwarePrimaryRayCaster, Shading.slang:283-334; projection from
GetUniformBufferObject, Engine.cpp:746-754): uv = ipos / size * 2 - 1 with no half-pixel
offset, a vertical field of view and the window's aspect ratio, Vulkan's flipped Y (row 0
looks up). Every pixel's direction is compared with the pinhole closed form
normalize(f + uv.x * aspect * tan(fov/2) * s - uv.y * tan(fov/2) * u) from the look-at
basis (f forward, s right, u up).

REF_V2 (FPathTracingRendererV2, Shading.slang:1003-1184):
  * zero emission (no emitter, sky or sun, no probes) renders zero;
  * an open plane under a uniform sky returns the sky on every path: the first bounce
    leaves the plane, so diffuse + specular = sky exactly, whichever lobe is drawn (a
    bounce below the surface would hit the plane's back and return 0.5 grey);
  * the Cornell box without its two blocks is symmetric about x = 1/2; swapping the red
    and green walls mirrors the image. With uv = x/W*2-1 the mirror of pixel x is W - x;
    the per-pixel means of the two renders agree there within their sampling error, and
    do not without the mirror.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import gknextengine_amd as gk
from gknextengine_amd.scenes import material
from oracle import pyoracle as po
from tests import scenes


def lookat_basis(eye, target, up):
    f = np.subtract(target, eye).astype(np.float64)
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    return f, s, u


@pytest.mark.parametrize("W,H,fov,eye,target,up", [
    (64, 48, 40.0, (0.5, 0.5, -1.4), (0.5, 0.5, 0.0), (0, 1, 0)),
    (96, 54, 50.0, (-20.0, 28.0, -44.0), (2.0, 0.0, 4.0), (0, 1, 0)),
    (33, 65, 75.0, (1.0, 2.0, 3.0), (-2.0, 0.5, 1.0), (0, 0, 1)),
])
def test_camera_rays_are_the_pinhole_closed_form(W, H, fov, eye, target, up):
    p = gk.make_frame_params(eye, target, W, H, fov=fov, up=up)
    rays = po.primary_rays(p, W, H).reshape(H, W, 8)
    f, s, u = lookat_basis(eye, target, up)
    t = math.tan(math.radians(fov) / 2)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    ux, uy = xx / W * 2 - 1, yy / H * 2 - 1
    d = f + (ux * (W / H) * t)[..., None] * s - (uy * t)[..., None] * u
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    np.testing.assert_allclose(rays[..., 4:7], d, atol=2e-6)
    np.testing.assert_allclose(rays[..., :3], np.broadcast_to(eye, (H, W, 3)), atol=1e-5)
    # the named pixels: centre = view direction, row 0 looks up by fov/2, column 0 left
    if W % 2 == 0 and H % 2 == 0:
        np.testing.assert_allclose(rays[H // 2, W // 2, 4:7], f, atol=1e-6)
        assert math.degrees(math.acos(np.dot(rays[0, W // 2, 4:7], f))) == pytest.approx(fov / 2, abs=1e-3)
        assert np.dot(rays[0, W // 2, 4:7], u) > 0
        assert np.dot(rays[H // 2, 0, 4:7], s) < 0


def test_camera_jitter_shifts_by_the_halton_offset():
    """TAA: the projection carries Halton(2,3) - 1/2 pixel offsets (Engine.cpp:771-778)."""
    W, H = 64, 64
    cam = scenes.CORNELL_CAM
    p0 = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    f, s, u = lookat_basis(cam["eye"], cam["target"], (0, 1, 0))
    t = math.tan(math.radians(cam["fov"]) / 2)

    def halton(i, b):
        r, fct = 0.0, 1.0
        while i > 0:
            fct /= b
            r += fct * (i % b)
            i //= b
        return r

    for frame in (0, 1, 5):
        p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"], frame=frame, taa=True)
        r = po.primary_rays(p, W, H).reshape(H, W, 8)[..., 4:7].astype(np.float64)
        r0 = po.primary_rays(p0, W, H).reshape(H, W, 8)[..., 4:7].astype(np.float64)
        assert not np.allclose(r, r0)
        # recover the image-plane offset of the centre pixel in pixels
        c = r[H // 2, W // 2] / np.dot(r[H // 2, W // 2], f)
        jx = np.dot(c, s) / t * (W / 2)
        jy = -np.dot(c, u) / t * (H / 2)
        idx = frame % p.temporal_frames if p.temporal_frames else frame
        hx, hy = halton(idx + 1, 2) - 0.5, halton(idx + 1, 3) - 0.5
        assert abs(abs(jx) - abs(hx)) < 1e-3 and abs(abs(jy) - abs(hy)) < 1e-3, (frame, jx, jy, hx, hy)


# ---------------------------------------------------------------- REF_V2 invariants
def refv2_params(W, H, eye, target, up=(0, 1, 0), fov=40.0, bounces=4, samples=1, frame=0, sky=None):
    p = gk.make_frame_params(eye, target, W, H, fov=fov, up=up)
    p.estimator = gk.ESTIMATOR_REF_V2
    p.bounces = p.max_bounces = bounces
    p.samples = samples
    p.total_frames = frame
    p.global_seed = 3
    p.has_sun = 0
    p.has_sky = 0 if sky is None else 1
    if sky is not None:
        p.sky_color[:] = [*sky, 1.0]
    return p


def test_refv2_without_emission_renders_zero():
    s = gk.procgen_cornell()
    for m in s.materials:
        if m.material_model == gk.MAT_DIFFUSE_LIGHT:
            m.material_model = gk.MAT_LAMBERTIAN
            m.diffuse[:] = [0.9, 0.9, 0.9, 1.0]
    osc = po.OracleScene(s)
    cam = scenes.CORNELL_CAM
    W = H = 32
    for frame in range(4):
        o = osc.render_refv2(refv2_params(W, H, cam["eye"], cam["target"], samples=4, frame=frame), W, H)
        assert np.all(o["diffuse"][..., :3] == 0.0) and np.all(o["spec"][..., :3] == 0.0)
        assert o["albedo"][..., :3].max() > 0.5  # the frame does see the box


@pytest.mark.parametrize("model,rough,metal", [(gk.MAT_LAMBERTIAN, 0.05, 0.0), (gk.MAT_MIXTURE, 0.05, 0.0),
                                               (gk.MAT_MIXTURE, 0.05, 1.0), (gk.MAT_METALLIC, 0.05, 1.0)])
def test_refv2_open_plane_under_sky_returns_the_sky(model, rough, metal):
    sky = (0.3, 0.6, 0.9)
    s = scenes.tri_scene(scenes.floor_tris(), albedo=(0.6, 0.5, 0.4))
    s.materials = [material((0.6, 0.5, 0.4), model, rough, metal)]
    osc = po.OracleScene(s)
    W = H = 48
    # looking straight down: reflections stay within fov/2 of the normal, so a rough lobe
    # of 0.05 around them never dips below the horizon
    p = refv2_params(W, H, (0.0, 5.0, 0.0), (0.0, 0.0, 0.0), up=(0, 0, 1), samples=1, sky=sky)
    for frame in range(3):
        p.total_frames = frame
        o = osc.render_refv2(p, W, H)
        assert np.all(o["objid"] != 65535)
        tot = o["diffuse"][..., :3] + o["spec"][..., :3]
        if metal == 1.0:  # metal paths multiply by the primary albedo (Shading.slang:1159-1162)
            want = np.array(sky, np.float32) * np.array([0.6, 0.5, 0.4], np.float32)
        else:
            want = np.array(sky, np.float32)
        np.testing.assert_allclose(tot, np.broadcast_to(want, tot.shape), rtol=1e-6)


def _open_box(swap):
    """The Cornell box without its two blocks and with the whole ceiling emitting (radiance
    2, so that REF_V2, which gathers light only by hitting emitters, converges quickly);
    walls as procgen_cornell, red and green swapped when swap."""
    s = gk.procgen_cornell()
    m = s.meshes[0]
    tm = m["tri_material"].copy()
    light = int(tm[-1])
    tm[2:4] = light                     # the ceiling
    keep = np.r_[0:10]                  # floor, ceiling, back, left and right walls
    m["indices"] = np.ascontiguousarray(m["indices"][keep])
    m["tri_material"] = np.ascontiguousarray(tm[keep])
    s.materials[light].diffuse[:] = [2.0, 2.0, 2.0, 1.0]
    if swap:
        s.materials[0], s.materials[1] = s.materials[1], s.materials[0]
    return s


def _mean_and_se(osc, p, W, H, frames):
    imgs = []
    for k in range(frames):
        p.total_frames = 1000 * k
        o = osc.render_refv2(p, W, H)
        imgs.append(o["diffuse"][..., :3] * o["albedo"][..., :3] + o["spec"][..., :3])
    a = np.stack(imgs).astype(np.float64)
    return a.mean(0), a.std(0, ddof=1) / np.sqrt(frames)


def test_refv2_mirror_symmetry_of_the_open_cornell_box():
    cam = scenes.CORNELL_CAM
    W = H = 32
    s_a, s_b = _open_box(False), _open_box(True)
    # the geometry mirrors about x = 1/2 exactly
    tri = s_a.meshes[0]["positions"][s_a.meshes[0]["indices"]]
    mirrored = tri.copy()
    mirrored[..., 0] = 1.0 - mirrored[..., 0]
    key = lambda t: np.sort(np.round(np.sort(t, axis=1), 6).reshape(len(t), -1), axis=0)  # noqa: E731
    np.testing.assert_array_equal(key(tri), key(mirrored))
    p = refv2_params(W, H, cam["eye"], cam["target"], samples=16, bounces=4)
    ma, sa = _mean_and_se(po.OracleScene(s_a), p, W, H, 16)
    mb, sb = _mean_and_se(po.OracleScene(s_b), p, W, H, 16)
    # pixel x mirrors to W - x (uv = x/W*2-1 has no half-pixel offset); column 0 has no mirror
    se = np.sqrt(sa[:, 1:] ** 2 + sb[:, :0:-1] ** 2)
    lit = se > 1e-6
    assert lit.mean() > 0.7
    z = (ma[:, 1:] - mb[:, :0:-1])[lit] / se[lit]
    # without the mirror the red and green walls disagree
    se_plain = np.sqrt(sa[:, 1:] ** 2 + sb[:, 1:] ** 2)[lit]
    z_plain = (ma[:, 1:] - mb[:, 1:])[lit] / se_plain
    chi2, chi2_plain = float((z ** 2).mean()), float((z_plain ** 2).mean())
    print(f"mean z^2: mirrored {chi2:.3f}, not mirrored {chi2_plain:.3f}")
    assert chi2 < 1.6 and chi2_plain > 4.0, (chi2, chi2_plain)


# ---------------------------------------------------------------- thin lens
def test_unit_disk_sampler_is_the_concentric_map():
    """RandomInUnitDisk (Const_Func.slang:260-280): Shirley and Chiu's concentric map of the
    square onto the disk. Known points of the map, and area-uniform samples (r^2 and the
    angle uniform)."""
    d = po.sample(3, 200_000, seed=5)[:, :2].astype(np.float64)
    r2 = (d ** 2).sum(1)
    assert r2.max() <= 1.0 + 1e-6
    q = np.sort(r2)
    assert np.abs(q - (np.arange(q.size) + 0.5) / q.size).max() < 0.005     # r^2 ~ U(0, 1)
    phi = np.sort((np.arctan2(d[:, 1], d[:, 0]) + np.pi) / (2 * np.pi))
    assert np.abs(phi - (np.arange(phi.size) + 0.5) / phi.size).max() < 0.005  # angle ~ U
    # u -> 2u - 1; the larger coordinate is the radius, the angle pi/4 * smaller / larger
    s8, c8 = math.sin(math.pi / 8), math.cos(math.pi / 8)
    for (ua, ub), want in (((0.5, 0.5), (0.0, 0.0)), ((0.75, 0.625), (0.5 * c8, 0.5 * s8)),
                           ((0.625, 0.75), (0.5 * s8, 0.5 * c8)), ((0.25, 0.375), (-0.5 * c8, -0.5 * s8)),
                           ((0.875, 0.5), (0.75, 0.0)), ((0.5, 0.125), (0.0, -0.75)),
                           ((0.0, 0.0), (-math.sqrt(0.5), -math.sqrt(0.5)))):
        np.testing.assert_allclose(po.concentric_disk(ua, ub), want, atol=1e-6)


def _lens_setup(W, H, aperture, focus, frame=0):
    cam = scenes.CORNELL_CAM
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"])
    p.aperture, p.focus_distance = aperture, focus
    p.total_frames, p.global_seed = frame, 7
    return p, lookat_basis(cam["eye"], cam["target"], (0, 1, 0)), np.asarray(cam["eye"], np.float64)


@pytest.mark.parametrize("focus", [1.0, 2.5])
def test_thin_lens_rays_start_on_the_lens_and_meet_on_the_focal_plane(focus):
    assert _focal_depth(focus) > 0


def _focal_depth(focus):
    """FHardwarePrimaryRayCaster's thin lens (Shading.slang:283-300): every ray of a pixel
    starts on the lens disk (radius aperture / 2, centred on the eye, perpendicular to the
    view), and the rays of one pixel drawn from different seeds all pass through one point
    of that pixel's pinhole ray; those points lie on one plane perpendicular to the view,
    whose depth is proportional to the focus distance."""
    W, H, A = 24, 16, 0.2
    pin = po.primary_rays(_lens_setup(W, H, 0.0, focus)[0], W, H).reshape(H, W, 8).astype(np.float64)
    depths = []
    for frame in range(6):
        p, (f, s, u), eye = _lens_setup(W, H, A, focus, frame)
        r = po.lens_rays(p, W, H).reshape(H, W, 8).astype(np.float64)
        off = r[..., :3] - eye
        assert np.abs(off @ f).max() < 1e-6                      # on the lens plane
        rad = np.linalg.norm(off, axis=-1)
        assert rad.max() <= A / 2 + 1e-6 and rad.mean() > A / 8  # inside the lens, and spread
        # where the lens ray crosses the pinhole ray: solve eye + a * dp = o + b * d
        dp, o, d = pin[..., 4:7], r[..., :3], r[..., 4:7]
        n = np.cross(dp, d)
        a = np.einsum("hwk,hwk->hw", np.cross(o - eye, d), n) / np.einsum("hwk,hwk->hw", n, n)
        b = np.einsum("hwk,hwk->hw", np.cross(o - eye, dp), n) / np.einsum("hwk,hwk->hw", n, n)
        x1 = eye + a[..., None] * dp
        x2 = o + b[..., None] * d
        assert np.abs(x1 - x2).max() < 1e-4 * focus             # the two rays meet
        depths.append((x1 - eye) @ f)
    depths = np.stack(depths)
    assert np.abs(depths - depths.mean()).max() < 1e-4 * focus  # one plane, for every pixel and seed
    return float(depths.mean())


def test_thin_lens_focal_depth_scales_with_focus_distance():
    d1, d2 = _focal_depth(1.0), _focal_depth(2.5)
    assert d2 / d1 == pytest.approx(2.5, rel=1e-5)


def test_thin_lens_with_zero_aperture_is_the_pinhole():
    p, _, _ = _lens_setup(20, 12, 0.0, 3.0)
    np.testing.assert_array_equal(po.lens_rays(p, 20, 12), po.primary_rays(p, 20, 12))
