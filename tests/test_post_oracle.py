"""Pins the oracle's post passes (oracle/oracle_post.cpp) by what the mathematics fixes.

Reprojection (Process.ReProject.comp.slang:71-223):
  * a static scene with zero motion is an exponential moving average with weight
    1/TemporalFrames on the current frame, towards history taken at the same pixel;
  * motion samples history bilinearly at pixel + motion, which is exact on a linear ramp;
  * history is dropped on a miss, a moving object (moment > 0), or a source position
    outside the image (previpos = -1 included: the reference's int2 < uint2 test compares
    as unsigned);
  * a bilinear tap of another object falls back to the current sample;
  * the 5x5 spatial prefilter is a normalised average: a uniform image passes unchanged,
    and neighbours of another object or with a normal outside the 0.98 cone get weight 0;
  * progressive mode is lerp(prev, current, 1/512) on all four channels.
Bilateral filter (Process.DenoiseJBF.comp.slang:75-181):
  * with both sigmas large every weight is 1, so the filter is the mean of its 36 taps
    (offsets +-1, +-3, +-5 on both axes, the centre excluded) of colour + 0.001, out-of-image
    taps reading 0 (+0.001);
  * a uniform image maps to (colour + 0.001) * albedo + spec + 0.001 for any sigmas;
  * bf_size 0 composes colour * albedo + spec.
GT tonemap (Const_Func.slang:84-121, Uchimura 2017 with the reference's contrast 0.7):
  known values on the toe, the linear section and the shoulder, its limit, monotonicity.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import pyoracle as po


def f16(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


def rgba(h, w, rgb, alpha=1.0):
    out = np.empty((h, w, 4), np.float32)
    out[..., :3] = rgb
    out[..., 3] = alpha
    return out


class Frame:
    """Inputs of one reprojection: a static plane (object 7, normal +z) by default."""

    def __init__(self, W=24, H=20, seed=0):
        rng = np.random.default_rng(seed)
        self.W, self.H = W, H
        yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
        self.xx, self.yy = xx, yy
        self.cd = rgba(H, W, (0.8, 0.3, 1.7), 0.5)
        self.cs = rgba(H, W, (0.2, 0.9, 0.05), 0.5)
        self.ca = f16(np.stack([0.1 + 0.02 * xx, 0.3 + 0.01 * yy, 0.2 + 0.01 * (xx + yy), np.ones_like(xx)], -1))
        self.normal = rgba(H, W, (0.0, 0.0, 1.0), 0.5)
        # history: linear ramps (exactly representable in fp16 at these sizes)
        self.pd = f16(np.stack([0.5 + 0.03125 * xx + 0.0625 * yy, 1.0 + 0.0625 * xx, 2.0 - 0.03125 * yy,
                                np.ones_like(xx)], -1))
        self.ps = f16(np.stack([0.25 + 0.0625 * yy, 0.125 + 0.03125 * xx, 0.5 + 0.0 * xx, np.ones_like(xx)], -1))
        self.pa = f16(np.stack([0.75 - 0.015625 * xx, 0.5 + 0.015625 * yy, 0.25 + 0.0 * xx, np.ones_like(xx)], -1))
        self.obj0 = np.full((H, W), 7, np.uint32)
        self.obj1 = np.full((H, W), 7, np.uint32)
        self.motion = np.zeros((H, W, 2), np.float32)
        self.moment = np.zeros((H, W), np.uint16)
        self.acc_init = [rgba(H, W, (0.0, 0.0, 0.0), 0.25) for _ in range(3)]
        self.rng = rng

    def run(self, T=16, progressive=False):
        return po.reproject(self.cd, self.cs, self.ca, self.normal, self.pd, self.ps, self.pa, self.obj0,
                            self.obj1, self.motion, self.moment, self.acc_init, temporal_frames=T,
                            progressive=progressive)


def lerp(a, b, t):
    return a * (1.0 - t) + b * t


# ---------------------------------------------------------------- reprojection
@pytest.mark.parametrize("T", [1, 4, 16, 256])
def test_reproject_zero_motion_static_scene_is_an_ema_of_the_same_pixel(T):
    f = Frame()
    d, s, a = f.run(T)
    k = 1.0 / T
    for out, hist, cur in ((d, f.pd, f.cd), (s, f.ps, f.cs), (a, f.pa, f.ca)):
        want = lerp(hist[..., :3].astype(np.float64), cur[..., :3].astype(np.float64), k)
        np.testing.assert_allclose(out[..., :3], want, rtol=2e-6, atol=1e-7)
        # .rgb writes keep the target's alpha
        assert np.all(out[..., 3] == 0.25)


def test_reproject_repeated_frames_converge_as_the_closed_form_ema():
    f = Frame()
    T, N = 16, 40
    h0 = f.pd.copy()
    for _ in range(N):
        d, s, a = f.run(T)
        f.pd, f.ps, f.pa = d.copy(), s.copy(), a.copy()
        for x in (f.pd, f.ps, f.pa):
            x[..., 3] = 1.0
    c = f.cd[..., :3].astype(np.float64)
    want = c + (h0[..., :3] - c) * (1.0 - 1.0 / T) ** N
    np.testing.assert_allclose(f.pd[..., :3], want, rtol=2e-5, atol=1e-6)


def test_reproject_motion_samples_history_bilinearly_at_pixel_plus_motion():
    f = Frame()
    W, H = f.W, f.H
    mx, my = 2.25, -1.5
    f.motion[..., 0], f.motion[..., 1] = mx, my
    T = 8
    k = 1.0 / T
    d, s, a = f.run(T)
    xx, yy = f.xx, f.yy
    ppx, ppy = (xx + mx).astype(int), np.floor(yy + my).astype(int)
    # all four taps inside: the bilinear history of a linear ramp is the ramp at x + motion
    full = (ppx + 1 <= W - 1) & (ppy >= 0) & (ppy + 1 <= H - 1)

    def ramp(img, x, y):  # the history ramps evaluated off-grid from their coefficients
        c0 = img[0, 0, :3].astype(np.float64)
        cx = (img[0, 1, :3] - img[0, 0, :3]).astype(np.float64)
        cy = (img[1, 0, :3] - img[0, 0, :3]).astype(np.float64)
        return c0 + cx * x[..., None] + cy * y[..., None]

    for out, hist, cur in ((d, f.pd, f.cd), (s, f.ps, f.cs), (a, f.pa, f.ca)):
        want = lerp(ramp(hist, xx + mx, yy + my), cur[..., :3].astype(np.float64), k)
        np.testing.assert_allclose(out[full][:, :3], want[full], rtol=1e-5, atol=1e-6)
    # a source position outside the image (ppx >= W, or ppy = -1 and below) keeps the
    # current frame, all four channels
    outside = (ppx >= W) | (ppy < 0)
    assert outside[1].all() and not outside[2, : W - 2].any()  # row y = 1 has ppy = -1
    for out, cur in ((d, f.cd), (s, f.cs), (a, f.ca)):
        np.testing.assert_array_equal(out[outside], cur[outside])
    # ppx = W - 1: the right-hand taps read outside the image, where the object id is 0,
    # so they take the current (spatially filtered) sample
    edge = (ppx == W - 1) & (ppy >= 0) & (ppy + 1 <= H - 1)
    assert edge.any()
    sx, sy = mx - math.floor(mx), my - math.floor(my)
    cur = f.cd[..., :3].astype(np.float64)
    h = f.pd[..., :3].astype(np.float64)
    ys, xs = np.nonzero(edge)
    for y, x in zip(ys, xs):
        px, py = ppx[y, x], ppy[y, x]
        top = lerp(h[py, px], cur[y, x], sx)
        bot = lerp(h[py + 1, px], cur[y, x], sx)
        np.testing.assert_allclose(d[y, x, :3], lerp(lerp(top, bot, sy), cur[y, x], k), rtol=1e-5)


def test_reproject_drops_history_on_miss_motion_moment_and_other_objects():
    f = Frame()
    f.obj0[3, 4] = 65535          # miss
    f.moment[5, 6] = 3            # the object moved recently
    f.motion[..., 0] = 1.0        # integer motion: history is the tap at (x + 1, y)
    f.obj1[8, 10] = 9             # the source of pixel (9, 8) was another object
    T = 4
    d, s, a = f.run(T)
    for y, x in ((3, 4), (5, 6)):
        for out, cur in ((d, f.cd), (s, f.cs), (a, f.ca)):
            np.testing.assert_array_equal(out[y, x], cur[y, x])
    # another object at the source: history is the current sample, so the EMA returns it
    np.testing.assert_allclose(d[8, 9, :3], f.cd[8, 9, :3], rtol=1e-6)
    np.testing.assert_allclose(a[8, 9, :3], f.ca[8, 9, :3], rtol=1e-6)
    # its neighbour reads the tap at (11, 8), which is the same object
    np.testing.assert_allclose(d[8, 10, :3], lerp(f.pd[8, 11, :3].astype(np.float64), f.cd[8, 10, :3], 1 / T),
                               rtol=2e-6)


def test_reproject_spatial_prefilter_is_normalised_and_rejects_other_surfaces():
    """The 5x5 prefilter of the current frame stands in for history taps of another object.
    With integer motion and another object id everywhere in the previous frame, every tap
    is rejected and the accumulation is lerp(spatial, current, 1/T)."""
    T = 2

    def frame():
        f = Frame()
        f.motion[..., 0] = 1.0      # integer motion (so the previous ids are read)
        f.obj1[:] = 9               # ... and none of them matches
        return f

    base = frame()
    d, _, _ = base.run(T)
    # uniform current frame: the prefilter passes it unchanged
    inner = np.s_[:, : base.W - 1]
    np.testing.assert_allclose(d[inner][..., :3], base.cd[inner][..., :3], rtol=2e-6)
    q = (10, 12)

    def impulse(obj=7, normal=(0.0, 0.0, 1.0)):
        f = frame()
        f.cd[q[0], q[1], :3] = 100.0
        f.obj0[q] = obj
        f.normal[q[0], q[1], :3] = normal
        d, _, _ = f.run(T)
        near = np.zeros((f.H, f.W), bool)
        near[q[0] - 2:q[0] + 3, q[1] - 2:q[1] + 3] = True
        near[q] = False
        return np.abs(d[..., 0] - f.cd[..., 0])[near]

    spread = impulse()
    assert spread.min() > 1.0                          # same surface: every neighbour sees it
    # closed form at a distance-1 neighbour: weights 0.4 (centre) and 2 / (1.5 d + 4) for the
    # 24 others, all on the same surface away from the border
    w = [0.4 if (x, y) == (0, 0) else 2.0 / (1.5 * math.hypot(x, y) + 4.0) for y in range(-2, 3)
         for x in range(-2, 3)]
    c = base.cd[0, 0, 0]
    f = frame()
    f.cd[q[0], q[1], :3] = 100.0
    d, _, _ = f.run(T)
    spatial = c + (100.0 - c) * (2.0 / 5.5) / sum(w)
    assert d[q[0], q[1] + 1, 0] == pytest.approx(0.5 * spatial + 0.5 * c, rel=1e-5)
    assert impulse(obj=8).max() < 1e-5                 # another object: weight 0
    # normal inside the 0.98 cone: weight scaled by (0.988 - 0.98) / 0.02 = 0.4, not 0
    assert impulse(normal=(0.0, 0.15, 0.988)).min() > 0.2
    assert impulse(normal=(0.0, 0.21, 0.9777)).max() < 1e-5  # outside it: weight 0


def test_reproject_progressive_is_lerp_by_1_over_512_on_all_channels():
    f = Frame(seed=3)
    rng = f.rng
    for x in (f.cd, f.cs, f.ca, f.pd, f.ps, f.pa):
        x[:] = f16(rng.uniform(0.0, 4.0, x.shape))
    f.motion[:] = rng.normal(0, 3, f.motion.shape)
    f.obj0[:] = rng.integers(0, 3, f.obj0.shape)
    f.moment[:] = rng.integers(0, 2, f.moment.shape)
    d, s, a = f.run(16, progressive=True)
    for out, prev, cur in ((d, f.pd, f.cd), (s, f.ps, f.cs), (a, f.pa, f.ca)):
        np.testing.assert_allclose(out, lerp(prev.astype(np.float64), cur, 1 / 512), rtol=1e-6, atol=1e-7)


# ---------------------------------------------------------------- bilateral filter
TAPS = [-5, -3, -1, 1, 3, 5]


def box36(img):
    """Mean over the 36 taps of (img + 0.001), reading 0 outside the image."""
    H, W = img.shape[:2]
    pad = np.zeros((H + 10, W + 10, 3), np.float64)
    pad[5:H + 5, 5:W + 5] = img[..., :3]
    acc = np.zeros((H, W, 3))
    for dy in TAPS:
        for dx in TAPS:
            acc += pad[5 + dy:5 + dy + H, 5 + dx:5 + dx + W] + 0.001
    return acc / 36.0


def test_jbf_with_large_sigmas_is_the_36_tap_box_mean():
    rng = np.random.default_rng(1)
    H, W = 18, 22
    d = f16(rng.uniform(0, 2, (H, W, 4)))
    s = f16(rng.uniform(0, 1, (H, W, 4)))
    a = f16(rng.uniform(0.1, 0.9, (H, W, 4)))
    out = po.jbf(d, s, a, sigma=1e6, sigma_lum=1e6, linear=True)
    want = box36(d) * a[..., :3] + s[..., :3] + 0.001
    np.testing.assert_allclose(out[..., :3], want, rtol=2e-6, atol=1e-7)


def test_jbf_tap_pattern_odd_offsets_without_the_centre():
    H = W = 21
    d = np.zeros((H, W, 4), np.float32)
    d[10, 10, :3] = 36.0
    a = np.ones((H, W, 4), np.float32)
    s = np.zeros((H, W, 4), np.float32)
    out = po.jbf(d, s, a, sigma=1e6, sigma_lum=1e6, linear=True)[..., 0] - 0.001
    dy, dx = np.mgrid[0:H, 0:W] - 10
    tap = np.isin(dy, TAPS) & np.isin(dx, TAPS)
    np.testing.assert_allclose(out[tap], 1.0 + 0.001, rtol=1e-5)
    np.testing.assert_allclose(out[~tap], 0.001, atol=1e-6)
    assert not tap[10, 10]


@pytest.mark.parametrize("sigma,sigma_lum", [(2.0, 3.0), (0.5, 0.01), (10.0, 100.0)])
def test_jbf_uniform_image_composes_colour_albedo_and_spec(sigma, sigma_lum):
    H, W = 20, 20
    c, sp, al = (0.7, 1.3, 0.2), (0.05, 0.1, 0.4), (0.5, 0.25, 0.8)
    out = po.jbf(rgba(H, W, c), rgba(H, W, sp), rgba(H, W, al), sigma=sigma, sigma_lum=sigma_lum, linear=True)
    want = (np.array(c) + 0.001) * np.array(al) + np.array(sp) + 0.001
    np.testing.assert_allclose(out[5:-5, 5:-5, :3], np.broadcast_to(want, (10, 10, 3)), rtol=2e-6)


def test_jbf_disabled_composes_without_offsets():
    rng = np.random.default_rng(2)
    d, s, a = (f16(rng.uniform(0, 2, (9, 11, 4))) for _ in range(3))
    out = po.jbf(d, s, a, bf_size=0, linear=True)
    np.testing.assert_allclose(out[..., :3], d[..., :3].astype(np.float64) * a[..., :3] + s[..., :3], rtol=1e-6)


# ---------------------------------------------------------------- GT tonemap
# Uchimura's curve with P = 1, a = 0.7, m = 0.22, l = 0.4, c = 1.33, b = 0: the linear section
# runs from m to S0 = m + (P - m) l / a = 0.665714.., where it reaches S1 = m + a l0 = 0.532, and
# the shoulder is P - (P - S1) exp(-C2 (x - S0) / P) with C2 = a P / (P - S1) = 0.7 / 0.468.
S0 = 0.22 + 0.78 * 0.4 / 0.7
C2 = 0.7 / 0.468


@pytest.mark.parametrize("x,want", [
    (0.0, 0.0),
    (0.22, 0.22),                       # toe meets the linear section
    (0.3, 0.22 + 0.7 * 0.08),
    (0.5, 0.22 + 0.7 * 0.28),
    (0.6, 0.22 + 0.7 * 0.38),
    (1.0, 1.0 - 0.468 * math.exp(-0.5)),  # C2 (1 - S0) = 1/2 exactly
    (2.0, 1.0 - 0.468 * math.exp(-C2 * (2.0 - S0))),
    (100.0, 1.0),
    # toe at m/2: smoothstep weight 1/2 between m (x/m)^c and the line
    (0.11, 0.5 * 0.22 * 0.5 ** 1.33 + 0.5 * (0.22 + 0.7 * (0.11 - 0.22))),
])
def test_gt_tonemap_known_values(x, want):
    assert po.oracle().oracle_gt_tonemap(x) == pytest.approx(want, rel=2e-6, abs=1e-7)


def test_gt_tonemap_is_monotone_and_continuous():
    xs = np.linspace(0.0, 8.0, 80001)
    ys = np.array([po.oracle().oracle_gt_tonemap(float(x)) for x in xs])
    dy = np.diff(ys)
    assert (dy >= -1e-7).all()
    assert np.abs(dy).max() < 2e-4   # no jump at m or S0 (the slope stays below 2)
    assert ys[-1] < 1.0
