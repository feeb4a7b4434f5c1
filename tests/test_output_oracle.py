"""Pins the oracle's output stage: ST 2084 encoding, the SDR/HDR compose transform, the
linear-blit and FSR EASU resolves, and the SuperResolution render extents. CPU only."""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po


def test_st2084_known_values():
    o = po.oracle()
    assert o.oracle_st2084(10000.0) == 1.0                      # peak luminance -> code 1
    np.testing.assert_allclose(o.oracle_st2084(100.0), 0.5081, atol=2e-4)   # SDR white, PQ 100 nits
    np.testing.assert_allclose(o.oracle_st2084(1000.0), 0.7518, atol=2e-4)
    np.testing.assert_allclose(o.oracle_st2084(0.0), 0.8359375 ** 78.84375, rtol=1e-4)
    v = [o.oracle_st2084(x) for x in np.geomspace(0.01, 10000, 40)]
    assert np.all(np.diff(v) > 0)


def test_compose_transform_sdr_and_hdr():
    o = po.oracle()
    for v in (0.0, 3.0, 150.0, 4000.0):
        assert o.oracle_display(v, 600.0, 0) == o.oracle_gt_tonemap(np.float32(v) * np.float32(600.0 / 40000.0))
        t = np.float32(o.oracle_gt_tonemap(np.float32(v) / np.float32(2000.0))) * np.float32(2000.0)
        assert o.oracle_display(v, 600.0, 1) == o.oracle_st2084(t * np.float32(600.0) / np.float32(230.0))
    # HDR saturates where the tonemap does: 2000 * 600 / 230 nits
    np.testing.assert_allclose(o.oracle_display(1e7, 600.0, 1), o.oracle_st2084(2000.0 * 600.0 / 230.0), rtol=1e-6)


def test_render_extent_follows_superresolution():
    assert gk.render_extent(1920, 1080, 0) == (960, 540)
    assert gk.render_extent(1920, 1080, 1) == (1280, 720)
    assert gk.render_extent(1920, 1080, 2) == (1920, 1080)


@pytest.mark.parametrize("mode", [0, 1])
def test_upscale_constant_and_bounds(mode):
    c = np.zeros((12, 20, 4), np.float32)
    c[:] = [0.3, 0.6, 0.9, 1.0]
    assert np.array_equal(po.upscale(c, 30, 18, mode), np.broadcast_to(c[0, 0], (18, 30, 4)))
    rng = np.random.default_rng(mode)
    img = rng.uniform(0, 1, (12, 20, 4)).astype(np.float32)
    out = po.upscale(img, 30, 18, mode)
    assert out.min() >= img.min() and out.max() <= img.max()     # no overshoot (dering / convex weights)
    assert np.array_equal(po.upscale(img, 20, 12, 1), img)         # equal sizes: copy


def test_linear_blit_reproduces_a_ramp():
    w = 16
    img = np.zeros((4, w, 4), np.float32)
    img[..., 0] = np.arange(w)[None, :]
    out = po.upscale(img, 2 * w, 8, 0)
    u = (np.arange(2 * w) + 0.5) * 0.5 - 0.5
    want = np.clip(u, 0, w - 1)
    np.testing.assert_allclose(out[3, :, 0], want, atol=1e-6)


def test_easu_keeps_a_step_edge_sharp():
    img = np.zeros((16, 16, 4), np.float32)
    img[:, 8:, :3] = 1.0
    out = po.upscale(img, 32, 32, 1)[..., 0]
    row = out[16]
    assert np.all(row[:12] == 0) and np.all(row[-12:] == 1)        # far from the edge: the levels exactly
    assert np.all(np.diff(row) >= 0)                               # monotone across the edge
    bil = po.upscale(img, 32, 32, 0)[16, :, 0]
    # EASU's transition is no wider than bilinear's
    assert ((row > 0.05) & (row < 0.95)).sum() <= ((bil > 0.05) & (bil < 0.95)).sum()


def test_easu_flat_regions_stay_finite_and_exact():
    """In a flat neighbourhood the edge length is 0 / 0; clamped as the GPU clamps (NaN ->
    0) the kernel stays finite and a flat image is reproduced exactly."""
    img = np.zeros((10, 10, 4), np.float32)
    img[:, :, :3] = 0.25
    img[4:6, 4:6, :3] = 0.75           # a bump inside a flat field
    out = po.upscale(img, 15, 15, 1)
    assert np.isfinite(out).all()
    assert np.all(out[:3, :, :3] == 0.25) and np.all(out[-3:, :, :3] == 0.25)
