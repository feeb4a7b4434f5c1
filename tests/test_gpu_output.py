"""Parity of the output stage with the oracle: HDR compose, the linear-blit and FSR EASU
resolves, and screenshots."""
from __future__ import annotations

import struct
import zlib

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests.test_gpu_render import cornell_params

pytestmark = pytest.mark.gpu


def h2f(a):
    return a.astype(np.float32)


def frame(W, H, hdr=False, denoiser=gk.DENOISER_BILATERAL):
    s = gk.procgen_cornell()
    ctx = gk.Context(W, H)
    ctx.upload(s)
    p = cornell_params(W, H, estimator=gk.ESTIMATOR_PT_MIS)
    p.denoiser = denoiser
    p.hdr = 1 if hdr else 0
    p.paper_white_nit = 600.0
    p.bf_size = 5
    ctx.frame(p)
    ctx.render()
    ctx.sync()
    return ctx


@pytest.mark.parametrize("hdr", [False, True])
def test_compose_matches_oracle(hdr):
    W, H = 96, 64
    ctx = frame(W, H, hdr)
    acc = [h2f(ctx.read(k)) for k in ("ACCUM_DIFFUSE", "ACCUM_SPECULAR", "ACCUM_ALBEDO")]
    den = h2f(ctx.read("DENOISED"))
    ctx.close()
    od = po.jbf(*acc, paper_white=600.0, hdr=hdr).astype(np.float16).astype(np.float32)
    fin = np.isfinite(od) & np.isfinite(den)
    assert np.isclose(den[fin], od[fin], rtol=4e-3, atol=2e-3).mean() > 0.999
    if hdr:
        assert den[..., :3][fin[..., :3]].max() <= 1.0
        sdr = frame(W, H, False)
        assert not np.allclose(h2f(sdr.read("DENOISED")), den)
        sdr.close()


@pytest.mark.parametrize("mode", [gk.UPSCALE_LINEAR, gk.UPSCALE_FSR_EASU])
def test_upscale_matches_oracle(mode):
    out_w, out_h = 144, 96
    W, H = gk.render_extent(out_w, out_h, 1)
    ctx = frame(W, H)
    src = h2f(ctx.read("DENOISED"))
    got = ctx.upscale(out_w, out_h, mode)
    same = ctx.upscale(W, H, mode)
    ctx.close()
    want = po.upscale(src, out_w, out_h, int(mode)).astype(np.float16)
    fin = np.isfinite(want.astype(np.float32))
    eq = got.view(np.uint16) == want.view(np.uint16)
    assert eq[fin].mean() > 0.999, eq[fin].mean()
    np.testing.assert_allclose(got.astype(np.float32)[fin], want.astype(np.float32)[fin], atol=2e-3)
    if mode == gk.UPSCALE_FSR_EASU:
        assert np.array_equal(same.view(np.uint16), src.astype(np.float16).view(np.uint16))


def read_png_rgb(path):
    d = open(path, "rb").read()
    assert d[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, w = 8, b"", 0
    while p < len(d):
        n = struct.unpack(">I", d[p:p + 4])[0]
        t = d[p + 4:p + 8]
        body = d[p + 8:p + 8 + n]
        assert struct.unpack(">I", d[p + 8 + n:p + 12 + n])[0] == zlib.crc32(t + body) & 0xFFFFFFFF
        if t == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2)
        elif t == b"IDAT":
            idat += body
        p += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 3 * w + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 3)


def test_screenshot_writes_the_display_image(tmp_path):
    W, H = 64, 48
    ctx = frame(W, H)
    den = h2f(ctx.read("DENOISED"))
    ctx.screenshot(tmp_path / "shot.png")
    ctx.close()
    img = read_png_rgb(tmp_path / "shot.png")
    d = np.nan_to_num(den[..., :3], nan=0.0)
    want = np.round(np.clip(d, 0, 1) * 255).astype(np.uint8)
    assert np.array_equal(img, want)
