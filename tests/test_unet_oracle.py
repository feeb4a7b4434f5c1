"""CPU checks of the U-Net oracle (oracle/unet_oracle.py).

The oracle has no reference counterpart, so it is pinned here by what the mathematics
fixes: each convolution equals the direct per-pixel sum over taps and channels; the
analytic gradient equals central finite differences of the loss (float64) for weights and
biases of every layer; pooling and upsampling are adjoint pairs; bf16 rounding is nearest
even; one Adam step from zero moments moves each parameter by lr * g / (|g| + eps).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import unet_oracle as U


def test_layer_table_for_base_32():
    assert U.channels(32) == [32, 48, 64, 96, 128]
    L = U.layers(32)
    assert [l["cout"] for l in L] == [32, 32, 48, 64, 96, 128, 96, 64, 48, 32, 32, 16]
    assert [l["cin_a"] + l["cin_b"] for l in L] == [16, 32, 32, 48, 64, 96, 224, 160, 112, 80, 32, 32]
    assert all(l["kpad"] % 32 == 0 and l["kpad"] > 9 * (l["cin_a"] + l["cin_b"]) for l in L)
    # real (unpadded) weights and biases of the network
    assert U.real_params(32) == 587_331  # sum of cout * (9 cin + 1)
    assert U.num_params(32) == sum(l["cout"] * l["kpad"] for l in L)


def _direct_conv(x, W, b):
    """y[n,y,x,co] = b[co] + sum_{dy,dx,ci} x[n,y+dy,x+dx,ci] W[3(dy+1)+(dx+1), ci, co] (zero outside)."""
    n, h, w, c = x.shape
    co = W.shape[2]
    y = np.zeros((n, h, w, co))
    for i in range(n):
        for yy in range(h):
            for xx in range(w):
                acc = b.astype(np.float64).copy()
                for tap in range(9):
                    sy, sx = yy + tap // 3 - 1, xx + tap % 3 - 1
                    if 0 <= sy < h and 0 <= sx < w:
                        acc += x[i, sy, sx] @ W[tap]
                y[i, yy, xx] = acc
    return y


def test_conv_equals_direct_sum():
    rng = np.random.default_rng(0)
    p = U.init_params(8, 1)
    p += rng.normal(0, 0.1, p.shape).astype(np.float32)  # non-zero biases too
    net = U.Net(p, 8, dtype=np.float64)
    for l in (0, 6):
        cin = net.W[l].shape[1]
        x = rng.normal(size=(1, 6, 5, cin))
        np.testing.assert_allclose(net.conv(x, l), _direct_conv(x, net.W[l], net.b[l]), rtol=1e-12, atol=1e-12)
    # the parameter layout: column tap*cin + ci of row co, bias at column 9*cin
    L = U.layers(8)[6]
    cin = L["cin_a"] + L["cin_b"]
    M = p[L["offset"]:L["offset"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
    assert net.W[6][4, 3, 5] == M[5, 4 * cin + 3] and net.b[6][7] == M[7, 9 * cin]


def test_pool_and_upsample_are_adjoint():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(2, 4, 6, 3))
    b = rng.normal(size=(2, 8, 12, 3))
    up = U.Net.up(a)
    assert up.shape == b.shape and np.allclose(up[:, 5, 7], a[:, 2, 3])
    q = b.reshape(2, 4, 2, 6, 2, 3).sum(axis=(2, 4))
    assert np.isclose((up * b).sum(), (a * q).sum())
    assert np.allclose(U.Net.pool(b), q / 4)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(2)
    base = 8
    p = U.init_params(base, 3).astype(np.float64)
    p[p != 0] += rng.normal(0, 0.05, np.count_nonzero(p))
    for L in U.layers(base):  # non-zero biases
        cin = L["cin_a"] + L["cin_b"]
        M = p[L["offset"]:L["offset"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
        M[:L["cout_real"], 9 * cin] = rng.normal(0, 0.1, L["cout_real"])
    feats = np.concatenate([rng.gamma(0.5, 2.0, (1, 16, 16, 3)), rng.uniform(0, 1, (1, 16, 16, 3)),
                            rng.normal(0, 1, (1, 16, 16, 3)), rng.uniform(0, 1, (1, 16, 16, 1))], axis=3)
    target = rng.gamma(1.0, 1.0, (1, 16, 16, 3))
    loss, g = U.loss_grad(p, feats, target, base, dtype=np.float64)
    picks = []
    for l, L in enumerate(U.layers(base)):
        cin = L["cin_a"] + L["cin_b"]
        cin_real = U.IN_CH if l == 0 else cin
        for _ in range(3):
            co = int(rng.integers(L["cout_real"]))
            picks.append(L["offset"] + co * L["kpad"] + int(rng.integers(9)) * cin + int(rng.integers(cin_real)))
        picks.append(L["offset"] + int(rng.integers(L["cout_real"])) * L["kpad"] + 9 * cin)  # bias
    h = 1e-6
    for i in picks:
        q = p.copy()
        q[i] += h
        lp, _ = U.loss_grad(q, feats, target, base, dtype=np.float64)
        q[i] -= 2 * h
        lm, _ = U.loss_grad(q, feats, target, base, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), (i, g[i], fd)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = U.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even
    assert r[2] == 1.0 + 2 ** -6                # tie -> even (up)
    assert r[3] == 1.0 + 2 ** -7                # below half -> down
    assert r[4] == -2.5 and np.isfinite(r[5])


def test_first_adam_step():
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    p0 = np.ones(4, np.float32)
    p1, m, v = U.adam(p0, g, np.zeros(4, np.float32), np.zeros(4, np.float32), 1, 1e-3)
    np.testing.assert_allclose(p1, p0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-5)
