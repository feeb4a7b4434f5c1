"""CPU oracle of the learned denoiser (test infrastructure only: the product never calls it).

Restates, in numpy, the U-Net that gk_unet_* runs on the GPU: the layer table and the
parameter layout stated in include/gknx_rt.h (gk_unet_layer), the input transform, the
forward pass, the training loss, its gradient with respect to every parameter and one
Adam step. There is no reference counterpart (the reference's learned denoiser is a
prebuilt library, src/Rendering/PathTracing/PathTracingRenderer.cpp:244-268); the
restatement is pinned by finite differences of its own loss (float64) and by direct
per-pixel convolution sums on small layers (tests/test_unet_oracle.py).

``bf16=True`` rounds to bfloat16 (nearest even) at every point where the GPU stores
bfloat16: weights, packed input, each layer's output, pooled tensors, the loss gradient,
input gradients, upsampled-gradient sums and pooled-gradient accumulations. Products are
then summed in float32, as the MFMA accumulates them, so GPU and oracle differ only by
summation order.
"""
from __future__ import annotations

import math

import numpy as np

IN_CH = 10
IN_PAD = 16
LAYERS = 12


def round16(c: int) -> int:
    return (c + 15) // 16 * 16


def round32(c: int) -> int:
    return (c + 31) // 32 * 32


def channels(base: int) -> list[int]:
    return [round16(int(math.ceil(base * m))) for m in (1.0, 1.5, 2.0, 3.0, 4.0)]


def layers(base: int) -> list[dict]:
    """Layer table: first input, skip input, output channels, upsampling, ReLU, level."""
    c = channels(base)
    table = [
        (IN_PAD, 0, c[0], 0, 1, 0), (c[0], 0, c[0], 0, 1, 0), (c[0], 0, c[1], 0, 1, 1), (c[1], 0, c[2], 0, 1, 2),
        (c[2], 0, c[3], 0, 1, 3), (c[3], 0, c[4], 0, 1, 4), (c[4], c[3], c[3], 1, 1, 3), (c[3], c[2], c[2], 1, 1, 2),
        (c[2], c[1], c[1], 1, 1, 1), (c[1], c[0], c[0], 1, 1, 0), (c[0], 0, c[0], 0, 1, 0), (c[0], 0, 16, 0, 0, 0),
    ]
    out, off = [], 0
    for i, (ca, cb, co, up, relu, lev) in enumerate(table):
        kpad = round32(9 * (ca + cb) + 1)
        out.append(dict(cin_a=ca, cin_b=cb, cout=co, cout_real=3 if i == LAYERS - 1 else co, kpad=kpad, up=up,
                        relu=relu, level=lev, offset=off))
        off += co * kpad
    return out


def num_params(base: int) -> int:
    L = layers(base)[-1]
    return L["offset"] + L["cout"] * L["kpad"]


def real_params(base: int) -> int:
    """Weights and biases that are not padding."""
    n = 0
    for i, L in enumerate(layers(base)):
        cin = IN_CH if i == 0 else L["cin_a"] + L["cin_b"]
        n += L["cout_real"] * (9 * cin + 1)
    return n


def init_params(base: int, seed: int) -> np.ndarray:
    """He-normal weights of the real inputs, zero bias (numpy generator; the GPU draws its own)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(num_params(base), np.float32)
    for i, L in enumerate(layers(base)):
        cin = L["cin_a"] + L["cin_b"]
        cin_real = IN_CH if i == 0 else cin
        sd = math.sqrt((1.0 if i == LAYERS - 1 else 2.0) / (9 * cin_real))
        W = p[L["offset"]:L["offset"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
        for tap in range(9):
            W[:L["cout_real"], tap * cin:tap * cin + cin_real] = rng.normal(0, sd, (L["cout_real"], cin_real))
    return p


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


class Net:
    def __init__(self, params: np.ndarray, base: int, bf16_mode: bool = False, dtype=np.float32):
        self.base = base
        self.L = layers(base)
        self.bf = bf16_mode
        self.dt = dtype
        p = np.asarray(params, dtype)
        if bf16_mode:
            p = bf16(p.astype(np.float32)).astype(dtype)
        self.W, self.b = [], []
        for L in self.L:
            cin = L["cin_a"] + L["cin_b"]
            M = p[L["offset"]:L["offset"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
            self.W.append(M[:, :9 * cin].reshape(L["cout"], 9, cin).transpose(1, 2, 0).copy())  # (tap, cin, cout)
            self.b.append(M[:, 9 * cin].copy())

    def r(self, x):
        return bf16(x).astype(self.dt) if self.bf else x

    # ------------------------------------------------------------ building blocks
    @staticmethod
    def up(a):
        return np.repeat(np.repeat(a, 2, axis=1), 2, axis=2)

    @staticmethod
    def pool(s):
        n, h, w, c = s.shape
        q = s.reshape(n, h // 2, 2, w // 2, 2, c)
        return ((q[:, :, 0, :, 0] + q[:, :, 0, :, 1]) + q[:, :, 1, :, 0] + q[:, :, 1, :, 1]) * 0.25

    def conv(self, x, l):
        """3x3 'same' convolution with zero padding; x: (n, h, w, cin)."""
        n, h, w, c = x.shape
        xp = np.zeros((n, h + 2, w + 2, c), self.dt)
        xp[:, 1:h + 1, 1:w + 1] = x
        y = np.zeros((n * h * w, self.W[l].shape[2]), self.dt)
        for tap in range(9):
            dy, dx = tap // 3 - 1, tap % 3 - 1
            xs = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w].reshape(-1, c)
            y += xs @ self.W[l][tap]
        y += self.b[l]
        return y.reshape(n, h, w, -1)

    def conv_dx(self, g, l):
        n, h, w, co = g.shape
        cin = self.W[l].shape[1]
        dxp = np.zeros((n, h + 2, w + 2, cin), self.dt)
        g2 = g.reshape(-1, co)
        for tap in range(9):
            dy, dx = tap // 3 - 1, tap % 3 - 1
            dxp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] += (g2 @ self.W[l][tap].T).reshape(n, h, w, cin)
        return dxp[:, 1:h + 1, 1:w + 1]

    def conv_dw(self, x, g, l):
        n, h, w, c = x.shape
        xp = np.zeros((n, h + 2, w + 2, c), self.dt)
        xp[:, 1:h + 1, 1:w + 1] = x
        g2 = g.reshape(-1, g.shape[3])
        dW = np.zeros((9, c, g.shape[3]), self.dt)
        for tap in range(9):
            dy, dx = tap // 3 - 1, tap % 3 - 1
            dW[tap] = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w].reshape(-1, c).T @ g2
        return dW, g2.sum(axis=0)

    # ------------------------------------------------------------ network
    def pack(self, feats):
        f = np.asarray(feats, self.dt)
        x = np.zeros(f.shape[:3] + (IN_PAD,), self.dt)
        x[..., :3] = np.log1p(np.maximum(f[..., :3], 0))
        x[..., 3:IN_CH] = f[..., 3:IN_CH]
        return self.r(x)

    def layer_input(self, l, T):
        L = self.L[l]
        src = {0: "X", 1: "A0", 2: "P0", 3: "P1", 4: "P2", 5: "P3", 6: "B", 7: "D3", 8: "D2", 9: "D1", 10: "D0",
               11: "E"}[l]
        skip = {6: "S3", 7: "S2", 8: "S1", 9: "S0"}.get(l)
        a = T[src]
        if L["up"]:
            a = self.up(a)
        return a if skip is None else np.concatenate([a, T[skip]], axis=3)

    OUT = ["A0", "S0", "S1", "S2", "S3", "B", "D3", "D2", "D1", "D0", "E", "Y"]

    def forward(self, feats):
        T = {"X": self.pack(feats)}
        for l in range(LAYERS):
            y = self.conv(self.layer_input(l, T), l)
            if self.L[l]["relu"]:
                y = np.maximum(y, 0)
            name = self.OUT[l]
            T[name] = y if name == "Y" else self.r(y)
            if 1 <= l <= 4:
                T["P%d" % (l - 1)] = self.r(self.pool(T[name]))
        self.T = T
        return T["Y"][..., :3]

    def loss_grad(self, feats, target):
        """Loss mean((z - log1p(max(t,0)))^2) and its gradient in the parameter layout."""
        z = self.forward(feats)
        T = self.T
        t = np.log1p(np.maximum(np.asarray(target, self.dt), 0))
        d = z - t
        cnt = d.size
        loss = float((d.astype(np.float64) ** 2).sum() / cnt)
        gY = np.zeros(T["Y"].shape, self.dt)
        gY[..., :3] = self.r(2.0 / cnt * d) if self.bf else 2.0 / cnt * d
        G = {"Y": gY}
        grads = [None] * LAYERS
        skip_of = {6: "S3", 7: "S2", 8: "S1", 9: "S0"}
        pooled_from = {2: "S0", 3: "S1", 4: "S2", 5: "S3"}
        for l in range(LAYERS - 1, -1, -1):
            L = self.L[l]
            name = self.OUT[l]
            g = G[name]
            if L["relu"]:
                g = g * (T[name] > 0)
            x = self.layer_input(l, T)
            dW, db = self.conv_dw(x, g, l)
            grads[l] = (dW, db)
            if l == 0:
                continue
            dx = self.r(self.conv_dx(g, l))
            ca = L["cin_a"]
            da = dx[..., :ca]
            if l in skip_of:
                G[skip_of[l]] = dx[..., ca:]
            if L["up"]:
                n, h, w, c = da.shape
                q = da.reshape(n, h // 2, 2, w // 2, 2, c)
                da = self.r(((q[:, :, 0, :, 0] + q[:, :, 0, :, 1]) + q[:, :, 1, :, 0]) + q[:, :, 1, :, 1])
            src = {1: "A0", 2: "P0", 3: "P1", 4: "P2", 5: "P3", 6: "B", 7: "D3", 8: "D2", 9: "D1", 10: "D0",
                   11: "E"}[l]
            if src.startswith("P"):
                s = pooled_from[l]
                G[s] = self.r(G[s] + self.up(da) * 0.25)
            else:
                G[src] = da
        flat = np.zeros(num_params(self.base), np.float64)
        for l, L in enumerate(self.L):
            cin = L["cin_a"] + L["cin_b"]
            dW, db = grads[l]
            M = flat[L["offset"]:L["offset"] + L["cout"] * L["kpad"]].reshape(L["cout"], L["kpad"])
            M[:, :9 * cin] = dW.transpose(2, 0, 1).reshape(L["cout"], 9 * cin)
            M[:, 9 * cin] = db
        return loss, flat


def loss_grad(params, feats, target, base, bf16_mode=False, dtype=np.float32):
    return Net(params, base, bf16_mode, dtype).loss_grad(feats, target)


def infer(params, feats, base, bf16_mode=False, dtype=np.float32):
    """Linear radiance max(expm1(z), 0) for an (n, h, w, 10) feature batch."""
    z = Net(params, base, bf16_mode, dtype).forward(feats)
    return np.maximum(np.expm1(z), 0)


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new params, m, v (float32)."""
    g = np.asarray(grad, np.float32)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = np.float32(1.0 / (1.0 - b1 ** step))
    c2 = np.float32(1.0 / (1.0 - b2 ** step))
    p = params - lr * (m * c1) / (np.sqrt(v * c2) + np.float32(eps))
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
