"""torch.autograd binding of the differentiable renderer (gk_render_radiance / gk_adjoint).

Parameters are a (num_materials, 5) float32 CUDA tensor: albedo r, g, b, roughness
(Material.fuzziness), metallic. ``render(params)`` returns the (H, W, 3) radiance image
(mean of ``samples`` paths per pixel); its backward runs the path-replay adjoint.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import Context, FrameParams, Material, check, lib


def apply_params(ctx: Context, base: list[Material], params: torch.Tensor) -> None:
    """Writes albedo / roughness / metallic from params into the context's materials."""
    p = params.detach().to("cpu", torch.float32).numpy()
    mats = (Material * len(base))()
    for i, m in enumerate(base):
        mats[i] = m
        mats[i].diffuse[0], mats[i].diffuse[1], mats[i].diffuse[2] = (float(v) for v in p[i, :3])
        mats[i].fuzziness = float(p[i, 3])
        mats[i].metalness = float(p[i, 4])
    check(lib().gk_materials_update(ctx.h, mats, 0, len(base)), "gk_materials_update")


def params_of(materials: list[Material]) -> torch.Tensor:
    rows = [[m.diffuse[0], m.diffuse[1], m.diffuse[2], m.fuzziness, m.metalness] for m in materials]
    return torch.tensor(rows, dtype=torch.float32)


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(fctx, params, ctx, base, frame):
        apply_params(ctx, base, params)
        ctx.frame(frame)
        img = torch.empty((ctx.height, ctx.width, 4), dtype=torch.float32, device=params.device)
        torch.cuda.synchronize()
        check(lib().gk_render_radiance(ctx.h, img.data_ptr()), "gk_render_radiance")
        ctx.sync()
        fctx.gk = (ctx, params.shape)
        return img[..., :3].clone()

    @staticmethod
    def backward(fctx, grad_img):
        ctx, shape = fctx.gk
        dl = torch.zeros((ctx.height, ctx.width, 4), dtype=torch.float32, device=grad_img.device)
        dl[..., :3] = grad_img
        grad = torch.empty(shape, dtype=torch.float32, device=grad_img.device)
        torch.cuda.synchronize()
        check(lib().gk_adjoint(ctx.h, dl.data_ptr(), grad.data_ptr()), "gk_adjoint")
        ctx.sync()
        return grad, None, None, None


def render(params: torch.Tensor, ctx: Context, base: list[Material], frame: FrameParams) -> torch.Tensor:
    """Differentiable render; base holds the non-parameter fields of every material."""
    return _Render.apply(params, ctx, base, frame)
