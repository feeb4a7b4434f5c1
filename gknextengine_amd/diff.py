"""torch.autograd binding of the differentiable renderer (gk_render_radiance / gk_adjoint).

Parameters are a (num_materials, 5) float32 CUDA tensor: albedo r, g, b, roughness
(Material.fuzziness), metallic. ``render(params)`` returns the (H, W, 3) radiance image
(mean of ``samples`` paths per pixel); its backward runs the path-replay adjoint.

The backward replays the materials and frame parameters of its own forward call (not
whatever the context holds when backward runs), so several renders may precede one
``backward()``.

Seeds of the adjoint. A Monte Carlo image I and its derivative dI/dθ drawn from the same
paths give a biased gradient of a non-linear loss: for L = ||I - T||², E[2 (I - T) dI/dθ]
differs from 2 (E[I] - T) dE[I]/dθ by the covariance of I and dI/dθ. By default
(``decorrelate=True``) the adjoint therefore replays paths from an independent stream
(global seed ``frame.global_seed ^ ADJOINT_SEED_XOR``), which makes the gradient an unbiased
estimate of the gradient of the loss of the converged image. ``decorrelate=False`` gives the
exact derivative of the rendered estimate itself (same paths), as finite differences of
the same-seed render measure it.
"""
from __future__ import annotations

import torch

from . import Context, FrameParams, Material, check, lib

ADJOINT_SEED_XOR = 0x5BD1E995


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


def _check_params(params: torch.Tensor, base: list[Material]) -> None:
    if not isinstance(params, torch.Tensor) or not params.is_cuda:
        raise ValueError("params must be a CUDA tensor (the kernels read it as device memory)")
    if params.dtype != torch.float32 or tuple(params.shape) != (len(base), 5):
        raise ValueError(f"params must be float32 of shape ({len(base)}, 5), got {params.dtype} {tuple(params.shape)}")


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(fctx, params, ctx, base, frame, decorrelate):
        _check_params(params, base)
        apply_params(ctx, base, params)
        ctx.frame(frame)
        img = torch.empty((ctx.height, ctx.width, 4), dtype=torch.float32, device=params.device)
        torch.cuda.synchronize()
        check(lib().gk_render_radiance(ctx.h, img.data_ptr()), "gk_render_radiance")
        ctx.sync()
        # what the backward must replay: this call's parameters and frame constants
        fctx.gk = (ctx, base, params.detach().clone(), FrameParams.from_buffer_copy(frame), bool(decorrelate))
        return img[..., :3].clone()

    @staticmethod
    def backward(fctx, grad_img):
        ctx, base, params, frame, decorrelate = fctx.gk
        apply_params(ctx, base, params)
        replay = FrameParams.from_buffer_copy(frame)
        if decorrelate:
            replay.global_seed = (frame.global_seed ^ ADJOINT_SEED_XOR) & 0xFFFFFFFF
        ctx.frame(replay)
        dl = torch.zeros((ctx.height, ctx.width, 4), dtype=torch.float32, device=grad_img.device)
        dl[..., :3] = grad_img
        grad = torch.empty(tuple(params.shape), dtype=torch.float32, device=grad_img.device)
        torch.cuda.synchronize()
        check(lib().gk_adjoint(ctx.h, dl.data_ptr(), grad.data_ptr()), "gk_adjoint")
        ctx.sync()
        return grad, None, None, None, None


def render(params: torch.Tensor, ctx: Context, base: list[Material], frame: FrameParams,
           decorrelate: bool = True) -> torch.Tensor:
    """Differentiable render; base holds the non-parameter fields of every material.

    decorrelate: the adjoint replays independent paths (unbiased gradient of a loss of the
    converged image); False replays the forward's own paths (exact derivative of this
    estimate)."""
    return _Render.apply(params, ctx, base, frame, decorrelate)
