"""torch.autograd.Function over the differentiable render of libfvc.so.

``render_rays(renderer, rm, pix, pos_enc, mlp_base, mlp_head)`` returns the
colour (C + T, white background) of the pixels ``pix`` (int32 tensor,
y * W + x, on the GPU) as an [n, 3] float32 tensor; its backward calls the
library's adjoint kernels (compositing, colour MLP, density MLP, hash-grid
scatter).  The three parameter tensors are float32 on the GPU in the nerfacc
layout.  The forward state lives in the renderer, so a renderer serves one
forward / backward pair at a time.
"""
import torch

from . import lib


class DiffRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, rm, pix, pos_enc, mlp_base, mlp_head):
        for t in (pix, pos_enc, mlp_base, mlp_head):
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError("render_rays takes contiguous GPU tensors")
        if pix.dtype != torch.int32:
            raise ValueError("pix must be int32")
        params = tuple(t.detach() for t in (pos_enc, mlp_base, mlp_head))
        rgb = torch.empty(pix.numel(), 3, dtype=torch.float32, device=pix.device)
        renderer.diff_forward(rm, pix, params, rgb)
        ctx.renderer = renderer
        ctx.params = params
        return rgb

    @staticmethod
    def backward(ctx, d_rgb):
        d_rgb = d_rgb.contiguous().float()
        grads = tuple(torch.zeros_like(p) for p in ctx.params)
        ctx.renderer.diff_backward(d_rgb, grads)
        return (None, None, None) + grads


def render_rays(renderer, rm, pix, pos_enc, mlp_base, mlp_head):
    return DiffRender.apply(renderer, rm, pix, pos_enc, mlp_base, mlp_head)


class DiffRenderCached(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, rm, pix, g_pool, pos_enc, mlp_base, mlp_head):
        params = tuple(t.detach() for t in (pos_enc, mlp_base, mlp_head))
        rgb = torch.empty(pix.numel(), 3, dtype=torch.float32, device=pix.device)
        renderer.diff_cached_forward(rm, pix, params, rgb)
        ctx.renderer = renderer
        ctx.params = params
        ctx.g_pool = g_pool
        return rgb

    @staticmethod
    def backward(ctx, d_rgb):
        grads = tuple(torch.zeros_like(p) for p in ctx.params)
        ctx.renderer.diff_cached_backward(d_rgb.contiguous().float(), grads, ctx.g_pool)
        return (None, None, None, None) + grads


def render_rays_cached(renderer, rm, pix, g_pool, pos_enc, mlp_base, mlp_head):
    """render_rays against the renderer's render cache.  The gradients of the
    cache contents (per voxel 8 latents and density) are added to g_pool, a
    float32 GPU tensor [n_bricks_used, 512, 10], by the backward."""
    return DiffRenderCached.apply(renderer, rm, pix, g_pool, pos_enc, mlp_base, mlp_head)
