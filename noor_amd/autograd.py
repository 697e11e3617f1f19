"""torch.autograd binding of the differentiable render.

`render(renderer, stream, albedo, roughness, env)` returns the one-sample
image (H, W, 3) of the renderer's scene with the given parameters; its
backward runs the path-replay adjoint (noor_adjoint) and returns gradients
for albedo (M, 3), roughness (M,) and env (env_h, env_w, 3).  Tensors are
float32 on the renderer's GPU.  The forward writes the parameters into the
renderer, so a backward must run before the next forward on the same
renderer.  torch provides device memory and the loss; the render and its
adjoint are the library's HIP kernels (no fallback).
"""
import ctypes as C

import torch

from . import abi
from ._lib import lib


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _check(r, st):
    if st != 0:
        m = lib().noor_last_error(r._h)
        raise abi.NoorError(st, m.decode() if m else "")


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, stream, albedo, roughness, env):
        albedo = albedo.detach().contiguous()
        roughness = roughness.detach().contiguous()
        env_c = env.detach().contiguous() if env is not None else None
        M = albedo.shape[0]
        assert albedo.shape == (M, 3) and roughness.shape == (M,)
        eh, ew = (env_c.shape[0], env_c.shape[1]) if env_c is not None else (0, 0)
        torch.cuda.synchronize(albedo.device)
        _check(r, lib().noor_set_params_device(r._h, _ptr(albedo), _ptr(roughness), M, _ptr(env_c), ew, eh))
        img = torch.empty((r.camera.h, r.camera.w, 4), dtype=torch.float32, device=albedo.device)
        _check(r, lib().noor_render_sample(r._h, stream, _ptr(img)))
        ctx.save_for_backward(img)
        ctx.r, ctx.stream, ctx.M, ctx.env_shape = r, stream, M, (eh, ew)
        ctx.want_env = env is not None
        return img[..., :3].clone()

    @staticmethod
    def backward(ctx, grad_out):
        (img,) = ctx.saved_tensors
        r = ctx.r
        g = torch.zeros_like(img)
        g[..., :3] = grad_out
        ga = torch.zeros((ctx.M, 3), dtype=torch.float32, device=img.device)
        gr = torch.zeros((ctx.M,), dtype=torch.float32, device=img.device)
        ge = None
        if ctx.want_env:
            eh, ew = ctx.env_shape
            ge = torch.zeros((eh, ew, 3), dtype=torch.float32, device=img.device)
        torch.cuda.synchronize(img.device)
        _check(r, lib().noor_adjoint(r._h, ctx.stream, _ptr(img), _ptr(g), _ptr(ga), _ptr(gr), ctx.M, _ptr(ge)))
        return None, None, ga, gr, ge


def render(renderer, stream, albedo, roughness, env=None):
    """Differentiable one-sample render (see the module docstring)."""
    return _Render.apply(renderer, stream, albedo, roughness, env)
