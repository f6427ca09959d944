"""torch.autograd harness over the C ABI (ngi_render / ngi_render_adjoint).

`render_fn(reflectance, Le)` renders the scene with the given per-primitive
reflectance (D.R / G.R) and emitter radiance tensors (num_primitives x 3,
float64, CPU) and returns the image (H x W x 3); its backward pass calls
the path-replay adjoint with a decorrelated seed.  Torch is only the
differentiation harness: all rendering and gradients run in libngi.so.
"""
import copy

import numpy as np
import torch

from . import render, render_adjoint


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, reflectance, Le, scene, params, adjoint_seed):
        scene.set_params(reflectance.detach().cpu().double().numpy(), Le.detach().cpu().double().numpy())
        img, _ = render(scene, params)
        ctx.scene = scene
        ctx.params = params
        ctx.adjoint_seed = adjoint_seed
        return torch.from_numpy(img)

    @staticmethod
    def backward(ctx, grad_img):
        p = copy.copy(ctx.params)
        p.seed = ctx.adjoint_seed
        gR, gL, _ = render_adjoint(ctx.scene, p, grad_img.detach().cpu().double().numpy())
        return torch.from_numpy(gR), torch.from_numpy(gL), None, None, None


def render_fn(scene, params, adjoint_seed=None):
    """A differentiable function (reflectance, Le) -> image for this scene."""
    def fn(reflectance, Le):
        seed = params.seed + 0x9E3779B9 if adjoint_seed is None else adjoint_seed
        return _Render.apply(reflectance, Le, scene, params, int(seed) & 0xFFFFFFFF)
    return fn


def scene_parameters(scene):
    """(reflectance, Le, reflectance_mask, Le_mask) tensors of a scene."""
    from . import PRIM_D, PRIM_G, PRIM_L
    pr = scene.desc()["primitives"]
    is_r = (pr["type"] & (PRIM_D | PRIM_G)) > 0
    refl = np.where((pr["type"] & PRIM_D)[:, None] > 0, pr["D_R"], pr["G_R"])
    is_l = (pr["type"] & PRIM_L) > 0
    return (torch.from_numpy(np.ascontiguousarray(refl)), torch.from_numpy(np.ascontiguousarray(pr["Le"])),
            torch.from_numpy(is_r), torch.from_numpy(is_l))
