"""A thin torch.autograd.Function over the HIP replay / adjoint entry points (tests only).

forward(albedo [T,3], texels [h,w]) -> image [H,W,3]: ahip_set_eval_params + ahip_replay, the
per-pixel mean radiance of the detached path-replay estimator of the prepared scene.
backward(dL/dimage) -> (dL/dalbedo, dL/dtexels): ahip_adjoint at the same parameters.
Tensors live on the CPU: the library owns its device copies and takes host pointers.
"""
from __future__ import annotations

import numpy as np
import torch


class ReplayRender(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, texels, backend, sample_start, sample_count, texture):
        a = albedo.detach().cpu().numpy().astype(np.float32)
        t = texels.detach().cpu().numpy().astype(np.float32)
        backend.set_eval_params(a, texture, t)
        img = backend.replay(sample_start, sample_count)
        ctx.backend = backend
        ctx.args = (sample_start, sample_count, texture, tuple(t.shape))
        ctx.params = (a, t)
        return torch.from_numpy(img[..., :3] / img[..., 3:4])

    @staticmethod
    def backward(ctx, grad_img):
        be = ctx.backend
        s0, ns, texture, tshape = ctx.args
        a, t = ctx.params
        be.set_eval_params(a, texture, t)          # the adjoint is taken at the forward's parameters
        da, dt = be.adjoint(s0, ns, grad_img.detach().cpu().numpy().astype(np.float32), tshape)
        return torch.from_numpy(da), torch.from_numpy(dt), None, None, None, None


def render(backend, albedo, texels, sample_start, sample_count, texture=0):
    return ReplayRender.apply(albedo, texels, backend, sample_start, sample_count, texture)
