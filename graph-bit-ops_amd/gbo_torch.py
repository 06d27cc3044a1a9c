"""torch.autograd.Function over the engine's personalized PageRank.

forward runs gbo_ppr_forward (FP32 fixed point over A^T), backward runs the
hand-written adjoint gbo_ppr_adjoint (one transposed fixed-point solve over
A plus one product for the teleport factor) -- both HIP kernels through the
C ABI.  Tensors cross as host numpy arrays, so the function accepts CPU or
GPU tensors and returns tensors on the input's device.
"""
from __future__ import annotations

import numpy as np
import torch

import gbo


class PersonalizedPageRank(torch.autograd.Function):
    """P = PPR(S, alpha) for S of shape [n, K] (K <= 64), alpha a 0-d tensor."""

    @staticmethod
    def forward(ctx, S, alpha, A, AT, outdeg, tol=1e-7, max_iter=1000):
        s_np = S.detach().to("cpu", torch.float32).contiguous().numpy()
        a = float(alpha)
        P, it, _ = gbo.ppr_forward(AT, outdeg, s_np, a, tol, max_iter)
        ctx.meta = (A, AT, outdeg, a, tol, max_iter, S.device, S.dtype)
        ctx.s_np = s_np
        ctx.p_np = P
        ctx.iters = it
        return torch.from_numpy(P).to(device=S.device, dtype=S.dtype)

    @staticmethod
    def backward(ctx, grad_p):
        A, AT, outdeg, a, tol, max_iter, dev, dt = ctx.meta
        w = grad_p.detach().to("cpu", torch.float32).contiguous().numpy()
        gS, ga, _ = gbo.ppr_adjoint(A, AT, outdeg, ctx.s_np, ctx.p_np, w, a, tol, max_iter)
        return (torch.from_numpy(gS).to(device=dev, dtype=dt), torch.tensor(ga, dtype=dt, device=dev),
                None, None, None, None, None)


def personalized_pagerank(S, alpha, A, AT, outdeg, tol=1e-7, max_iter=1000):
    """Differentiable personalized PageRank of the columns of S."""
    if not torch.is_tensor(alpha):
        alpha = torch.tensor(float(alpha))
    return PersonalizedPageRank.apply(S, alpha, A, AT, np.asarray(outdeg, np.int32), tol, max_iter)
