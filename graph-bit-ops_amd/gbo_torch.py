"""torch.autograd.Function over the engine's personalized PageRank.

forward runs gbo_ppr_forward (FP32 fixed point over A^T), backward runs the
hand-written adjoint gbo_ppr_adjoint (one transposed fixed-point solve over
A plus one product for the teleport factor) -- both HIP kernels through the
C ABI.

GPU tensors stay on the device: their addresses go straight to the C ABI,
the work is ordered on torch's current stream, and the outputs are tensors
allocated by torch.  That needs libgbo.so and torch on one HIP runtime
(torch wheels bundle their own; import torch before the first gbo call and
libgbo.so binds to it) -- otherwise the call raises instead of passing
device addresses between two runtimes.  CPU tensors cross as host numpy
arrays.

Both solves run with strict=True: an iteration cap reached before the
tolerance raises GboError (GBO_ENOCONV) rather than returning an
unconverged iterate or gradient.
"""
from __future__ import annotations

import numpy as np
import torch

import gbo


def _device_ok(t: torch.Tensor):
    rt = gbo.hip_runtimes()
    if len(rt) != 1:
        raise RuntimeError("GPU tensors need libgbo.so and torch on one HIP runtime, found "
                           f"{rt}: import torch before the first gbo call")
    if t.device.index not in (None, 0):
        raise RuntimeError("the engine runs on device 0")


class PersonalizedPageRank(torch.autograd.Function):
    """P = PPR(S, alpha) for S of shape [n, K] (K <= 64), alpha a 0-d tensor."""

    @staticmethod
    def forward(ctx, S, alpha, A, AT, outdeg, tol=1e-7, max_iter=1000):
        a = float(alpha)
        n, K = S.shape
        ctx.meta = (A, AT, a, tol, max_iter, S.device, S.dtype)
        if S.is_cuda:
            _device_ok(S)
            s32 = S.detach().to(torch.float32).contiguous()
            od = torch.as_tensor(np.asarray(outdeg, np.int32) if not torch.is_tensor(outdeg) else outdeg,
                                 dtype=torch.int32, device=S.device).contiguous()
            P = torch.empty_like(s32)
            st = torch.cuda.current_stream(S.device).cuda_stream
            it, _ = gbo.ppr_forward_ptr(AT, od.data_ptr(), K, a, s32.data_ptr(), P.data_ptr(), tol, max_iter,
                                        strict=True, stream=st)
            ctx.save_for_backward(s32, P, od)
            ctx.iters = it
            return P.to(S.dtype)
        s_np = S.detach().to("cpu", torch.float32).contiguous().numpy()
        od = np.asarray(outdeg.cpu() if torch.is_tensor(outdeg) else outdeg, np.int32)
        P, it, _ = gbo.ppr_forward(AT, od, s_np, a, tol, max_iter, strict=True)
        ctx.host = (s_np, P, od)
        ctx.iters = it
        return torch.from_numpy(P).to(device=S.device, dtype=S.dtype)

    @staticmethod
    def backward(ctx, grad_p):
        A, AT, a, tol, max_iter, dev, dt = ctx.meta
        if dev.type == "cuda":
            s32, P, od = ctx.saved_tensors
            w = grad_p.detach().to(torch.float32).contiguous()
            gS = torch.empty_like(s32)
            st = torch.cuda.current_stream(dev).cuda_stream
            ga, _ = gbo.ppr_adjoint_ptr(A, AT, od.data_ptr(), s32.shape[1], a, s32.data_ptr(), P.data_ptr(),
                                        w.data_ptr(), gS.data_ptr(), tol, max_iter, strict=True, stream=st)
            return gS.to(dt), torch.tensor(ga, dtype=dt, device=dev), None, None, None, None, None
        s_np, P, od = ctx.host
        w = grad_p.detach().to("cpu", torch.float32).contiguous().numpy()
        gS, ga, _ = gbo.ppr_adjoint(A, AT, od, s_np, P, w, a, tol, max_iter, strict=True)
        return (torch.from_numpy(gS).to(device=dev, dtype=dt), torch.tensor(ga, dtype=dt, device=dev),
                None, None, None, None, None)


def personalized_pagerank(S, alpha, A, AT, outdeg, tol=1e-7, max_iter=1000):
    """Differentiable personalized PageRank of the columns of S."""
    if not torch.is_tensor(alpha):
        alpha = torch.tensor(float(alpha))
    return PersonalizedPageRank.apply(S, alpha, A, AT, outdeg, tol, max_iter)
