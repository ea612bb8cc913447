"""torch.autograd.Function over the device log-likelihood of the compare.

PoissonLogLik.apply(logE, ctx, use_quals) returns the summed Poisson
log-likelihood of the abundances of the last compare's aligned pairs
(d2g_loglik_forward) as a float64 scalar; its backward is the hand-written
gradient kernel (d2g_loglik_backward).  logE is the 16 x ncol log error matrix
(any device; the 16 x ncol values cross to the library as host arrays).
"""
from __future__ import annotations

import numpy as np
import torch


class PoissonLogLik(torch.autograd.Function):
    @staticmethod
    def forward(fctx, logE: torch.Tensor, dctx, use_quals: bool = True):
        e = logE.detach().to("cpu", torch.float32).contiguous().numpy()
        loss = dctx.loglik_forward(e, use_quals)
        fctx.dctx = dctx
        fctx.logE = e
        fctx.use_quals = use_quals
        fctx.token = getattr(dctx, "_ll_token", 0) + 1
        dctx._ll_token = fctx.token
        fctx.out_device = logE.device
        fctx.out_dtype = logE.dtype
        return torch.tensor(loss, dtype=torch.float64, device=logE.device)

    @staticmethod
    def backward(fctx, grad_out: torch.Tensor):
        dctx = fctx.dctx
        if getattr(dctx, "_ll_token", None) != fctx.token:
            # another forward ran on this context since: redo this one's
            dctx.loglik_forward(fctx.logE, fctx.use_quals)
            dctx._ll_token = fctx.token
        _, g64 = dctx.loglik_backward()
        g = torch.from_numpy(np.ascontiguousarray(g64)).to(fctx.out_device)
        return (g * grad_out.to(g.dtype)).to(fctx.out_dtype), None, None


def poisson_loglik(logE: torch.Tensor, dctx, use_quals: bool = True) -> torch.Tensor:
    return PoissonLogLik.apply(logE, dctx, use_quals)
