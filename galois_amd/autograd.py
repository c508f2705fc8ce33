"""torch.autograd wrapper of the differentiable PageRank (the only PyTorch-facing
part of the package).

Forward: gx_pagerank_weighted (Jacobi on the transpose).  Backward: the
hand-written adjoint gx_pagerank_adjoint (transposed power iteration on the
same CSR) with c = dL/dx; it returns dL/dw (CSR edge order) and dL/dt.  Only
the converged forward vector is saved for the backward pass.
"""
from __future__ import annotations

import torch

from . import pagerank_adjoint, pagerank_weighted


class PageRankFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weights, teleport, graph, alpha, tol, max_iter, precision):
        w = weights.detach().float().contiguous()
        t = teleport.detach().float().contiguous()
        x, _ = pagerank_weighted(graph, w, t, alpha, tol, max_iter, precision)
        ctx.save_for_backward(w, x)
        ctx.graph, ctx.alpha, ctx.tol, ctx.max_iter, ctx.precision = graph, alpha, tol, max_iter, precision
        ctx.wdtype, ctx.tdtype = weights.dtype, teleport.dtype
        return x

    @staticmethod
    def backward(ctx, grad_x):
        w, x = ctx.saved_tensors
        gw, gt, _ = pagerank_adjoint(ctx.graph, w, x, grad_x.detach().float().contiguous(), ctx.alpha, ctx.tol,
                                     ctx.max_iter, ctx.precision)
        return gw.to(ctx.wdtype), gt.to(ctx.tdtype), None, None, None, None, None


def pagerank_fn(graph, weights, teleport, alpha=0.85, tol=1e-9, max_iter=200, precision="fp32"):
    if weights.numel() != graph.m or teleport.numel() != graph.n:
        raise ValueError("weights must have m entries and teleport n entries")
    return PageRankFunction.apply(weights, teleport, graph, alpha, tol, max_iter, precision)
