"""torch.autograd binding of Batch-OMP with the fixed-support adjoint.

Forward runs the engine's Batch-OMP (support selection is piecewise constant,
so it carries no gradient) and returns the dense coefficient matrix
Gamma (N, K).  Backward differentiates the fixed-support least squares
gamma_I = (D_I^T D_I)^{-1} D_I^T x with respect to D and X through the HIP
adjoint kernel (ksvd_omp_adjoint); both directions go through the C-ABI.
"""
from __future__ import annotations

import torch

from .engine import Engine


class BatchOMPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, X, eng: Engine):
        D = D.detach().contiguous()
        X = X.detach().contiguous()
        G = eng.gram(D)
        A0 = eng.correlate(D, X)
        idx, val, cnt = eng.batch_omp_compact(A0, G)
        N, K = X.shape[0], D.shape[0]
        dense = torch.zeros((N, K), dtype=torch.float32, device=X.device)
        valid = idx >= 0
        rows = torch.arange(N, device=X.device)[:, None].expand_as(idx)
        dense[rows[valid], idx[valid].long()] = val[valid]
        ctx.save_for_backward(D, X, idx, val, cnt)
        ctx.eng = eng
        ctx.mark_non_differentiable(idx, cnt)
        return dense, idx, cnt

    @staticmethod
    def backward(ctx, g_dense, g_idx=None, g_cnt=None):
        D, X, idx, val, cnt = ctx.saved_tensors
        if g_dense is None:
            return None, None, None
        g_dense = g_dense.contiguous().float()
        safe = idx.clamp_min(0).long()
        gval = torch.gather(g_dense, 1, safe) * (idx >= 0)
        gD, gX = ctx.eng.omp_adjoint(D, X, idx, val, cnt, gval.contiguous())
        return gD, gX, None


def batch_omp(D: torch.Tensor, X: torch.Tensor, eng: Engine):
    """Differentiable Batch-OMP: returns (Gamma (N, K), idx, cnt)."""
    return BatchOMPFunction.apply(D, X, eng)
