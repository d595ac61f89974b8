"""torch.autograd binding of Batch-OMP with the fixed-support adjoint.

`batch_omp(D, X, eng)` normalises the atoms (ksvd_normcols: every Batch-OMP
kernel, like the reference's parOMP, assumes unit-norm atoms, S/OMP.cu:317,
S/KSVD.cu:57), runs Batch-OMP on them and returns the dense coefficients
Gamma (N, K).  The support selection is piecewise constant and carries no
gradient; backward differentiates the fixed-support least squares
gamma_I = (d_I^T d_I)^{-1} d_I^T x with respect to the normalised atoms and X
(ksvd_omp_adjoint_factored, reusing the Cholesky factors the forward stored,
or ksvd_omp_adjoint recomputing them when store_factors=False), then through
the normalisation to the caller's D (ksvd_normcols_adjoint).  Every step,
the layout changes included, is a call through the C-ABI; torch holds the
buffers.
"""
from __future__ import annotations

import torch

from .engine import Engine


class BatchOMPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, X, eng: Engine, store_factors: bool = True):
        D = D.detach().contiguous()
        X = X.detach().contiguous()
        Dn = eng.normcols(D)
        G = eng.gram(Dn)
        A0 = eng.correlate(Dn, X)
        if store_factors:
            idx, val, cnt, Lfac = eng.batch_omp_compact_factors(A0, G)
        else:
            idx, val, cnt = eng.batch_omp_compact(A0, G)
            Lfac = None
        del A0
        dense = eng.compact_to_dense(idx, val, cnt)
        ctx.save_for_backward(D, Dn, X, idx, val, cnt, *( [Lfac] if Lfac is not None else []))
        ctx.eng = eng
        ctx.mark_non_differentiable(idx, cnt)
        return dense, idx, cnt

    @staticmethod
    def backward(ctx, g_dense, g_idx=None, g_cnt=None):
        saved = ctx.saved_tensors
        D, Dn, X, idx, val, cnt = saved[:6]
        Lfac = saved[6] if len(saved) > 6 else None
        if g_dense is None:
            return None, None, None, None
        eng = ctx.eng
        gval = eng.compact_gather(g_dense.contiguous().float(), idx, cnt)
        gDn, gX = eng.omp_adjoint(Dn, X, idx, val, cnt, gval, Lfac=Lfac)
        gD = eng.normcols_adjoint(D, gDn)
        return gD, gX, None, None


def batch_omp(D: torch.Tensor, X: torch.Tensor, eng: Engine, store_factors: bool = True):
    """Differentiable Batch-OMP on the normalised atoms of D: returns
    (Gamma (N, K), idx, cnt)."""
    return BatchOMPFunction.apply(D, X, eng, store_factors)
