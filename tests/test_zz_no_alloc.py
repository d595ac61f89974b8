"""Allocation-free Batch-OMP (ksvd_batch_omp_from_signals_noalloc) against the
allocating entry point (ksvd_batch_omp_from_signals), through the C-ABI.

Both run the same kernels on the same inputs, so the results must be identical
bit for bit (integer supports and counts exactly; coefficients as the same
FP32 values).  Two calls in a row on one handle, the second with fewer signals,
so that the second call runs over the buffer the first one left behind.
"""
import sys

import numpy as np
import pytest
import torch

from ksvd_amd import Engine, synth

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _inputs(N, seed):
    wl = synth.make("C2", N=N, seed=seed)
    X = wl.X.to(DEV).contiguous()
    D = wl.D_true.to(DEV).contiguous()
    return X, D


def _reference(X, D, N, form):
    eng = Engine(256, 1024, N, 32, gemm_form=form)
    G = eng.gram(D)
    idx, val, cnt = eng.BatchOMP(D, X[:N].contiguous(), G)
    torch.cuda.synchronize()
    out = (idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy(), G)
    eng.close()
    return out


@pytest.mark.parametrize("form", ["bf16x3", "fp32"])
def test_noalloc_matches_allocating_entry_point(form):
    N1, N2 = 20000, 3000  # 3 chunks of up to 8192 signals, then 1 smaller chunk
    say("noalloc vs allocating Batch-OMP, C2 shape, form", form, "N", N1, "then", N2)
    X1, D = _inputs(N1, 101)
    X2, _ = _inputs(N2, 202)
    eng = Engine(256, 1024, N1, 32, gemm_form=form)
    G = eng.gram(D)
    for X, N in ((X1, N1), (X2, N2)):
        idx, val, cnt = eng.BatchOMP_noalloc(D, X, G, N=N)
        torch.cuda.synchronize()
        ri, rv, rc, Gr = _reference(X, D, N, form)
        assert torch.equal(G, Gr)
        gi, gv, gc = idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy()
        say(f"N={N}: counts equal {np.array_equal(gc, rc)}, supports equal {np.array_equal(gi, ri)}, "
            f"coefficients equal {np.array_equal(gv, rv)}")
        np.testing.assert_array_equal(gc, rc)
        np.testing.assert_array_equal(gi, ri)
        np.testing.assert_array_equal(gv, rv)
        assert eng.omp_stats() == (0, 0)
    eng.close()


def test_noalloc_rejects_more_signals_than_the_handle():
    eng = Engine(64, 256, 100, 8)
    X = torch.zeros((200, 64), device=DEV)
    D = torch.nn.functional.normalize(torch.randn((256, 64), device=DEV), dim=1)
    G = eng.gram(D)
    with pytest.raises(Exception, match="outside"):
        eng.BatchOMP_noalloc(D, X, G, N=200, out=(torch.empty((200, 8), dtype=torch.int32, device=DEV),
                                                  torch.empty((200, 8), device=DEV),
                                                  torch.empty((200,), dtype=torch.int32, device=DEV)))
    eng.close()
