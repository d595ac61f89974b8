"""The launch shapes and the project autotuner (ksvd_autotune) through the
C-ABI.  The register-basis Batch-OMP shapes and the GEMM builds compute every
value with the same operations in the same order, so their results must be
identical bit for bit; the dictionary-update grid changes only the order of
its block sums (results within 1e-6).  The autotuner keeps the fastest
candidate, which includes the built-in rule, and caches its choice."""
import sys

import numpy as np
import pytest
import torch

from ksvd_amd import Engine, synth
from ksvd_amd._lib import KsvdError

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _shapes(eng):
    out = []
    for kpl in (1, 2, 4, 8):
        W = 1
        while 64 * kpl * W < eng.K and W < 8:
            W *= 2
        try:
            eng.set_tuning(omp_kpl=kpl, omp_waves=W)
            out.append((kpl, W))
        except KsvdError:
            pass
    eng.set_tuning(omp_kpl=0, omp_waves=0)
    return out


@pytest.mark.parametrize("cfg,N,eps", [("C2", 4096, False), ("C1", 4096, False), ("C3", 4096, True),
                                       ("C4", 4096, False)])
def test_batch_omp_shapes_identical(cfg, N, eps):
    wl = synth.make(cfg, N=N)
    p = synth.CONFIGS[cfg]
    n, K, T = p["n"], p["K"], p["T"]
    eng = Engine(n, K, N, T)
    X = wl.X.to(DEV).contiguous()
    D = eng.normcols(wl.D_true.to(DEV).contiguous())
    G = eng.gram(D)
    e2 = wl.eps2.to(DEV).contiguous() if eps else None

    def run():
        r = eng.BatchOMP_eps(D, X, G, e2) if eps else eng.BatchOMP(D, X, G)
        torch.cuda.synchronize()
        return [t.clone() for t in r]

    ref = run()
    shapes = _shapes(eng)
    say(cfg, "register-basis shapes", shapes)
    assert len(shapes) >= 2
    for kpl, W in shapes:
        eng.set_tuning(omp_kpl=kpl, omp_waves=W)
        got = run()
        for a, b in zip(ref, got):
            assert torch.equal(a, b), (kpl, W)
    eng.close()


@pytest.mark.parametrize("form", ["fp32", "bf16", "bf16x3", "bf16x6"])
def test_gemm_builds_identical(form):
    rng = np.random.default_rng(5)
    A = torch.as_tensor(rng.standard_normal((512, 240)).astype(np.float32)).to(DEV)
    B = torch.as_tensor(rng.standard_normal((3000, 240)).astype(np.float32)).to(DEV)
    eng = Engine(240, 512, 3000, 8, gemm_form=form)
    eng.set_tuning(gemm_wpe=3)
    c3 = eng.AT_times_B(A, B)
    eng.set_tuning(gemm_wpe=2)
    c2 = eng.AT_times_B(A, B)
    torch.cuda.synchronize()
    assert torch.equal(c2, c3)
    eng.close()


def test_du_grid_changes_only_rounding():
    wl = synth.make("C1", N=6000)
    X, D0 = wl.X.to(DEV).contiguous(), wl.D_init.to(DEV).contiguous()
    outs = []
    for blocks in (256, 128, 512, 1024):
        eng = Engine(240, 512, 6000, 16, perm_seed=3)
        eng.set_tuning(du_blocks=blocks)
        D, idx, val, cnt, rmse = eng.ExecuteIterations(X, D0, 2)
        torch.cuda.synchronize()
        outs.append((D.cpu().numpy(), np.asarray(rmse)))
        eng.close()
    for D, r in outs[1:]:
        say("du grid: D max diff", np.abs(D - outs[0][0]).max(), "rmse", r, outs[0][1])
        assert np.abs(D - outs[0][0]).max() <= 1e-5
        np.testing.assert_allclose(r, outs[0][1], rtol=1e-5)


@pytest.mark.parametrize("cfg,N,eps", [("C1", 20000, False), ("C3", 20000, True)])
def test_autotune_picks_and_caches(cfg, N, eps):
    wl = synth.make(cfg, N=N)
    p = synth.CONFIGS[cfg]
    n, K, T = p["n"], p["K"], p["T"]
    X = wl.X.to(DEV).contiguous()
    D0 = wl.D_init.to(DEV).contiguous()
    e2 = wl.eps2.to(DEV).contiguous() if eps else None
    eng = Engine(n, K, N, T)
    t = eng.autotune(D0, X, e2, force=True)
    say(cfg, "tuning", t)
    assert t["source"] == 2
    assert t["omp_ms"] > 0 and t["gemm_ms"] > 0 and t["du_ms"] > 0
    # the built-in rule is one of the candidates
    assert t["omp_ms"] <= t["default_omp_ms"] and t["gemm_ms"] <= t["default_gemm_ms"]
    assert t["du_ms"] <= t["default_du_ms"]
    eng2 = Engine(n, K, N, T)
    t2 = eng2.autotune(D0, X, e2)
    assert t2["source"] == 3
    for k in ("omp_kpl", "omp_waves", "omp_eps_launch", "gemm_wpe", "du_blocks"):
        assert t2[k] == t[k]
    # a tuned loop gives the untuned loop's results (up to the DU grid's rounding)
    if eps:
        a = eng2.ExecuteIterations_eps(X, D0, e2, 2)
        b = Engine(n, K, N, T).ExecuteIterations_eps(X, D0, e2, 2)
    else:
        a = eng2.ExecuteIterations(X, D0, 2)
        b = Engine(n, K, N, T).ExecuteIterations(X, D0, 2)
    torch.cuda.synchronize()
    np.testing.assert_allclose(a[4], b[4], rtol=1e-5)
    eng.close()
    eng2.close()
