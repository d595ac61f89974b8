"""Error-constrained path (the comment at S/KSVD.cu:79: sparse coding with a
threshold instead of a sparsity) through the C-ABI, against the FP64 oracle,
on a subset of the batching configuration (n=128, K=512, T_max=48, per-signal
sparsity 1..48, eps_s = 1.1 sigma_s sqrt(n)).

Tolerances (stated here):
  Batch-OMP    no mismatches; a signal may stop one atom apart from the
               oracle (same prefix) when ||r||^2 sits within FP32 error of
               eps^2 -- at most 1 % of the signals; coefficients on identical
               supports within 1e-4 relative
  phases       dictionary update D <= 1e-6 and gamma <= 1e-5 * max|gamma|,
               rmse <= 1e-6 relative, clearDict the same picks, D <= 1e-6 --
               each phase fed the engine's own state
  forms        the register-basis kernel and the wave form's persistent
               launch, each to the Batch-OMP bar above
  loop         ksvd_execute_iterations_eps (graph replay of iteration 2..)
               equals its phases (rmse 1e-6, D 1e-6) and equals its own
               eager launches bit for bit
"""
import sys

import numpy as np
import pytest
import torch

from ksvd_amd import Engine, synth, _lib
from tests import helpers as H

pytestmark = pytest.mark.gpu
DEV = "cuda"
n, K, T = 128, 512, 48


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def cu(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV)


def _c3(N):
    wl = synth.make("C3", N=N)
    X = wl.X.numpy()
    rng = np.random.default_rng(4)
    D0 = wl.D_init.numpy()
    D0 = (D0 + 0.05 * np.linalg.norm(D0, axis=1, keepdims=True) / np.sqrt(n) * rng.standard_normal(D0.shape))
    return X, D0.astype(np.float32), wl.eps2.numpy().astype(np.float32)


def _check_eps_supports(o64, D, X, eps2, got, what):
    gi, gv, gc = got
    D64, X64 = D.astype(np.float64), X.astype(np.float64)
    A064 = o64.correlate(D64, X64)
    ref = o64.batch_omp(A064, o64.gram(D64), T, xnorm2=(X64 ** 2).sum(1), eps2=eps2.astype(np.float64), gaps=True)
    eq, tie, mis, bad = H.support_parity(gi, gc, ref["idx"], ref["cnt"], ref["gap"], 1e-4 * np.abs(A064).max(axis=1))
    near = sum(1 for s in bad if abs(int(gc[s]) - int(ref["cnt"][s])) == 1 and
               np.array_equal(gi[s, :min(gc[s], ref["cnt"][s])], ref["idx"][s, :min(gc[s], ref["cnt"][s])]))
    same = (gc == ref["cnt"]) & np.all(gi == ref["idx"], axis=1)
    rel = np.linalg.norm(gv[same] - ref["val"][same], axis=1) / (np.linalg.norm(ref["val"][same], axis=1) + 1e-30)
    say(f"  {what}: supports equal {eq}/{len(gc)}, ties {tie}, off-by-one stops {near}, other {mis - near}; "
        f"coef rel err max {rel.max():.2e}; mean count {gc.mean():.2f}")
    assert mis - near == 0 and near <= 0.01 * len(gc)
    assert rel.max() <= 1e-4


def test_c3_ksvd_iteration_step_locked(o64):
    N, iters, seed = 4000, 2, 13
    say("C3-shape error-constrained K-SVD step-locked, N", N)
    X, D0, eps2 = _c3(N)
    X64 = X.astype(np.float64)
    eng = Engine(n, K, N, T, perm_seed=seed)
    Xg, e2g = cu(X), cu(eps2)
    Dg = eng.normcols(cu(D0))
    perms = _lib.perm_sequence(seed, K, iters)
    rmse_steps = []
    for it in range(iters):
        D = Dg.cpu().numpy()
        G = eng.gram(Dg)
        idx, val, cnt = eng.BatchOMP_eps(Dg, Xg, G, e2g)
        torch.cuda.synchronize()
        gi, gv, gc = idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy()
        _check_eps_supports(o64, D, X, eps2, (gi, gv, gc), f"iteration {it} Batch-OMP")
        eng.du_init()
        eng.updateDictionary(Xg, Dg, idx, val, cnt, perms[it])
        torch.cuda.synchronize()
        Dr, vr, st = o64.dict_update(X64, D.astype(np.float64), gi, gv.astype(np.float64), gc, perms[it])
        D1, v1 = Dg.cpu().numpy(), val.cpu().numpy()
        dd, dv = np.abs(D1 - Dr).max(), np.abs(v1 - vr).max()
        say(f"  iteration {it} dictionary update: D max diff {dd:.2e}, gamma max diff {dv:.2e}, unused atoms "
            f"{st['usc'][0]}")
        assert dd <= 1e-6 and dv <= 1e-5 * max(1.0, np.abs(vr).max())
        rmse, err = eng.errorComputation(Xg, Dg, idx, val, cnt)
        r_ref, e_ref = o64.error(X64, D1.astype(np.float64), gi, v1.astype(np.float64), gc, bitmap=st["bitmap"])
        say(f"  iteration {it} rmse {rmse:.8f} oracle {r_ref:.8f}")
        assert abs(rmse / r_ref - 1) <= 1e-6
        rmse_steps.append(rmse)
        err_before = err.cpu().numpy().astype(np.float64)
        eng.clearDict(Xg, Dg, idx, val, cnt, err)
        torch.cuda.synchronize()
        Dc, ec, flags = o64.clear_dict(X64, D1.astype(np.float64), gi, v1.astype(np.float64), gc, st["replaced"],
                                       err_before)
        say(f"  iteration {it} clearDict: atoms replaced {int(flags.sum())}")
        np.testing.assert_array_equal(err.cpu().numpy() == 0, ec == 0)
        assert np.abs(Dg.cpu().numpy() - Dc).max() <= 1e-6
    eng.close()
    outs = []
    for graph in (True, False):
        e2 = Engine(n, K, N, T, perm_seed=seed, use_graph=graph)
        D, idx, val, cnt, rmse = e2.ExecuteIterations_eps(Xg, cu(D0), e2g, iters)
        torch.cuda.synchronize()
        outs.append((D.cpu().numpy(), idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy(), np.asarray(rmse)))
        e2.close()
    say(f"  ksvd_execute_iterations_eps: rmse {outs[0][4]} vs phases {rmse_steps}, D max diff "
        f"{np.abs(outs[0][0] - Dg.cpu().numpy()).max():.2e}")
    np.testing.assert_allclose(outs[0][4], rmse_steps, rtol=1e-6)
    assert np.abs(outs[0][0] - Dg.cpu().numpy()).max() <= 1e-6
    for a, b in zip(*outs):  # graph replay == eager launches
        assert np.array_equal(a, b)


def test_c3_ksvd_iterations_end_to_end(o64):
    """Two free-running error-constrained iterations against the oracle's own
    trajectory (stops within FP32 error of eps^2 may part them slightly)."""
    N = 4000
    X, D0, eps2 = _c3(N)
    say("C3-shape error-constrained K-SVD, 2 iterations, N", N)
    eng = Engine(n, K, N, T, perm_seed=13)
    D, idx, val, cnt, rmse = eng.ExecuteIterations_eps(cu(X), cu(D0), cu(eps2), 2)
    torch.cuda.synchronize()
    ref = o64.ksvd(X.astype(np.float64), D0.astype(np.float64), T, 2, seed=13, eps2=eps2.astype(np.float64))
    rel = np.abs(np.asarray(rmse) / ref["rmse"] - 1)
    say(f"rmse {rmse} oracle {ref['rmse']} rel {rel}; mean count {cnt.float().mean().item():.2f} oracle "
        f"{ref['cnt'].mean():.2f}")
    assert np.all(rel <= 1e-4)
    eng.close()


@pytest.mark.parametrize("launch", [0, 1])
def test_batch_omp_eps_both_forms(o64, launch):
    """Both error-mode kernels the tuning chooses between against the oracle:
    the register-basis form (one workgroup per signal) and the G-streaming
    wave form's persistent work-stealing launch."""
    N = 3000
    X, D0, eps2 = _c3(N)
    what = "register basis, one workgroup per signal" if launch == 0 else "wave form, persistent launch"
    say("eps form:", what)
    eng = Engine(n, K, N, T)
    Dg = eng.normcols(cu(D0))
    G = eng.gram(Dg)
    eng.set_tuning(omp_eps_launch=launch)
    a = eng.BatchOMP_eps(Dg, cu(X), G, cu(eps2))
    torch.cuda.synchronize()
    _check_eps_supports(o64, Dg.cpu().numpy(), X, eps2, tuple(t.cpu().numpy() for t in a), what)
    eng.close()
