"""Parity at the headline shape (n=256, K=1024, T=32), through the C-ABI,
against the FP64 oracle on the same seeded inputs.

* Two full K-SVD iterations (Gram, Batch-OMP, dictionary update, error,
  clearDict) on 32768 signals, FP32 and split-BF16x3 GEMM forms, phase by
  phase on identical state and free-running: S/KSVD.cu:55-131,
  S/OMP.cu:254-405, S/DictionaryUpdate.cu:45-203.
* Batch-OMP on 65536 signals of the benchmark's workload (signals coded with
  the dictionary drawn from the signals, as the benchmark does), every GEMM
  form, with the exact one-atom fits counted on their own.

Tolerances (stated here):
  phase by phase          Batch-OMP: no mismatches, coefficients <= 1e-5;
                          dictionary update: D <= 1e-6, gamma <= 1e-5 * max|gamma|;
                          rmse <= 1e-6 relative, err <= 1e-4 relative;
                          clearDict: same picks, D <= 1e-6; the loop entry
                          point equals its phases (rmse 1e-6, D 1e-6)
  rmse, free-running      iteration 0 within 2e-5, iteration 1 within 1e-4
                          (near-tie supports part the trajectories, see the test)
  rmse, exact-fit start   iteration 0 within 1e-3 (the FP64 oracle stops on
                          alpha^2 < 1e-14 at the 1024 signals that are atoms,
                          FP32 goes on with ~1e-7 coefficients); iteration 1
                          within 1e-4, or twice the oracle's own FP32/FP64
                          spread when that is larger
  Batch-OMP supports      identical except documented ties (oracle top-2 gap
                          below tie_rel * max|alpha0| at the first differing
                          step) and exact fits; no mismatches; identical
                          among signals that are not exact fits: >= 99.9 %
                          (FP32, BF16x6), >= 99.8 % (BF16x3, whose operands
                          carry 16 significant bits)
  coefficients            max relative error on identical supports
                          <= 1e-5 (FP32, BF16x3, BF16x6), <= 5e-3 (BF16).  Plain BF16
                          operands (8-bit mantissa) resolve only gaps of about
                          1e-2 * max|alpha0|: >= 60 % identical, every other
                          signal a documented tie at that gap
"""
import sys

import numpy as np
import pytest
import torch

from ksvd_amd import Engine, synth
from tests import helpers as H

pytestmark = pytest.mark.gpu
DEV = "cuda"
n, K, T = 256, 1024, 32


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def cu(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(DEV)


def _perturbed_init(D0, rel=0.05, seed=3):
    rng = np.random.default_rng(seed)
    scale = rel * np.linalg.norm(D0, axis=1, keepdims=True) / np.sqrt(D0.shape[1])
    return (D0 + scale * rng.standard_normal(D0.shape)).astype(np.float32)


_ORACLE = {}


def _oracle_ksvd(o, X, D0, key, iters=2, seed=11):
    if key not in _ORACLE:
        _ORACLE[key] = o.ksvd(X.astype(o.dtype), D0.astype(o.dtype), T, iters, seed=seed)
    return _ORACLE[key]


def _run_gpu(X, D0, form, iters=2, seed=11):
    eng = Engine(n, K, X.shape[0], T, gemm_form=form, perm_seed=seed)
    D, idx, val, cnt, rmse = eng.ExecuteIterations(cu(X), cu(D0), iters)
    torch.cuda.synchronize()
    out = D.cpu().numpy(), np.asarray(rmse)
    eng.close()
    return out


def _check_supports(o64, D, X, got, form, what):
    """Batch-OMP of the engine vs the FP64 oracle on the same FP32 D and X:
    no mismatches (documented ties only), coefficients within 1e-5."""
    gi, gv, gc = got
    D64, X64 = D.astype(np.float64), X.astype(np.float64)
    A064 = o64.correlate(D64, X64)
    ref = o64.batch_omp(A064, o64.gram(D64), T, gaps=True)
    st = {}
    eq, tie, mis, bad = H.support_parity(gi, gc, ref["idx"], ref["cnt"], ref["gap"], 1e-4 * np.abs(A064).max(axis=1),
                                         gv, ref["val"], st)
    same = (gc == ref["cnt"]) & np.all(gi == ref["idx"], axis=1)
    rel = np.linalg.norm(gv[same] - ref["val"][same], axis=1) / (np.linalg.norm(ref["val"][same], axis=1) + 1e-30)
    say(f"  {what} {form} Batch-OMP: identical {eq}/{len(gc)}, ties {st}, mismatches {mis}; coef rel err max "
        f"{rel.max():.2e}")
    assert mis == 0
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "bf16x6"])
def test_c2_ksvd_iteration_step_locked(o64, form):
    """Two K-SVD iterations at n=256, K=1024, T=32 (N = 32768), phase by
    phase: every phase of the engine is fed the engine's own state and checked
    against the FP64 oracle run on that same state (normcols, Gram +
    Batch-OMP, dictionary update, error, clearDict; S/KSVD.cu:55-131).  The
    engine's whole-loop entry point (ksvd_execute_iterations, hipGraph replay
    of iteration 2) is then checked against the phases it is made of."""
    from ksvd_amd import _lib
    N, iters, seed = 32768, 2, 11
    say("C2-shape K-SVD step-locked, N", N, "form", form)
    wl = synth.make("C2", N=N)
    X = wl.X.numpy()
    X64 = X.astype(np.float64)
    D0 = _perturbed_init(wl.D_init.numpy())
    eng = Engine(n, K, N, T, gemm_form=form, perm_seed=seed)
    Xg = cu(X)
    Dg = eng.normcols(cu(D0))
    np.testing.assert_allclose(Dg.cpu().numpy(), o64.normcols(D0.astype(np.float64)), rtol=0, atol=1e-6)
    perms = _lib.perm_sequence(seed, K, iters)
    rmse_steps = []
    for it in range(iters):
        D = Dg.cpu().numpy()
        G = eng.gram(Dg)
        idx, val, cnt = eng.BatchOMP(Dg, Xg, G)
        torch.cuda.synchronize()
        gi, gv, gc = idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy()
        _check_supports(o64, D, X, (gi, gv, gc), form, f"iteration {it}")
        # dictionary update on the engine's Gamma
        eng.du_init()
        eng.updateDictionary(Xg, Dg, idx, val, cnt, perms[it])
        torch.cuda.synchronize()
        Dr, vr, st = o64.dict_update(X64, D.astype(np.float64), gi, gv.astype(np.float64), gc, perms[it])
        D1, v1 = Dg.cpu().numpy(), val.cpu().numpy()
        dd, dv = np.abs(D1 - Dr).max(), np.abs(v1 - vr).max()
        say(f"  iteration {it} dictionary update: D max diff {dd:.2e}, gamma max diff {dv:.2e}, unused atoms {st['usc'][0]}")
        assert dd <= 1e-6 and dv <= 1e-5 * max(1.0, np.abs(vr).max())
        # error and clearDict on the updated state
        rmse, err = eng.errorComputation(Xg, Dg, idx, val, cnt)
        r_ref, e_ref = o64.error(X64, D1.astype(np.float64), gi, v1.astype(np.float64), gc,
                                 bitmap=st["bitmap"])
        say(f"  iteration {it} rmse {rmse:.8f} oracle {r_ref:.8f}")
        assert abs(rmse / r_ref - 1) <= 1e-6
        np.testing.assert_allclose(err.cpu().numpy(), e_ref, rtol=1e-4, atol=1e-6 * e_ref.max())
        rmse_steps.append(rmse)
        err_before = err.cpu().numpy().astype(np.float64)
        eng.clearDict(Xg, Dg, idx, val, cnt, err)
        torch.cuda.synchronize()
        Dc, ec, flags = o64.clear_dict(X64, D1.astype(np.float64), gi, v1.astype(np.float64), gc, st["replaced"],
                                       err_before)
        say(f"  iteration {it} clearDict: atoms replaced {int(flags.sum())}, D max diff "
            f"{np.abs(Dg.cpu().numpy() - Dc).max():.2e}")
        np.testing.assert_array_equal(err.cpu().numpy() == 0, ec == 0)
        assert np.abs(Dg.cpu().numpy() - Dc).max() <= 1e-6
    # the loop entry point runs the same phases
    D_loop, rmse_loop = _run_gpu(X, D0, form, iters=iters, seed=seed)
    say(f"  ksvd_execute_iterations: rmse {rmse_loop} vs phases {rmse_steps}, D max diff "
        f"{np.abs(D_loop - Dg.cpu().numpy()).max():.2e}")
    np.testing.assert_allclose(rmse_loop, rmse_steps, rtol=1e-6)
    assert np.abs(D_loop - Dg.cpu().numpy()).max() <= 1e-6
    eng.close()


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "bf16x6"])
def test_c2_ksvd_iterations_end_to_end(o64, form):
    """Two free-running iterations (N = 32768) against the oracle's own
    trajectory.  Iteration 0 agrees to 2e-5.  After it, the trajectories part
    where a few signals (7 of 32768 here) sit on Batch-OMP near-ties (the
    oracle's top-2 gap below 1e-6 of max|alpha0|): either choice is exact,
    they just differ, and through the dictionary update they move atoms by up
    to ~1e-3 and later supports with them.  So iteration 1 is held to 1e-4,
    and the atoms are not compared one by one (the step-locked test does
    that)."""
    N = 32768
    say("C2-shape K-SVD, 2 iterations, N", N, "perturbed start, form", form)
    wl = synth.make("C2", N=N)
    X = wl.X.numpy()
    D0 = _perturbed_init(wl.D_init.numpy())
    ref = _oracle_ksvd(o64, X, D0, ("pert", N))
    D, rmse = _run_gpu(X, D0, form)
    rel = np.abs(rmse / ref["rmse"] - 1)
    colmax = np.abs(D - ref["D"]).max(axis=1)
    say(f"{form}: rmse {rmse} oracle {ref['rmse']} rel {rel}; D column max diff median "
        f"{np.median(colmax):.3e} max {colmax.max():.3e}")
    assert rel[0] <= 2e-5 and rel[1] <= 1e-4


def test_c2_ksvd_iterations_exact_fit_start(o64, o32):
    """The reference's own start: D_init = K of the signals, normalised."""
    N = 32768
    say("C2-shape K-SVD, 2 iterations, N", N, "exact-fit start (atoms drawn from the signals)")
    wl = synth.make("C2", N=N)
    X = wl.X.numpy()
    D0 = wl.D_init.numpy()
    ref = _oracle_ksvd(o64, X, D0, ("exact", N))
    ref32 = _oracle_ksvd(o32, X, D0, ("exact32", N))
    D, rmse = _run_gpu(X, D0, "fp32")
    rel = np.abs(rmse / ref["rmse"] - 1)
    spread = np.abs(ref32["rmse"] / ref["rmse"] - 1)
    say(f"rmse {rmse} oracle {ref['rmse']} rel {rel}; oracle f32/f64 spread {spread}")
    assert rel[0] <= 1e-3
    assert rel[1] <= max(1e-4, 2 * spread[1])


_OMP = {}


def _bench_workload(o64, N=65536):
    if "ref" not in _OMP:
        wl = synth.make("C2", N=N, seed=0x5EED0002 + 77)
        X = wl.X.numpy()
        D0 = wl.D_init.numpy()
        D = (D0 / np.linalg.norm(D0.astype(np.float64), axis=1, keepdims=True)).astype(np.float32)
        X64, D64 = X.astype(np.float64), D.astype(np.float64)
        A064 = o64.correlate(D64, X64)
        _OMP["ref"] = (X, D, A064, o64.batch_omp(A064, o64.gram(D64), T, gaps=True))
    return _OMP["ref"]


@pytest.mark.parametrize("form,tie_rel,coef_tol,min_frac", [("bf16x6", 1e-4, 1e-5, 0.999), ("bf16x3", 1e-4, 1e-5, 0.998),
                                                             ("fp32", 1e-4, 1e-5, 0.999), ("bf16", 1e-2, 5e-3, 0.6)])
def test_c2_batch_omp_on_benchmark_workload(o64, form, tie_rel, coef_tol, min_frac):
    X, D, A064, ref = _bench_workload(o64)
    N = X.shape[0]
    say("C2 Batch-OMP on the benchmark workload, N", N, "form", form)
    eng = Engine(n, K, N, T, gemm_form=form)
    Dg = cu(D)
    G = eng.gram(Dg)
    idx, val, cnt = eng.BatchOMP(Dg, cu(X), G)
    torch.cuda.synchronize()
    gi, gv, gc = idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy()
    eng.close()
    tol = tie_rel * np.abs(A064).max(axis=1)
    st = {}
    eq, tie, mis, bad = H.support_parity(gi, gc, ref["idx"], ref["cnt"], ref["gap"], tol, gv, ref["val"], st)
    exact = st["precision_stops"]
    frac = eq / (N - exact)
    same = (gc == ref["cnt"]) & np.all(gi == ref["idx"], axis=1)
    rel = np.linalg.norm(gv[same] - ref["val"][same], axis=1) / (np.linalg.norm(ref["val"][same], axis=1) + 1e-30)
    say(f"{form}: identical {eq}/{N} ({frac:.5f} of the {N - exact} non-exact-fit signals), exact fits {exact}, "
        f"gap ties {st['gap_ties']}, mismatches {mis} {bad[:5]}; coef rel err max {rel.max():.3e} "
        f"median {np.median(rel):.3e}")
    assert mis == 0
    assert frac >= min_frac
    assert rel.max() <= coef_tol
