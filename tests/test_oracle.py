"""Pins the CPU oracle (oracle/ksvd_oracle.c) before anything is checked
against it.  Every check derives its expected answer independently: closed
forms, textbook algorithms (explicit least squares, SVD), brute force, known
answers, the reference's own data files, and mathematical invariants."""
import json
import os

import numpy as np
import pytest

from ksvd_amd import synth
from tests import helpers as H

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def c0(N=400, seed=1):
    wl = synth.make("C0", N=N, seed=seed)
    return wl.X.numpy().astype(np.float64), wl


# ---- normcols / GEMM --------------------------------------------------------------
def test_normcols_unit_and_scaled_identity(o64):
    rng = np.random.default_rng(0)
    D = rng.standard_normal((20, 7))
    out = o64.normcols(D)
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose(out, D / np.linalg.norm(D, axis=1, keepdims=True), rtol=1e-14)
    S = np.eye(6)[:4] * np.array([[3.0], [0.5], [7.0], [2.0]])  # scaled identity rows
    np.testing.assert_allclose(o64.normcols(S), np.eye(6)[:4], atol=0)


def test_at_times_b_brute_force_long_double(o64):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((13, 37))
    B = rng.standard_normal((29, 37))
    C = o64.at_times_b(A, B)  # (colsB, colsA)
    ref = np.einsum("mt,jt->jm", A.astype(np.longdouble), B.astype(np.longdouble))
    np.testing.assert_allclose(C, ref.astype(np.float64), rtol=1e-13, atol=1e-13)
    I = np.eye(16)
    X = rng.standard_normal((5, 16))
    np.testing.assert_array_equal(o64.gram(I), np.eye(16))
    np.testing.assert_array_equal(o64.correlate(I, X), X)


def test_at_times_b_blocked(o64, o32):
    """The blocked product of the CPU baseline: in FP64 (no reassociation) the
    same sums in the same order as the plain loop, so bit-identical; in the
    FP32 port within FP32 rounding of the long-double product."""
    rng = np.random.default_rng(4)
    for rows, ca, cb in [(37, 13, 29), (64, 255, 130), (256, 66, 67)]:
        A = rng.standard_normal((ca, rows))
        B = rng.standard_normal((cb, rows))
        np.testing.assert_array_equal(o64.at_times_b_blocked(A, B), o64.at_times_b(A, B))
        ref = np.einsum("mt,jt->jm", A.astype(np.longdouble), B.astype(np.longdouble)).astype(np.float64)
        bound = np.abs(B) @ np.abs(A).T
        C32 = o32.at_times_b_blocked(A.astype(np.float32), B.astype(np.float32))
        A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
        ref32 = np.einsum("mt,jt->jm", A32.astype(np.longdouble), B32.astype(np.longdouble)).astype(np.float64)
        assert np.all(np.abs(C32 - ref32) <= rows * 1.2e-7 * bound)
        assert np.abs(ref32 - ref).max() < 1e-4 * bound.max()


# ---- Batch-OMP --------------------------------------------------------------------
def test_omp_orthonormal_dictionary_closed_form(o64):
    """G = I: alpha after each step is alpha0 with the chosen entries zeroed,
    so the support is the top-T |alpha0| in order and gamma = alpha0 there."""
    rng = np.random.default_rng(2)
    n, T, N = 32, 6, 50
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    D = Q.T  # K = n orthonormal atoms
    X = rng.standard_normal((N, n))
    A0 = X @ D.T
    r = o64.batch_omp(A0, D @ D.T, T)
    for s in range(N):
        order = np.argsort(-np.abs(A0[s]), kind="stable")[:T]
        assert r["cnt"][s] == T
        np.testing.assert_array_equal(r["idx"][s], order)
        np.testing.assert_allclose(r["val"][s], A0[s, order], atol=1e-12)


def test_omp_exact_recovery_incoherent(o64):
    """mu([I, H/sqrt(n)]) = 1/sqrt(n) = 1/8 < 1/(2T-1) for T <= 4 (Tropp 2004):
    OMP recovers any T-sparse combination exactly."""
    rng = np.random.default_rng(3)
    n, T, N = 64, 4, 200
    D = H.dirac_hadamard(n)
    K = D.shape[0]
    S = np.stack([rng.choice(K, T, replace=False) for _ in range(N)])
    c = rng.standard_normal((N, T)) + np.sign(rng.standard_normal((N, T))) * 0.5
    X = np.einsum("st,stn->sn", c, D[S])
    r = o64.batch_omp(X @ D.T, D @ D.T, T)
    for s in range(N):
        assert r["cnt"][s] == T
        assert set(r["idx"][s]) == set(S[s])
        got = dict(zip(r["idx"][s], r["val"][s]))
        for a, v in zip(S[s], c[s]):
            assert abs(got[a] - v) < 1e-10


def test_omp_matches_textbook_least_squares(o64):
    """Tiny random problems: the Cholesky Batch-OMP selects the same ordered
    support as OMP with an explicit least-squares solve per step."""
    rng = np.random.default_rng(4)
    for trial in range(200):
        n, K, T = 8, int(rng.integers(5, 13)), int(rng.integers(1, 5))
        D = H.unit_rows(rng.standard_normal((K, n)))
        x = rng.standard_normal(n)
        r = o64.batch_omp((D @ x)[None, :], D @ D.T, T)
        S, c = H.textbook_omp(D, x, T)
        cnt = int(r["cnt"][0])
        assert list(r["idx"][0, :cnt]) == S, trial
        np.testing.assert_allclose(r["val"][0, :cnt], c, atol=1e-10)


def test_omp_invariants_on_c0(o64):
    """Residual orthogonal to the chosen atoms; residual non-increasing in T."""
    X, wl = c0(300)
    D = o64.normcols(wl.D_init.numpy())
    G, A0 = o64.gram(D), o64.correlate(D, X)
    prev = np.linalg.norm(X, axis=1)
    for T in (1, 2, 4, 8):
        r = o64.batch_omp(A0, G, T)
        res = np.empty(len(X))
        for s in range(len(X)):
            c = r["cnt"][s]
            I = r["idx"][s, :c]
            rs = X[s] - D[I].T @ r["val"][s, :c]
            if c:
                assert np.abs(D[I] @ rs).max() < 1e-9
                # the Cholesky factor L L^T = G_II  <=>  gamma solves the normal equations
                np.testing.assert_allclose(G[np.ix_(I, I)] @ r["val"][s, :c], A0[s, I], atol=1e-9)
            res[s] = np.linalg.norm(rs)
        assert np.all(res <= prev + 1e-12)
        prev = res


def test_omp_special_cases(o64):
    rng = np.random.default_rng(5)
    n, K = 16, 24
    D = H.unit_rows(rng.standard_normal((K, n)))
    G = D @ D.T
    # zero signal -> empty support
    r = o64.batch_omp(np.zeros((1, K)), G, 5)
    assert r["cnt"][0] == 0 and np.all(r["idx"][0] == -1)
    # a signal equal to an atom -> that atom alone
    r = o64.batch_omp((D @ (3.0 * D[7]))[None, :], G, 5)
    assert r["cnt"][0] == 1 and r["idx"][0, 0] == 7 and abs(r["val"][0, 0] - 3.0) < 1e-12


def test_omp_pivot_break_marks_slot_like_reference(o64):
    """Duplicated atom (d0 = d1 = e0) with an inconsistent correlation: step 0
    takes atom 1 (alpha^2 = 16), then alpha_0 = 3.5 - 4 = -0.5 wins, w = G(1,0)
    = 1 and 1 - ||w||^2 = 0 <= 1e-14, so the loop breaks with atom 0 marked in
    the slot map (slot 2) but not accepted (S/OMP.cu:313,331-333)."""
    E = np.eye(8)
    D = np.stack([E[0], E[0], E[1], E[2]])
    a0 = np.array([[3.5, 4.0, 0.1, 0.1]])
    r = o64.batch_omp(a0, D @ D.T, 3, dense=True, c_init=np.full((1, 3), 7.0))
    assert r["cnt"][0] == 1 and r["idx"][0, 0] == 1 and r["val"][0, 0] == 4.0
    assert r["pending"][0] == 0
    assert r["sel"][0, 1] == 1 and r["sel"][0, 0] == 2
    assert r["c"][0, 0] == 4.0 and r["c"][0, 1] == 7.0  # slot 2 never written: stale


def test_omp_error_mode(o64):
    X, wl = c0(300)
    D = o64.normcols(wl.D_init.numpy())
    G, A0 = o64.gram(D), o64.correlate(D, X)
    xn2 = (X * X).sum(1)
    # eps = 0 on noisy data: identical to sparsity mode
    r0 = o64.batch_omp(A0, G, 8)
    re = o64.batch_omp(A0, G, 8, xnorm2=xn2, eps2=np.zeros(len(X)))
    np.testing.assert_array_equal(r0["idx"], re["idx"])
    # eps > 0: residual <= eps at stop, or T_max reached; not stopped early
    eps2 = 0.3 * xn2
    re = o64.batch_omp(A0, G, 8, xnorm2=xn2, eps2=eps2)
    for s in range(len(X)):
        c = re["cnt"][s]
        I = re["idx"][s, :c]
        res2 = np.sum((X[s] - D[I].T @ re["val"][s, :c]) ** 2)
        assert res2 <= eps2[s] * (1 + 1e-9) or c == 8
        if c > 0:  # one atom fewer was not enough
            rp = o64.batch_omp(A0[s:s + 1], G, c - 1) if c > 1 else None
            if rp is not None:
                I1 = rp["idx"][0, :c - 1]
                prev = np.sum((X[s] - D[I1].T @ rp["val"][0, :c - 1]) ** 2)
            else:
                prev = xn2[s]
            assert prev > eps2[s]


# ---- dictionary update ------------------------------------------------------------
def _omp_state(o64, X, D, T):
    G, A0 = o64.gram(D), o64.correlate(D, X)
    return o64.batch_omp(A0, G, T)


def test_du_error_nonincreasing_and_unit_atoms(o64):
    X, wl = c0(600, seed=2)
    D = o64.normcols(wl.D_init.numpy())
    r = _omp_state(o64, X, D, 8)
    Gd = H.dense_gamma(r["idx"], r["val"], r["cnt"], D.shape[0])
    e0 = np.sum((X - Gd @ D) ** 2)
    perm = np.random.default_rng(0).permutation(D.shape[0])
    D1, v1, st = o64.dict_update(X, D, r["idx"], r["val"], r["cnt"], perm)
    used = np.bincount(r["idx"][r["idx"] >= 0], minlength=D.shape[0]) > 0
    np.testing.assert_allclose(np.linalg.norm(D1[used], axis=1), 1.0, atol=1e-12)
    Gd1 = H.dense_gamma(r["idx"], v1, r["cnt"], D.shape[0])
    if st["usc"][0] == 0:  # no unused-atom replacement in this sweep
        assert np.sum((X - Gd1 @ D1) ** 2) <= e0 * (1 + 1e-12)
    # supports unchanged: only values move
    assert np.all((v1 != 0) <= (r["idx"] >= 0))


def test_du_converges_to_rank1_svd(o64):
    """Repeated sweeps (block coordinate descent) reach a fixed point where every
    atom is the top left singular vector of its restricted error E_p and its
    coefficients are sigma_1 times the right singular vector."""
    rng = np.random.default_rng(7)
    n, K, N, T = 12, 6, 80, 2
    D = H.unit_rows(rng.standard_normal((K, n)))
    X = rng.standard_normal((N, n))
    r = _omp_state(o64, X, D, T)
    idx, cnt, val = r["idx"], r["cnt"], r["val"]
    perm = np.arange(K)
    for _ in range(3000):
        D, val, _ = o64.dict_update(X, D, idx, val, cnt, perm)
    Gd = H.dense_gamma(idx, val, cnt, K)
    for p in range(K):
        I = np.nonzero(Gd[:, p])[0]
        if len(I) == 0:
            continue
        E = (X[I] - Gd[I] @ D + np.outer(Gd[I, p], D[p])).T  # n x |I|
        U, S, Vt = np.linalg.svd(E, full_matrices=False)
        assert abs(abs(U[:, 0] @ D[p]) - 1) < 1e-8
        np.testing.assert_allclose(np.abs(Gd[I, p]), S[0] * np.abs(Vt[0]), atol=1e-6)


def test_du_single_atom_one_sweep_closed_form(o64):
    """K = 1: the restricted error of the only atom is X itself, so one sweep is
    d = X gamma^T / ||X gamma^T|| and gamma'_s = d^T x_s (one power-iteration
    step), computed here directly."""
    rng = np.random.default_rng(11)
    n, N = 20, 50
    X = rng.standard_normal((N, n))
    D = H.unit_rows(rng.standard_normal((1, n)))
    idx = np.zeros((N, 1), np.int32)
    cnt = np.ones(N, np.int32)
    cnt[[3, 17]] = 0  # two signals without the atom stay out of it
    idx[[3, 17]] = -1
    val = rng.standard_normal((N, 1)) * (cnt[:, None] > 0)
    a = (val[:, 0] * (cnt > 0)) @ X
    d = a / np.linalg.norm(a)
    D1, v1, st = o64.dict_update(X, D, idx, val, cnt, np.array([0], np.uint32))
    np.testing.assert_allclose(D1[0], d, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(v1[:, 0], np.where(cnt > 0, X @ d, 0.0), rtol=1e-12, atol=1e-14)


def test_du_threads_do_not_change_results(o64):
    """The parallel loops of the update give each output to one thread."""
    X, wl = c0(500, seed=6)
    D = o64.normcols(wl.D_init.numpy())
    r = _omp_state(o64, X, D, 8)
    perm = np.random.default_rng(1).permutation(D.shape[0])
    before = o64.set_threads(0)
    o64.set_threads(1)
    D1, v1, _ = o64.dict_update(X, D, r["idx"], r["val"], r["cnt"], perm)
    o64.set_threads(max(2, before))
    D2, v2, _ = o64.dict_update(X, D, r["idx"], r["val"], r["cnt"], perm)
    o64.set_threads(before)
    np.testing.assert_array_equal(D1, D2)
    np.testing.assert_array_equal(v1, v2)


def test_du_unused_atom_replacement(o64):
    """An atom nobody uses (visited first) is replaced by the normalised signal
    with the largest error among the first min(N, max_signals) candidates."""
    X, wl = c0(300, seed=3)
    D = o64.normcols(wl.D_init.numpy())
    K = D.shape[0]
    r = _omp_state(o64, X, D, 8)
    used = np.bincount(r["idx"][r["idx"] >= 0], minlength=K)
    p = int(np.argmin(used))
    idx, val, cnt = r["idx"].copy(), r["val"].copy(), r["cnt"].copy()
    # drop atom p from every support (keep slot order compact)
    for s in range(len(X)):
        keep = [j for j in range(cnt[s]) if idx[s, j] != p]
        i2, v2 = idx[s, keep], val[s, keep]
        idx[s] = -1; val[s] = 0
        idx[s, :len(keep)] = i2; val[s, :len(keep)] = v2
        cnt[s] = len(keep)
    perm = np.concatenate([[p], np.setdiff1d(np.arange(K), [p])])
    for max_signals in (5000, 50):
        Gd = H.dense_gamma(idx, val, cnt, K)
        err = np.sum((X - Gd @ D) ** 2, axis=1)[:min(len(X), max_signals)]
        best = int(np.argmax(err))
        D1, _, st = o64.dict_update(X, D, idx, val, cnt, perm, max_signals=max_signals)
        np.testing.assert_allclose(D1[p], X[best] / np.linalg.norm(X[best]), atol=1e-14)
        assert st["replaced"][p] == 1 and st["bitmap"][best] == 0 and st["usc"][0] == 1
        assert st["unused"][best] == len(X) - 1  # swap-remove with the last live entry


# ---- error / clearDict ----------------------------------------------------------------
def test_error_closed_form(o64):
    X, wl = c0(200, seed=4)
    D = o64.normcols(wl.D_init.numpy())
    r = _omp_state(o64, X, D, 8)
    Gd = H.dense_gamma(r["idx"], r["val"], r["cnt"], D.shape[0])
    bitmap = (np.arange(len(X)) % 3 != 0).astype(np.int32)
    rmse, err = o64.error(X, D, r["idx"], r["val"], r["cnt"], bitmap)
    e = np.sum((X - Gd @ D) ** 2, axis=1)
    assert abs(rmse - np.sqrt(e.sum() / X.size)) < 1e-12
    np.testing.assert_allclose(err, e * bitmap, rtol=1e-12)


def test_clear_dict_rules(o64):
    rng = np.random.default_rng(8)
    n, K, N, T = 16, 8, 60, 2
    D = H.unit_rows(rng.standard_normal((K, n)))
    D[1] = H.unit_rows((D[0] + 0.05 * rng.standard_normal(n))[None])[0]  # coherent with atom 0
    X = rng.standard_normal((N, n))
    r = _omp_state(o64, X, D, T)
    rmse, err = o64.error(X, D, r["idx"], r["val"], r["cnt"])
    replaced = np.zeros(K, np.int32)
    D1, err1, flags = o64.clear_dict(X, D, r["idx"], r["val"], r["cnt"], replaced, err)
    assert flags[0] == 1  # |d0^T d1| > 0.8
    worst = int(np.argmax(err))
    np.testing.assert_allclose(D1[0], X[worst] / np.linalg.norm(X[worst]), atol=1e-14)
    assert err1[worst] == 0
    # replaced atoms are skipped even when coherent
    replaced[:] = 0
    replaced[0] = 1
    _, _, flags = o64.clear_dict(X, D, r["idx"], r["val"], r["cnt"], replaced, err)
    assert flags[0] == 0
    # usecount < 4 -> replaced: atom 5 used by exactly two signals
    idx = np.empty((N, T), np.int32)
    for s_ in range(N):
        idx[s_] = [s_ % 5, 6 + (s_ % 2)]
    idx[0, 0], idx[1, 0] = 5, 5
    val = np.ones((N, T))
    cnt = np.full(N, T, np.int32)
    _, _, flags = o64.clear_dict(X, D, idx, val, cnt, np.zeros(K, np.int32), err)
    assert np.sum(idx == 5) == 2 and flags[5] == 1
    val[:, :] = 1e-4  # gamma^2 = 1e-8 <= 1e-7: nothing counts as used
    _, _, flags = o64.clear_dict(X, D, idx, val, cnt, np.zeros(K, np.int32), err)
    assert np.all(flags[2:] == 1)


# ---- adjoint ------------------------------------------------------------------------
def test_ls_adjoint_matches_finite_differences(o64):
    rng = np.random.default_rng(9)
    n, K, N, T = 10, 12, 6, 4
    D = H.unit_rows(rng.standard_normal((K, n)))
    X = rng.standard_normal((N, n))
    r = _omp_state(o64, X, D, T)
    idx, cnt = r["idx"], r["cnt"]
    val = o64.ls_fixed(D, X, idx, cnt)
    np.testing.assert_allclose(val, r["val"], atol=1e-10)  # OMP's gamma is the LS solution
    W = rng.standard_normal((N, T))

    def loss(Dm, Xm):
        return float(np.sum(W * o64.ls_fixed(Dm, Xm, idx, cnt)))

    gD, gX = o64.ls_adjoint(D, X, idx, cnt, val, W)
    h = 1e-6
    for (k, t) in [(idx[0, 0], 0), (idx[1, 1], 3), (idx[2, 0], 7), (0, 0)]:
        Dp, Dm = D.copy(), D.copy()
        Dp[k, t] += h
        Dm[k, t] -= h
        fd = (loss(Dp, X) - loss(Dm, X)) / (2 * h)
        assert abs(fd - gD[k, t]) <= 1e-6 * max(1.0, abs(fd))
    for (s, t) in [(0, 0), (3, 5), (5, 9)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[s, t] += h
        Xm[s, t] -= h
        fd = (loss(D, Xp) - loss(D, Xm)) / (2 * h)
        assert abs(fd - gX[s, t]) <= 1e-6 * max(1.0, abs(fd))


# ---- permutation ----------------------------------------------------------------------
def test_perm_golden_vectors(o64):
    g = json.load(open(os.path.join(GOLD, "perm_golden.json")))
    assert g["splitmix64_seed0_first2"] == ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4"]
    for st in g["streams"]:
        got = o64.perms(st["K"], st["iters"], st["seed"])
        np.testing.assert_array_equal(got, np.array(st["perms"], dtype=np.uint32))


# ---- full loop and reference fixture ------------------------------------------------------
def test_ksvd_rmse_decreases(o64):
    X, wl = c0(1000, seed=5)
    res = o64.ksvd(X, wl.D_init.numpy(), 8, 6, seed=3)
    assert np.all(np.diff(res["rmse"]) < 0), res["rmse"]
    np.testing.assert_allclose(np.linalg.norm(res["D"], axis=1), 1.0, atol=1e-12)


def test_reference_fixture_structure(o64):
    """The reference's recorded output (CUDA_Gamma, 9 x 3600 for its D.mat of
    300 x 9 and T = 6, 30 iterations) has exactly 6 nonzeros in every column and
    uses every atom.  Its X.mat is not in the tree, so signals are synthesised
    from the reference D (6 random atoms, N(0,1) coefficients, 20 dB noise) and
    the same structure is required of the oracle's run."""
    st = json.load(open(os.path.join(GOLD, "ref_gamma_structure.json")))
    for rec in st.values():
        assert rec["shape"] == [9, 3600] and rec["col_nnz"] == [6] and rec["finite"]
        assert min(rec["row_nnz"]) > 0
    Dref = np.load(os.path.join(GOLD, "ref_D.npy")).T.copy()  # (9, 300)
    rng = np.random.default_rng(10)
    N = 3600
    S = np.stack([rng.choice(9, 6, replace=False) for _ in range(N)])
    c = rng.standard_normal((N, 6))
    Dn = Dref / np.linalg.norm(Dref, axis=1, keepdims=True)
    Xc = np.einsum("st,stn->sn", c, Dn[S])
    X = Xc + rng.standard_normal(Xc.shape) * (np.linalg.norm(Xc, axis=1, keepdims=True) / np.sqrt(300) * 0.1)
    res = o64.ksvd(X, Dref, 6, 30, seed=0)
    assert np.all(res["cnt"] == 6)
    assert np.all(np.bincount(res["idx"].ravel(), minlength=9) > 0)
    assert np.all(np.isfinite(res["val"]))


def test_ksvd_error_mode_with_zero_threshold_is_sparsity_mode(o64):
    """eps^2 = 0 never stops a signal whose residual is not exactly zero, so
    the error-constrained loop reduces to the sparsity-mode loop."""
    X, wl = c0(300, seed=8)
    D0 = wl.D_init.numpy().astype(np.float64) + 0.01  # no signal is an atom
    a = o64.ksvd(X, D0, 6, 2, seed=3)
    b = o64.ksvd(X, D0, 6, 2, seed=3, eps2=np.zeros(len(X)))
    # (the RMSE total is an OpenMP reduction: last-bit differences between runs)
    np.testing.assert_allclose(a["rmse"], b["rmse"], rtol=1e-12)
    np.testing.assert_array_equal(a["D"], b["D"])
    np.testing.assert_array_equal(a["cnt"], b["cnt"])
    # and a threshold above every signal's energy codes nothing
    c = o64.ksvd(X, D0, 6, 1, seed=3, eps2=np.full(len(X), 1e30))
    assert np.all(c["cnt"] == 0)
