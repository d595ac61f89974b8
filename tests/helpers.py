"""Independent checks shared by the CPU and GPU tests (numpy only).

These are NOT restatements of the oracle: each computes its answer by a
different route (explicit least squares, SVD, closed forms) so that a slip in
the oracle or in the engine shows up as a mismatch.
"""
from __future__ import annotations

import numpy as np


def unit_rows(A):
    return A / np.linalg.norm(A, axis=1, keepdims=True)


def textbook_omp(D, x, T):
    """Textbook OMP with explicit least squares each step (numpy.linalg.lstsq).

    Selection: argmax |D^T r| over all atoms, lowest index on ties; stops when
    the winner is already selected or its correlation^2 < 1e-14 (the reference's
    rules, S/OMP.cu:292,304).  D: (K, n) unit rows.  Returns (support, coef)."""
    r = x.astype(np.float64).copy()
    S, c = [], np.zeros(0)
    for _ in range(T):
        a = D @ r
        a2 = a * a
        pos = int(np.argmax(a2))  # argmax returns the lowest index among maxima
        if pos in S or a2[pos] < 1e-14:
            break
        S.append(pos)
        c = np.linalg.lstsq(D[S].T, x, rcond=None)[0]
        r = x - D[S].T @ c
    return S, c


def two_ortho_bases(n, rng):
    """[I, Q] with Q a random orthogonal matrix: mutual coherence <= max|Q|."""
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return np.concatenate([np.eye(n), Q.T], axis=0)  # (2n, n) rows are atoms


def dirac_hadamard(n):
    """[I, H/sqrt(n)]: mutual coherence exactly 1/sqrt(n) (n a power of 2)."""
    H = np.array([[1.0]])
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    return np.concatenate([np.eye(n), H / np.sqrt(n)], axis=0)


def dense_gamma(idx, val, cnt, K):
    N, T = idx.shape
    G = np.zeros((N, K))
    for s in range(N):
        for j in range(cnt[s]):
            G[s, idx[s, j]] = val[s, j]
    return G  # (N, K): row s = Gamma(:, s)


def support_parity(idx_a, cnt_a, idx_b, cnt_b, gap, tol, val_a=None, val_b=None, stats=None):
    """Compares ordered supports signal by signal (a = engine, b = oracle).

    Returns (n_equal, n_tie, n_mismatch, mismatch_rows).  A signal that differs
    counts as a documented tie when
      * the oracle's top-2 |alpha| gap at the first differing step is below
        `tol[s]` (an arg-max that FP32 cannot resolve), or
      * the oracle stopped on alpha^2 < 1e-14 (an exact fit in FP64) while the
        engine, whose FP32 residual is ~1e-7 |x| rather than 0, went on: the
        engine's support extends the oracle's and the extra coefficients are
        below 1e-3 of the oracle's coefficient norm (needs val_a, val_b).
    `stats`, if a dict, receives the count of each kind."""
    N = idx_a.shape[0]
    eq = tie = mis = 0
    gap_ties = prec_stops = 0
    bad = []
    for s in range(N):
        ca, cb = int(cnt_a[s]), int(cnt_b[s])
        if ca == cb and np.array_equal(idx_a[s, :ca], idx_b[s, :cb]):
            eq += 1
            continue
        m = min(ca, cb)
        first = m
        for j in range(m):
            if idx_a[s, j] != idx_b[s, j]:
                first = j
                break
        g = gap[s, first] if first < gap.shape[1] else np.inf
        if (val_a is not None and first == cb and ca > cb and
                np.abs(val_a[s, cb:ca]).max() <= 1e-3 * max(np.linalg.norm(val_b[s, :cb]), 1e-30)):
            tie += 1
            prec_stops += 1
        elif 0 <= g < tol[s]:
            tie += 1
            gap_ties += 1
        else:
            mis += 1
            bad.append(s)
    if stats is not None:
        stats.update(gap_ties=gap_ties, precision_stops=prec_stops)
    return eq, tie, mis, bad
