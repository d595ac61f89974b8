"""Seeded synthetic workloads for the five benchmark configurations.

Every generator returns float32 torch tensors in the engine's layout (the
reference's column-major matrices seen as row-per-column arrays):
    X (N, n)  -- signal s is row s        (X n x N column-major)
    D (K, n)  -- atom k is row k          (D n x K column-major)

Generators run on any torch device; the same seed on the same device gives the
same data.  Tests generate on the CPU and copy to the GPU so that the oracle
and the engine see bit-identical inputs.

Configurations (signal length n, atoms K, signals N, sparsity T):
    C0  64 x 256,   N=10_000,    T=8,  Gaussian unit atoms, 20 dB SNR
    C1  240 x 512,  N=100_000,   T=16, AR(1) rho=0.9 unit atoms, N(0, 0.1^2)
                                       noise, columns mean-removed and unit-normalised
    C2  256 x 1024, N=1_000_000, T=32, Gaussian unit atoms, 25 dB SNR
    C3  128 x 512,  N=2_000_000, T_s ~ U{1..48}, sigma_s log-uniform [1e-3, 1e-1],
                                       eps_s = 1.1 sigma_s sqrt(n)
    C4  64 x 256,   B=65_536,    T=10, C0 generator; linear head W (10 x K) ~ N(0, 1/K)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

SEED_BASE = 0x5EED0000

CONFIGS = {
    "C0": dict(n=64, K=256, N=10_000, T=8, iters=10, snr_db=20.0),
    "C1": dict(n=240, K=512, N=100_000, T=16, iters=10),
    "C2": dict(n=256, K=1024, N=1_000_000, T=32, snr_db=25.0),
    "C3": dict(n=128, K=512, N=2_000_000, T=48),
    "C4": dict(n=64, K=256, N=65_536, T=10, snr_db=20.0, classes=10),
}


@dataclass
class Workload:
    X: torch.Tensor          # (N, n) signals
    D_init: torch.Tensor     # (K, n) initial dictionary (not normalised for C0/C2/C4)
    D_true: torch.Tensor     # (K, n) generating dictionary (unit columns)
    support: torch.Tensor    # (N, Tmax) int64 generating atoms (-1 pad)
    coef: torch.Tensor       # (N, Tmax) generating coefficients (0 pad)
    sigma: torch.Tensor      # (N,) per-signal noise std
    eps2: torch.Tensor | None = None   # (N,) squared stopping threshold (error mode)
    W: torch.Tensor | None = None      # (classes, K) linear head (C4)
    y: torch.Tensor | None = None      # (N, classes) one-hot targets (C4)


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def _distinct_atoms(N, K, T, g, device, chunk=65536):
    """T distinct uniform atoms per signal (top-T of uniform keys)."""
    out = torch.empty((N, T), dtype=torch.int64, device=device)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        keys = torch.rand((e - s, K), generator=g, device=device)
        out[s:e] = keys.topk(T, dim=1).indices
    return out


def _combine(D, support, coef, chunk=32768):
    """x_s = sum_j coef[s,j] d_{support[s,j]} (padding: coef 0, support clamped)."""
    N, T = support.shape
    n = D.shape[1]
    X = torch.empty((N, n), dtype=torch.float32, device=D.device)
    sup = support.clamp_min(0)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        X[s:e] = torch.bmm(coef[s:e, None, :].to(torch.float32), D[sup[s:e]]).squeeze(1)
    return X


def gaussian_dictionary(n, K, g, device):
    D = torch.randn((K, n), generator=g, device=device, dtype=torch.float32)
    return D / D.norm(dim=1, keepdim=True)


def ar1_dictionary(n, K, g, device, rho=0.9):
    e = torch.randn((K, n), generator=g, device=device, dtype=torch.float32)
    a = torch.empty_like(e)
    a[:, 0] = e[:, 0]
    for t in range(1, n):
        a[:, t] = rho * a[:, t - 1] + e[:, t]
    return a / a.norm(dim=1, keepdim=True)


def _signals_snr(D, N, T, snr_db, g, device):
    K, n = D.shape
    support = _distinct_atoms(N, K, T, g, device)
    coef = torch.randn((N, T), generator=g, device=device, dtype=torch.float32)
    Xc = _combine(D, support, coef)
    sigma = Xc.norm(dim=1) / math.sqrt(n) * 10.0 ** (-snr_db / 20.0)
    X = Xc + sigma[:, None] * torch.randn((N, n), generator=g, device=device, dtype=torch.float32)
    return X, support, coef, sigma


def _init_from_signals(X, K, g, device):
    N = X.shape[0]
    cols = torch.randperm(N, generator=g, device=device)[:K]
    return X[cols].clone()


def make(config: str, device="cpu", N: int | None = None, seed: int | None = None, **over) -> Workload:
    """Build a configuration (optionally with fewer signals N for parity tests)."""
    p = dict(CONFIGS[config])
    p.update(over)
    if N is not None:
        p["N"] = int(N)
    idx = int(config[1])
    g = _gen(SEED_BASE + idx if seed is None else seed, device)
    n, K, Nn, T = p["n"], p["K"], p["N"], p["T"]
    if config in ("C0", "C2", "C4"):
        D = gaussian_dictionary(n, K, g, device)
        X, sup, coef, sigma = _signals_snr(D, Nn, T, p["snr_db"], g, device)
        D0 = _init_from_signals(X, K, g, device)
        wl = Workload(X, D0, D, sup, coef, sigma)
        if config == "C4":
            C = p["classes"]
            wl.W = torch.randn((C, K), generator=g, device=device, dtype=torch.float32) / math.sqrt(K)
            cls = torch.randint(0, C, (Nn,), generator=g, device=device)
            wl.y = torch.nn.functional.one_hot(cls, C).to(torch.float32)
        return wl
    if config == "C1":
        D = ar1_dictionary(n, K, g, device)
        sup = _distinct_atoms(Nn, K, T, g, device)
        coef = torch.randn((Nn, T), generator=g, device=device, dtype=torch.float32)
        X = _combine(D, sup, coef)
        X = X + 0.1 * torch.randn((Nn, n), generator=g, device=device, dtype=torch.float32)
        X = X - X.mean(dim=1, keepdim=True)
        X = X / X.norm(dim=1, keepdim=True)
        sigma = torch.full((Nn,), 0.1, device=device)
        D0 = _init_from_signals(X, K, g, device)
        return Workload(X, D0, D, sup, coef, sigma)
    if config == "C3":
        D = gaussian_dictionary(n, K, g, device)
        Tmax = T
        Ts = torch.randint(1, Tmax + 1, (Nn,), generator=g, device=device)
        sup = _distinct_atoms(Nn, K, Tmax, g, device)
        coef = torch.randn((Nn, Tmax), generator=g, device=device, dtype=torch.float32)
        keep = torch.arange(Tmax, device=device)[None, :] < Ts[:, None]
        coef = torch.where(keep, coef, torch.zeros_like(coef))
        sup = torch.where(keep, sup, torch.full_like(sup, -1))
        Xc = _combine(D, sup, coef)
        lo, hi = math.log(1e-3), math.log(1e-1)
        sigma = torch.exp(lo + (hi - lo) * torch.rand((Nn,), generator=g, device=device))
        X = Xc + sigma[:, None] * torch.randn((Nn, n), generator=g, device=device, dtype=torch.float32)
        eps2 = (1.1 * sigma * math.sqrt(n)) ** 2
        D0 = _init_from_signals(X, K, g, device)
        return Workload(X, D0, D, sup, coef, sigma, eps2=eps2)
    raise KeyError(config)
