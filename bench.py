"""Batch-OMP signals/s and K-SVD iterations/s on one MI355X.

Workload (fits one GPU): the precision-form configuration, n=256 signal
length, K=1024 atoms, N=1,000,000 signals, T=32 atoms per signal, Gaussian
unit-norm atoms with 25 dB SNR, synthetic and seeded, resident in HBM before
timing starts.  A step is one Batch-OMP pass over all N signals with the
headline form: G = D^T D and DtX = D^T X as split-BF16 MFMA GEMMs (operands
split into three BF16 parts, x = hi + mid + lo, six BF16 products per term,
FP32 accumulation: FP32-level accuracy), then the register-basis Batch-OMP
kernel (one workgroup per signal holding Q = G(:,I) L^-T on chip, FP32
throughout); --form bf16x3 uses the two-part split (three products, 16
significant bits), --form fp32 FP32 MFMA.  The plain FP32 form that streams G every step (FP32 GEMMs,
then the FP32-G wave kernel) is timed beside it, and the fraction of signals
whose supports it matches is reported.

    python bench.py [--steps 10] [--warmup 2] [--gpus 1]

Prints ONE JSON line: the metric, the roofline of the dominant kernel
(the Batch-OMP kernel), the CPU baseline (the FP32 port of the reference
algorithm in oracle/, timed on the host cores on bounded samples of the same
signals: a Batch-OMP pass and one K-SVD iteration) and, from the same leg, the
headline result checked against the FP64 oracle on a seeded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Batch-OMP signals/sec + K-SVD iters/sec at n=256,K=1024,T=32; % roofline"
FP32_PEAK_TF = 157.3
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def omp_flops_per_signal(K, T):
    """Reference algorithm's FP32 work per signal (alpha update + Cholesky/solves)."""
    return K * T * (T + 1) + T * (T + 1) * (2 * T + 1) / 2


BF16_MFMA_PEAK_TF = 2500.0


def ksvd_iteration_roof_ms(n, K, N, T, form):
    """Lower bound on one K-SVD iteration from its algorithmic work: the sum
    over phases of max(bytes / HBM peak, MFMA FLOP / MFMA peak + FP32 FLOP /
    FP32 peak).  Phases (s = 4 bytes):
      Gram + correlation + Batch-OMP: bytes N (n s + 8T) (X in, compact
        result out); MFMA 2nK(N + K) (x3 for the split form, on the FP32 MFMA
        peak for the FP32 form); FP32 N F_omp(T);
      dictionary update: bytes n s N (1 + T) + 16 T^2 N; FP32 4nTN + 4T^2 N + 4nK^2;
      error + clearDict: bytes n s N + 8TN; FP32 2nTN."""
    s = 4.0
    bw = HBM_PEAK_GBS * 1e9
    fv = FP32_PEAK_TF * 1e12
    f_mfma = 2.0 * n * K * (N + K)
    t_mfma = f_mfma * {"bf16x3": 3, "bf16x6": 6}[form] / (BF16_MFMA_PEAK_TF * 1e12) if form in ("bf16x3", "bf16x6") else \
        f_mfma / (BF16_MFMA_PEAK_TF * 1e12) if form == "bf16" else f_mfma / fv
    t_omp = max(N * (n * s + 8 * T) / bw, t_mfma + N * omp_flops_per_signal(K, T) / fv)
    t_du = max((n * s * N * (1 + T) + 16.0 * T * T * N) / bw, (4.0 * n * T * N + 4.0 * T * T * N + 4.0 * n * K * K) / fv)
    t_err = max((n * s * N + 8.0 * T * N) / bw, 2.0 * n * T * N / fv)
    return (t_omp + t_du + t_err) * 1000.0


def profiled_traffic(kernel_src):
    """HBM bytes per launch of the Batch-OMP kernel from the committed counter
    summary (profiles/omp_qform_n256_k1024_t32.md), used only when that summary
    was made from the current kernel source (sha256 recorded there)."""
    import hashlib
    import re
    path = os.path.join(ROOT, "profiles", "omp_qform_n256_k1024_t32.md")
    try:
        text = open(path).read()
        src = open(os.path.join(ROOT, "ksvd_amd", "csrc", kernel_src), "rb").read()
    except OSError:
        return None
    h = re.search(r"kernel source sha256: ([0-9a-f]{64})", text)
    t = re.search(r"traffic_bytes_per_launch: ([0-9.e+]+)", text)
    if not h or not t or h.group(1) != hashlib.sha256(src).hexdigest():
        return None
    return float(t.group(1))


def tuning_report(eng, t, T):
    """The autotuner's choice (ksvd_autotune) and its measured times, beside
    the built-in rules' times for the same work."""
    def r(x):
        return round(float(x), 4)
    TT = 8 if T <= 8 else 16 if T <= 16 else 32 if T <= 32 else 48 if T <= 48 else 64
    shape = (f"{t['omp_kpl']} atoms/lane x {t['omp_waves']} waves/signal (basis {TT})" if t["omp_kpl"]
             else "G-streaming wave form")
    return {"source": {2: "measured", 3: "cache"}.get(t["source"], t["source"]),
            "batch_omp_shape": shape,
            "error_mode_kernel": ("G-streaming wave form, persistent work-stealing launch" if t["omp_eps_launch"]
                                  else "register basis, one workgroup per signal"),
            "gemm_workgroups_per_cu": t["gemm_wpe"], "du_blocks": t["du_blocks"],
            "ms": {"batch_omp": r(t["omp_ms"]), "gemm": r(t["gemm_ms"]), "du_sweep": r(t["du_ms"])},
            "built_in_rule_ms": {"batch_omp": r(t["default_omp_ms"]), "gemm": r(t["default_gemm_ms"]),
                                 "du_sweep": r(t["default_du_ms"])},
            "candidates": t.get("log", ""),
            "note": "Batch-OMP timed on up to 65536 signals and the correlation on up to 131072, both scaled to "
                    "all N; the dictionary-update sweep over all N signals, launched kernel by kernel"}


def host_info():
    """Cores this process may use, the CPU model and whether the reference's
    own CPU path (Matlab/fmriKSVD.m under Octave with KSVD-Box) can run."""
    import platform
    import shutil
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, env) if env > 0 else avail
    octave = shutil.which("octave")
    return {"cpu_model": model, "nproc": os.cpu_count(), "cores_available": avail, "omp_num_threads_env": env or None,
            "threads": threads,
            "reference_cpu_path": ("octave found at %s, but KSVD-Box (ksvd.m) is not vendored with the reference "
                                   "(README.md:12 links it for download): not run" % octave) if octave else
                                  "unavailable offline: no MATLAB/Octave on the host and KSVD-Box is not vendored"}


def cpu_baseline(X, D, gpu_result, T, omp_sample, ksvd_sample, parity_sample):
    """The CPU leg: the FP32 port of the reference algorithm
    (oracle/liboracle_f32.so: OpenMP over signals and over the independent
    loops of the dictionary update, blocked GEMM) timed on the host cores,
    and the FP64 oracle as the checker of the GPU's headline result.

      * Batch-OMP signals/s: G, DtX and Batch-OMP on the first `omp_sample`
        signals (1 warm-up run on 4096 signals first);
      * K-SVD iterations/s: one full iteration (G, DtX, Batch-OMP, dictionary
        update, error, clearDict) of the port on the first `ksvd_sample`
        signals, scaled linearly to all N signals;
      * parity: the FP64 oracle's Batch-OMP on a seeded sample of
        `parity_sample` signals against the GPU's idx/val/cnt for them."""
    import numpy as np

    import oracle
    from tests import helpers as H
    info = host_info()
    o = oracle.load("f32")
    used = o.set_threads(info["threads"])
    N = X.shape[0]
    Dn = np.ascontiguousarray(D, dtype=np.float32)
    # Batch-OMP pass
    S = min(omp_sample, N)
    Xs = np.ascontiguousarray(X[:S], dtype=np.float32)
    G = o.at_times_b_blocked(Dn, Dn)
    o.batch_omp(o.at_times_b_blocked(Dn, Xs[:4096]), G, T)  # warm-up: threads, pages
    t0 = time.perf_counter()
    G = o.at_times_b_blocked(Dn, Dn)
    A0 = o.at_times_b_blocked(Dn, Xs)
    o.batch_omp(A0, G, T)
    dt = time.perf_counter() - t0
    del A0
    log(f"cpu Batch-OMP: {S} signals in {dt:.2f} s on {used} threads")
    # one K-SVD iteration on a sample
    Sk = min(ksvd_sample, N)
    t0 = time.perf_counter()
    r = o.ksvd(np.ascontiguousarray(X[:Sk]), Dn, T, 1, seed=0)
    dtk = time.perf_counter() - t0
    log(f"cpu K-SVD iteration: {Sk} signals in {dtk:.2f} s (rmse {r['rmse'][0]:.5f})")
    # FP64 checker on a seeded sample of the GPU's result
    o64 = oracle.load("f64")
    o64.set_threads(info["threads"])
    rng = np.random.default_rng(2024)
    sample = np.sort(rng.choice(N, min(parity_sample, N), replace=False))
    X64 = np.asarray(X[sample], dtype=np.float64)
    D64 = Dn.astype(np.float64)
    t0 = time.perf_counter()
    A064 = o64.correlate(D64, X64)
    ref = o64.batch_omp(A064, o64.gram(D64), T, gaps=True)
    gi, gv, gc = (a[sample] for a in gpu_result)
    st = {}
    eq, tie, mis, bad = H.support_parity(gi, gc, ref["idx"], ref["cnt"], ref["gap"], 1e-4 * np.abs(A064).max(axis=1),
                                         gv, ref["val"], st)
    same = (gc == ref["cnt"]) & np.all(gi == ref["idx"], axis=1)
    rel = np.linalg.norm(gv[same] - ref["val"][same], axis=1) / (np.linalg.norm(ref["val"][same], axis=1) + 1e-30)
    ns = len(sample)
    exact = st["precision_stops"]
    parity = {"sample": f"{ns} signals drawn without replacement (numpy default_rng(2024)) from the {N}",
              "identical_supports": eq, "identical_fraction": round(eq / ns, 6),
              "identical_fraction_excluding_exact_fits": round(eq / max(1, ns - exact), 6),
              "exact_fit_signals": exact, "documented_gap_ties": st["gap_ties"], "mismatches": mis,
              "coef_rel_err_max": float(rel.max()), "coef_rel_err_median": float(np.median(rel)),
              "rule": "ordered supports vs the FP64 oracle; a difference is a documented tie when the oracle's top-2 "
                      "|alpha| gap at the first differing step is below 1e-4 max|alpha0|, or an exact fit (the "
                      "oracle stopped on alpha^2 < 1e-14 at a signal that is an atom, the GPU went on with "
                      "coefficients below 1e-3 of the norm)",
              "check_s": round(time.perf_counter() - t0, 2)}
    log(f"parity vs FP64 oracle: identical {eq}/{ns}, exact fits {exact}, gap ties {st['gap_ties']}, "
        f"mismatches {mis}, coef rel err max {rel.max():.2e}")
    cpu = {"value": round(S / dt, 1), "unit": "signals/s", "cores": used, "kind": "port",
           "sample": f"Batch-OMP pass (G, DtX by blocked GEMM, Batch-OMP) over the first {S} of the {N} signals in "
                     f"FP32, {dt:.2f} s; the port of the reference algorithm in oracle/ (liboracle_f32.so), OpenMP "
                     f"on {used} threads",
           "ksvd_iters_per_s": round(Sk / (N * dtk), 5),
           "ksvd_sample": f"one K-SVD iteration of the port on the first {Sk} signals in {dtk:.2f} s, scaled "
                          f"linearly to {N} signals",
           "host": info}
    return cpu, parity


def compare(args):
    """The reference's end-to-end check (Matlab/fmriKSVD.m:26-70) on this
    project's CPU path: the same X and random initial D through the FP32 port
    of the algorithm (oracle/liboracle_f32.so, OpenMP) and through the GPU
    engine; prints the final errors, their difference and the speedup."""
    import numpy as np
    import torch

    import oracle
    from ksvd_amd import Engine, synth
    from ksvd_amd import _lib

    n, N, K, T = (int(v) for v in args.compare_shape.split(","))
    if args.compare_x:
        X = np.ascontiguousarray(_lib.mat_read(args.compare_x, "X").T, dtype=np.float32)  # (N, n)
        N, n = X.shape
    else:  # fMRI-shaped: AR(1) sources, T of them per signal, mean-removed unit columns
        X = synth.make("C1", N=N, n=n, K=K, T=T).X.numpy()
    D0 = np.random.default_rng(7).standard_normal((K, n)).astype(np.float32)  # D = randn(size(X,1), sources)
    iters = args.compare_iters
    threads = host_info()["threads"]
    o = oracle.load("f32")
    used = o.set_threads(threads)
    t0 = time.perf_counter()
    cpu = o.ksvd(X, D0, T, iters, seed=0)
    t_cpu = time.perf_counter() - t0
    dev = torch.device("cuda", 0)
    Xg, Dg = torch.from_numpy(X).to(dev), torch.from_numpy(D0).to(dev)
    eng = Engine(n, K, N, T, perm_seed=0)
    eng.ExecuteIterations(Xg, Dg, 1)  # warm-up (allocations, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, _, _, rmse = eng.ExecuteIterations(Xg, Dg, iters)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    eng.close()
    e_cpu, e_gpu = float(cpu["rmse"][-1]), float(rmse[-1])
    speedup = t_cpu / t_gpu
    print(f">Time for k-SVD training in CPU: {t_cpu:.6f}", file=sys.stderr)
    print(f">Time for k-SVD training in GPU: {t_gpu:.6f}", file=sys.stderr)
    print(f">Final errors\n  CPU = {e_cpu:.5f}, GPU = {e_gpu:.5f}", file=sys.stderr)
    print(f">Difference = {abs(e_cpu - e_gpu):.5f}", file=sys.stderr)
    print(f">Speedup = {speedup:.2f} ({(1 - 1 / speedup) * 100:.2f} %)", file=sys.stderr)
    print(json.dumps({"compare": {"n": n, "N": N, "K": K, "T": T, "iterations": iters,
                                  "cpu_s": round(t_cpu, 6), "cpu_threads": used, "gpu_s": round(t_gpu, 6),
                                  "rmse_cpu": e_cpu, "rmse_gpu": e_gpu, "difference": abs(e_cpu - e_gpu),
                                  "speedup": round(speedup, 3),
                                  "rmse_per_iter_cpu": [float(v) for v in cpu["rmse"]],
                                  "rmse_per_iter_gpu": [float(v) for v in rmse]}}), flush=True)


def other_config(args):
    """One JSON line for a secondary configuration, each at its full size:
    C1 (n=240, K=512, N=1e5, T=16): K-SVD iterations/s over its 10 iterations;
    C3 (n=128, K=512, N=2e6, per-signal sparsity up to 48, error-constrained
    stop): Gram + correlation + error-mode Batch-OMP signals/s, with the
    generating dictionary; C4 (n=64, K=256, B=65536, T=10): Batch-OMP forward
    storing the Cholesky factors + the fixed-support adjoint reusing them
    (gradients w.r.t. D and X) signals/s, and the recompute mode beside it."""
    import numpy as np
    import torch

    from ksvd_amd import Engine, synth

    if args.gpus != 1:
        print("bench.py runs on one GPU; --gpus must be 1", file=sys.stderr)
        return 2
    dev = torch.device("cuda", 0)
    cfg = args.config
    p = synth.CONFIGS[cfg]
    wl = synth.make(cfg, device=dev)
    X = wl.X.contiguous()
    n, K, N, T = p["n"], p["K"], p["N"], p["T"]
    reps = max(1, args.steps)
    out = {"config": cfg, "workload": dict(n=n, K=K, N=N, T=T), "n_gpus": 1, "data": "synthetic",
           "higher_is_better": True}

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    if cfg == "C1":
        its = p["iters"]
        eng = Engine(n, K, N, T, gemm_form=args.form, omp_variant="auto")
        D0 = wl.D_init.contiguous()
        out["tuning"] = tuning_report(eng, eng.autotune(D0, X), T)
        rm = []

        def run():
            rm[:] = eng.ExecuteIterations(X, D0, its)[4]
        dt = timed(run)
        out.update(metric="K-SVD iterations/s", value=round(its / dt, 2), unit="iterations/s",
                   ms_per_iteration=round(dt * 1000.0 / its, 3), iterations=its, gram_form=args.form,
                   rmse_first_last=[float(rm[0]), float(rm[-1])])
    elif cfg == "C3":
        eng = Engine(n, K, N, T, gemm_form=args.form, omp_variant="auto")
        D = eng.normcols(wl.D_true.contiguous())
        eps2 = wl.eps2.contiguous()
        out["tuning"] = tuning_report(eng, eng.autotune(D, X, eps2), T)
        G = eng.gram(D)
        res = eng.empty_compact()

        def run():
            eng.gram(D, G)
            eng.BatchOMP_eps(D, X, G, eps2, out=res)
        dt = timed(run)
        cnt = res[2].float()
        # the error-constrained K-SVD iteration (iterations after the first replayed from one hipGraph)
        D0 = wl.D_init.contiguous()
        Dout = torch.empty_like(D0)
        its = 3
        eng.ExecuteIterations_eps(X, D0, eps2, 1, D_out=Dout, out=res)  # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rm = eng.ExecuteIterations_eps(X, D0, eps2, its, D_out=Dout, out=res)[4]
        torch.cuda.synchronize()
        dti = (time.perf_counter() - t1) / its
        out.update(metric="error-constrained Batch-OMP signals/s", value=round(N / dt, 1), unit="signals/s",
                   ms_per_pass=round(dt * 1000.0, 3), gram_form=args.form,
                   mean_support=round(float(cnt.mean()), 3), max_support=int(cnt.max()),
                   ksvd_ms_per_iteration=round(dti * 1000.0, 2), ksvd_iterations=its,
                   ksvd_rmse=[float(v) for v in rm])
    else:  # C4
        eng = Engine(n, K, N, T, gemm_form=args.form, omp_variant="auto")
        D = eng.normcols(wl.D_init.contiguous())
        out["tuning"] = tuning_report(eng, eng.autotune(D, X), T)
        G = eng.gram(D)
        res = eng.empty_compact()
        gval = torch.randn((N, T), dtype=torch.float32, device=dev, generator=torch.Generator(device=dev).manual_seed(9))
        gD, gX = torch.empty_like(D), torch.empty_like(X)

        A0 = torch.empty((N, K), dtype=torch.float32, device=dev)
        Lfac = torch.empty((N, T * (T + 1) // 2), dtype=torch.float32, device=dev)

        def run():  # forward storing the Cholesky factors, adjoint reusing them
            eng.gram(D, G)
            eng.correlate(D, X, A0)
            idx, val, cnt, _ = eng.batch_omp_compact_factors(A0, G, out=res, Lfac=Lfac)
            eng.omp_adjoint(D, X, idx, val, cnt, gval, gD=gD, gX=gX, Lfac=Lfac)
        dt = timed(run)

        def run_recompute():  # the low-memory mode: factors recomputed in the adjoint
            eng.gram(D, G)
            idx, val, cnt = eng.BatchOMP(D, X, G, out=res)
            eng.omp_adjoint(D, X, idx, val, cnt, gval, gD=gD, gX=gX)
        dt_r = timed(run_recompute)
        out.update(metric="Batch-OMP forward + adjoint signals/s", value=round(N / dt, 1), unit="signals/s",
                   ms_per_pass=round(dt * 1000.0, 3), gram_form=args.form,
                   recompute_factors_ms_per_pass=round(dt_r * 1000.0, 3))
    eng.close()
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", default=None, choices=["fp32", "bf16", "bf16x3", "bf16x6"],
                    help="GEMM form of G and DtX (default: split-BF16x6 for C2, FP32 for --config C1/C3/C4)")
    ap.add_argument("--omp", default="qform", choices=["qform", "wave"],
                    help="headline Batch-OMP kernel")
    ap.add_argument("--N", type=int, default=1_000_000)
    ap.add_argument("--ksvd-iters", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="timed steps of the plain FP32 form (default: --steps; 0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=262_144, help="signals of the CPU Batch-OMP pass")
    ap.add_argument("--cpu-ksvd-sample", type=int, default=131_072, help="signals of the CPU K-SVD iteration")
    ap.add_argument("--parity-sample", type=int, default=65_536,
                    help="signals of the headline result checked against the FP64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in launch shapes (no ksvd_autotune)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run one device copy of GIB GiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--config", default="C2", choices=["C1", "C2", "C3", "C4"],
                    help="C2 (default): the headline line; C1/C3/C4: that configuration's own measurement "
                         "(K-SVD iterations/s, error-constrained Batch-OMP signals/s, forward + adjoint signals/s)")
    ap.add_argument("--compare", action="store_true",
                    help="only run the fMRI comparison: CPU (FP32 port) vs GPU K-SVD on one problem, then exit")
    ap.add_argument("--compare-x", default=None, help="MAT-file with variable X for --compare (default: synthetic)")
    ap.add_argument("--compare-shape", default="300,3600,9,6", help="n,N,K,T of the synthetic --compare problem")
    ap.add_argument("--compare-iters", type=int, default=30)
    args = ap.parse_args()
    if args.compare:
        return compare(args)
    if args.config != "C2":
        args.form = args.form or "fp32"
        return other_config(args)
    args.form = args.form or "bf16x6"
    if args.copy_probe > 0:
        import torch
        x = torch.ones(int(args.copy_probe * (1 << 28)), device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (DESIGN.md describes how it would shard)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ksvd_amd import Engine, synth

    n, K, T = 256, 1024, 32
    N = args.N
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log("generating workload N =", N)
    wl = synth.make("C2", device=dev, N=N)
    X, D0 = wl.X.contiguous(), wl.D_init.contiguous()
    del wl
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    eng = Engine(n, K, N, T, gemm_form=args.form, omp_variant=args.omp)
    D = eng.normcols(D0)
    tuning = None
    if not args.no_tune:
        log("autotune")
        tuning = tuning_report(eng, eng.autotune(D0, X), T)
        log(f"tuning {tuning}")
    G = torch.empty((K, K), dtype=torch.float32, device=dev)
    idx = torch.empty((N, T), dtype=torch.int32, device=dev)
    val = torch.empty((N, T), dtype=torch.float32, device=dev)
    cnt = torch.empty((N,), dtype=torch.int32, device=dev)
    A0 = torch.empty((N, K), dtype=torch.float32, device=dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def step(record):
        if record:
            ev[0].record(stream)
        eng.gram(D, G)
        eng.correlate(D, X, A0)
        if record:
            ev[1].record(stream)
        eng.batch_omp_compact(A0, G, (idx, val, cnt))
        if record:
            ev[2].record(stream)

    log("warmup", args.warmup)
    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    log("timing", args.steps, "steps")
    gram_ms, omp_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
        torch.cuda.synchronize()
        gram_ms.append(ev[0].elapsed_time(ev[1]))
        omp_ms.append(ev[1].elapsed_time(ev[2]))
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    value = N / (ms_per_step / 1000.0)
    omp_avg = float(np.mean(omp_ms))
    c = cnt.cpu().numpy()
    gpu_result = (idx.cpu().numpy(), val.cpu().numpy(), c)  # the headline result, checked below
    cands, fallbacks = eng.omp_stats()
    log(f"step {ms_per_step:.2f} ms (gemms {np.mean(gram_ms):.3f} ms, batch-omp {omp_avg:.2f} ms), "
        f"{value / 1e6:.2f} M signals/s, mean count {c.mean():.2f}, candidates {cands}, fallbacks {fallbacks}")

    # G-streaming wave form: FP32 GEMMs (G, DtX) + wave-per-signal Batch-OMP
    plain = None
    plain_steps = args.steps if args.plain_steps is None else args.plain_steps
    if plain_steps > 0:
        eng_p = Engine(n, K, N, T, gemm_form="fp32", omp_variant="wave")
        A0p = A0
        Gp = torch.empty_like(G)
        ip, vp, cp = torch.empty_like(idx), torch.empty_like(val), torch.empty_like(cnt)

        def pstep():
            eng_p.gram(D, Gp)
            eng_p.correlate(D, X, A0p)
            eng_p.batch_omp_compact(A0p, Gp, (ip, vp, cp))

        for _ in range(max(1, args.warmup)):
            pstep()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(plain_steps):
            pstep()
        torch.cuda.synchronize()
        p_ms = (time.perf_counter() - t1) * 1000.0 / plain_steps
        same = (cp == cnt) & (ip == idx).all(dim=1)
        plain = {"steps": plain_steps, "warmup": max(1, args.warmup),
                 "ms_per_step": round(p_ms, 3), "signals_per_s": round(N / (p_ms / 1000.0), 1),
                 "headline_speedup": round(p_ms / ms_per_step, 3),
                 "identical_supports_vs_headline": round(float(same.float().mean().item()), 6)}
        log(f"plain FP32 form {p_ms:.2f} ms/step, headline speedup {p_ms / ms_per_step:.2f}x, "
            f"identical supports {plain['identical_supports_vs_headline']:.4f}")
        del A0p, ip, vp, cp
        eng_p.close()

    # K-SVD iterations (G, Batch-OMP, dictionary update, error, clearDict)
    ksvd_ms = None
    if args.ksvd_iters > 0:
        Dout = torch.empty_like(D0)
        eng.ExecuteIterations(X, D0, 1, D_out=Dout, out=(idx, val, cnt))  # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, _, _, _, rmse = eng.ExecuteIterations(X, D0, args.ksvd_iters, D_out=Dout, out=(idx, val, cnt))
        torch.cuda.synchronize()
        ksvd_ms = (time.perf_counter() - t1) * 1000.0 / args.ksvd_iters
        log(f"K-SVD iteration {ksvd_ms:.1f} ms, rmse {rmse}")

    # roofline of the dominant kernel (Batch-OMP): the reference algorithm's
    # FP32 work per signal, F_omp = K T(T+1) + T(T+1)(2T+1)/2, for the signals'
    # actual support sizes, against the FP32 dense peak
    f_sig = np.array([omp_flops_per_signal(K, t) for t in range(T + 1)])
    flops = float(np.sum(f_sig[c]))
    achieved_tf = flops / (omp_avg / 1000.0) / 1e12
    tk = eng.get_tuning()
    kname = (f"omp_qform_kernel<{tk['omp_kpl'] or 4},32,{tk['omp_waves'] or 4},false>" if args.omp == "qform"
             else "omp_wave_kernel<16,false>")
    roofline = {
        "bound": "fp32_valu",
        "achieved": round(achieved_tf, 3),
        "peak": FP32_PEAK_TF,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / FP32_PEAK_TF, 4),
        "traffic": profiled_traffic("omp_qform.hpp") if args.omp == "qform" else None,
        "kernel": kname,
        "kernel_ms": round(omp_avg, 3),
        "work_per_launch": f"F_omp = K T(T+1) + T(T+1)(2T+1)/2 = {omp_flops_per_signal(K, T):.0f} FP32 FLOP per "
                           f"signal at T={T}, summed over the {N} signals' support sizes",
        "note": "compute-bound on FP32 vector FMAs (per-signal basis recurrence, no operand shared across "
                "signals for MFMA), priced at MI355X's 157.3 TFLOP/s FP32 peak; non-packed FP32 VALU issue tops "
                "out at half of that; traffic from the "
                "committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (null if the kernel source changed since)",
    }

    t_roof = ksvd_iteration_roof_ms(n, K, N, T, args.form)

    cpu = parity = None
    if not args.no_cpu:
        Xh = X.cpu().numpy()
        Dh = D.cpu().numpy()
        cpu, parity = cpu_baseline(Xh, Dh, gpu_result, T, args.cpu_sample, args.cpu_ksvd_sample, args.parity_sample)
        del Xh
        log(f"cpu {cpu['value']:.0f} signals/s, {cpu['ksvd_iters_per_s']} K-SVD iterations/s")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "signals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.form == "fp32" else f"f32 (GEMMs {args.form} MFMA, FP32 accumulation)",
        "data": "synthetic",
        "config": {"workload": "C2 precision form: n=256 K=1024 N=1000000 T=32, Gaussian unit atoms, 25 dB SNR",
                   "n": n, "K": K, "N": N, "T": T, "gram_form": args.form, "batch_omp": args.omp,
                   "parallelism": "single GPU"},
        "ksvd_iters_per_s": round(1000.0 / ksvd_ms, 3) if ksvd_ms else None,
        "ksvd_ms_per_iter": round(ksvd_ms, 2) if ksvd_ms else None,
        "ksvd_roofline": ({"t_roof_ms": round(t_roof, 2), "frac": round(t_roof / ksvd_ms, 4),
                           "bound": "sum over phases of max(algorithmic bytes / 8 TB/s, MFMA FLOP / MFMA peak + "
                                    "FP32 FLOP / 157.3 TFLOP/s)"} if ksvd_ms else None),
        "phase_ms": {"gram": round(float(np.mean(gram_ms)), 3), "batch_omp": round(omp_avg, 3)},
        "omp_candidates": cands,
        "omp_fallbacks": fallbacks,
        "plain_fp32_form": plain,
        "parity_vs_fp64_oracle": parity,
        "tuning": tuning,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
