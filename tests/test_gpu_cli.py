"""ksvd_main end to end on the GPU: MAT-files in, the reference's console lines
out (S/main.cu:117-119, S/ReductionKernels.cu:294), CUDA_Gamma / FinalD
written; RMSE checked against the FP64 oracle, Gamma and D against the
engine's own ExecuteIterations (bit for bit: same deterministic path)."""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.io as sio
import torch

from ksvd_amd import Engine, synth
from ksvd_amd import _lib

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "ksvd_amd", "ksvd_main")


def say(*a):
    print("[gpu-test]", *a, flush=True)


def test_driver_end_to_end(o64, tmp_path):
    say("ksvd_main C0 N=2000, 3 iterations")
    wl = synth.make("C0", N=2000)
    X = wl.X.numpy()  # (N, n): signal s is row s, i.e. the column-major n x N matrix
    rng = np.random.default_rng(3)
    D0 = wl.D_init.numpy()
    D0 = (D0 + 0.05 / 8 * np.linalg.norm(D0, axis=1, keepdims=True) * rng.standard_normal(D0.shape)).astype(np.float32)
    n, K, N, T, iters = 64, 256, 2000, 8, 3
    (tmp_path / "input").mkdir()
    (tmp_path / "output").mkdir()
    _lib.mat_write(str(tmp_path / "input" / "X.mat"), {"X": X.T.astype(np.float64)})
    sio.savemat(str(tmp_path / "input" / "D.mat"), {"D": D0.T.astype(np.float64)}, do_compression=True)
    r = subprocess.run([MAIN, "--iters", str(iters), "--T", str(T), "--seed", "11", "--dict-out", "output/D.mat"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=180)
    say(r.stdout.replace("\n", " | "))
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "Dictionary size: 64 x 256" in out and "Input data size: 64 x 2000" in out
    rmse = [float(m) for m in re.findall(r"Iteration \d\d / 3 complete, RMSE = ([0-9.]+)", out)]
    assert len(rmse) == iters
    assert re.search(r"Final Execution Time =>\t\t[0-9.]+ ms \([0-9.]+ sec\)", out)
    assert out.rstrip().endswith("# Exiting!")
    # the fmriKSVD harness reads the last 'RMSE = ' and the seconds in '(...)'
    k = out.rfind("RMSE = ")
    assert float(out[k + 7:k + 14]) == pytest.approx(rmse[-1], abs=1e-5)

    ref = o64.ksvd(X.astype(np.float64), D0.astype(np.float64), T, iters, seed=11)
    np.testing.assert_allclose(rmse, ref["rmse"], rtol=2e-5, atol=1e-5)  # printed with 5 decimals

    eng = Engine(n, K, N, T, perm_seed=11)
    Dg, idx, val, cnt, rm = eng.ExecuteIterations(torch.from_numpy(X).cuda(), torch.from_numpy(D0).cuda(), iters)
    torch.cuda.synchronize()
    idx, val, cnt = idx.cpu().numpy(), val.cpu().numpy(), cnt.cpu().numpy()
    G = np.zeros((K, N))
    for s in range(N):
        G[idx[s, :cnt[s]], s] = val[s, :cnt[s]]
    CG = sio.loadmat(str(tmp_path / "output" / "temp.mat"))["CUDA_Gamma"]
    assert CG.shape == (K, N)
    np.testing.assert_array_equal(CG, G.astype(np.float32).astype(np.float64))
    FD = sio.loadmat(str(tmp_path / "output" / "D.mat"))["FinalD"]
    np.testing.assert_array_equal(FD, Dg.cpu().numpy().T.astype(np.float64))
    eng.close()


def test_driver_phases_match_execute_iterations(tmp_path):
    """The C++ facade (include/ksvd_amd.hpp) driven phase by phase -- normcols,
    then per iteration the Gram, BatchOMP, initialize, updateDictionary,
    errorComputation, clearDict (ksvd_main --phases) -- against its
    KSVD::ExecuteIterations (the default): the same atom orders and kernels;
    only the error is formed differently (from X and D instead of the
    maintained residuals), so RMSE and dictionaries agree to rounding."""
    say("ksvd_main --phases vs ExecuteIterations, C0 N=2000, 3 iterations")
    wl = synth.make("C0", N=2000)
    X = wl.X.numpy()
    rng = np.random.default_rng(5)
    D0 = wl.D_init.numpy()
    D0 = (D0 + 0.05 / 8 * np.linalg.norm(D0, axis=1, keepdims=True) * rng.standard_normal(D0.shape)).astype(np.float32)
    (tmp_path / "input").mkdir()
    (tmp_path / "output").mkdir()
    _lib.mat_write(str(tmp_path / "input" / "X.mat"), {"X": X.T.astype(np.float64)})
    _lib.mat_write(str(tmp_path / "input" / "D.mat"), {"D": D0.T.astype(np.float64)})
    outs = {}
    for mode in ("loop", "phases"):
        args = [MAIN, "--iters", "3", "--T", "8", "--seed", "4", "--out", f"output/{mode}.mat",
                "--dict-out", f"output/{mode}_D.mat"] + (["--phases"] if mode == "phases" else [])
        r = subprocess.run(args, cwd=str(tmp_path), capture_output=True, text=True, timeout=180)
        say(mode, r.stdout.replace("\n", " | ")[-300:])
        assert r.returncode == 0, r.stderr
        rm = [float(m) for m in re.findall(r"Iteration \d\d / 3 complete, RMSE = ([0-9.]+)", r.stdout)]
        assert len(rm) == 3
        outs[mode] = (rm, sio.loadmat(str(tmp_path / "output" / f"{mode}_D.mat"))["FinalD"],
                      sio.loadmat(str(tmp_path / "output" / f"{mode}.mat"))["CUDA_Gamma"])
    (ra, Da, Ga), (rb, Db, Gb) = outs["loop"], outs["phases"]
    np.testing.assert_allclose(rb, ra, rtol=0, atol=2e-5)  # printed with 5 decimals
    np.testing.assert_allclose(Db, Da, rtol=0, atol=1e-5)
    assert np.array_equal(Ga != 0, Gb != 0)
    np.testing.assert_allclose(Gb, Ga, rtol=0, atol=1e-4)


def test_compare_mode():
    """bench.py --compare (the fmriKSVD.m check, Matlab/fmriKSVD.m:26-70): CPU FP32 port
    and GPU engine on the same problem end at the same RMSE, iteration by iteration."""
    import json
    import sys
    say("bench.py --compare 300x3600 K=9 T=6, 10 iterations")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--compare", "--compare-iters", "10"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    c = json.loads(r.stdout.strip().splitlines()[-1])["compare"]
    assert (c["n"], c["N"], c["K"], c["T"], c["iterations"]) == (300, 3600, 9, 6, 10)
    assert ">Difference = " in r.stderr and ">Speedup = " in r.stderr
    np.testing.assert_allclose(c["rmse_per_iter_gpu"], c["rmse_per_iter_cpu"], rtol=1e-4)
    assert c["difference"] < 1e-5
