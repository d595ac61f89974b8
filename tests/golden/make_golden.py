"""Writes the golden fixtures under tests/golden/ (run in the build container).

Reads the reference's own data files as DATA (scipy.io.loadmat, no code runs):
  K-SVD/K-SVD/input/D.mat                      -> ref_D.npy (300 x 9 float64),
                                                  and the file itself as ref_D.mat
  K-SVD/K-SVD/output/temp.mat,
  K-SVD/x64/Release/output/temp.mat (CUDA_Gamma) -> ref_gamma_structure.json, and
                                                  the first file itself as ref_temp.mat
The two .mat files are kept byte for byte (MATLAB-written, miCOMPRESSED) to
test the engine's own MAT reader against scipy's.
The reference's X.mat is not in the tree, so CUDA_Gamma's values cannot be
reproduced; only its structure (shape, nonzeros per column and per row) is kept.

Also writes perm_golden.json: SplitMix64 / Fisher-Yates permutation streams
computed by an independent pure-Python implementation below, which is itself
pinned to SplitMix64's published first outputs for seed 0.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
M64 = (1 << 64) - 1


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def perm_stream(seed, K, iters):
    buf = list(range(K))
    st = seed
    out = []
    for _ in range(iters):
        for i in range(K - 1, 0, -1):
            st, z = splitmix64(st)
            j = z % (i + 1)
            buf[i], buf[j] = buf[j], buf[i]
        out.append(list(buf))
    return out


def main():
    # published SplitMix64 outputs for state 0 (Vigna's reference implementation)
    st, a = splitmix64(0)
    st, b = splitmix64(st)
    assert a == 0xE220A8397B1DCDAF, hex(a)
    assert b == 0x6E789E6AA1B965F4, hex(b)
    perms = {
        "splitmix64_seed0_first2": [hex(a), hex(b)],
        "streams": [
            {"seed": s, "K": K, "iters": it, "perms": perm_stream(s, K, it)}
            for (s, K, it) in [(0, 9, 3), (12345, 16, 4), (0x5EED0002, 32, 2)]
        ],
    }
    with open(os.path.join(HERE, "perm_golden.json"), "w") as f:
        json.dump(perms, f)

    if not os.path.isdir(REF):
        print("reference tree absent: MAT fixtures left as committed")
        return 0
    import scipy.io as sio
    D = sio.loadmat(os.path.join(REF, "K-SVD/K-SVD/input/D.mat"))["D"]
    np.save(os.path.join(HERE, "ref_D.npy"), np.ascontiguousarray(D, dtype=np.float64))
    import shutil
    shutil.copyfile(os.path.join(REF, "K-SVD/K-SVD/input/D.mat"), os.path.join(HERE, "ref_D.mat"))
    shutil.copyfile(os.path.join(REF, "K-SVD/K-SVD/output/temp.mat"), os.path.join(HERE, "ref_temp.mat"))
    st = {}
    for key, rel in [("K-SVD/K-SVD/output/temp.mat", "K-SVD/K-SVD/output/temp.mat"),
                     ("K-SVD/x64/Release/output/temp.mat", "K-SVD/x64/Release/output/temp.mat")]:
        G = sio.loadmat(os.path.join(REF, rel))["CUDA_Gamma"]
        nz = G != 0
        st[key] = {
            "shape": list(G.shape),
            "col_nnz": sorted(set(int(v) for v in nz.sum(0))),
            "row_nnz": [int(v) for v in nz.sum(1)],
            "finite": bool(np.isfinite(G).all()),
            "abs_max": float(np.abs(G).max()),
            "abs_min_nonzero": float(np.abs(G[nz]).min()),
            "sum": float(G.sum()),
            "abs_sum": float(np.abs(G).sum()),
        }
    with open(os.path.join(HERE, "ref_gamma_structure.json"), "w") as f:
        json.dump(st, f, indent=1)
    print("wrote golden fixtures")
    return 0


if __name__ == "__main__":
    sys.exit(main())
