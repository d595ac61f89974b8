"""MI355X-native K-SVD engine: Batch-OMP and the approximate K-SVD dictionary
update on hand-written HIP/CDNA4 kernels behind a C-ABI (libksvd_amd.so).

    from ksvd_amd import Engine
    eng = Engine(n, K, N, T)          # one handle, bound to torch's current stream
    D, idx, val, cnt, rmse = eng.ExecuteIterations(X, D0, iternum=10)
"""
from ._lib import KsvdError, load  # noqa: F401
from .engine import Engine, version  # noqa: F401
