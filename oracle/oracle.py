"""ctypes front end of the C oracle (oracle/ksvd_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for every GPU result, and the FP32 port
timed as the CPU baseline.  The product never imports this module.

Array convention (numpy, C order) -- each is the reference's column-major
matrix seen row-by-column, so memory is bit-for-bit the reference layout:
    X   (N, n)   == X n x N col-major        (row s = signal s)
    D   (K, n)   == D n x K col-major        (row k = atom k)
    G   (K, K)
    DtX (N, K)   == DtX K x N col-major
    sel (N, K)   == SELECTED_ATOMS K x N
    c   (N, T)   == c T x N
    idx (N, T) int32, val (N, T), cnt (N,) int32   -- compact Gamma
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")


def build() -> None:
    """Compile both oracle libraries (gcc via oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_P = C.c_void_p
_I = C.c_int
_D = C.c_double


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One precision of the oracle: 'f64' (the checker) or 'f32' (CPU port)."""

    def __init__(self, prec: str = "f64"):
        assert prec in ("f64", "f32")
        self.prec = prec
        self.dtype = np.float64 if prec == "f64" else np.float32
        path = os.path.join(_BUILD, f"liboracle_{prec}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        sfx = "_" + prec
        L = self.lib

        def fn(name, res, *args):
            f = getattr(L, name + sfx)
            f.restype = res
            f.argtypes = list(args)
            return f

        self._threads = fn("or_set_threads", _I, _I)
        self._normcols = fn("or_normcols", None, _I, _I, _P, _P)
        self._atb = fn("or_at_times_b", None, _I, _I, _I, _P, _P, _P)
        self._atb_blk = fn("or_at_times_b_blocked", None, _I, _I, _I, _P, _P, _P)
        self._omp = fn("or_batch_omp", None, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P)
        self._perm = fn("or_perm_shuffle", None, _P, _I, _P)
        self._du_init = fn("or_du_init", None, _I, _I, _P, _P, _P, _P)
        self._du = fn("or_dict_update", None, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I)
        self._err = fn("or_error", _D, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P)
        self._clear = fn("or_clear_dict", None, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _D, _I, _P)
        self._ksvd = fn("or_ksvd", _I, _I, _I, _I, _I, _P, _P, _P, _I, C.c_uint64, _D, _I, _I, _P, _P, _P, _P, _P)
        self._lsfix = fn("or_ls_fixed", None, _I, _I, _I, _I, _P, _P, _P, _P, _P)
        self._lsadj = fn("or_ls_adjoint", None, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P)

    # ---- helpers -----------------------------------------------------------
    def _f(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    @staticmethod
    def _i(a):
        return np.ascontiguousarray(a, dtype=np.int32)

    def set_threads(self, t: int) -> int:
        return self._threads(int(t))

    # ---- primitives ------------------------------------------------------------
    def normcols(self, D):
        D = self._f(D)
        K, n = D.shape
        out = np.empty_like(D)
        self._normcols(n, K, _ptr(D), _ptr(out))
        return out

    def at_times_b(self, A, B):
        """C = A^T B with A (colsA, rows), B (colsB, rows) -> C (colsB, colsA)."""
        A = self._f(A)
        B = self._f(B)
        colsA, rows = A.shape
        colsB, rows2 = B.shape
        assert rows == rows2
        Cm = np.empty((colsB, colsA), dtype=self.dtype)
        self._atb(rows, colsA, colsB, _ptr(A), _ptr(B), _ptr(Cm))
        return Cm

    def at_times_b_blocked(self, A, B):
        """The same product by the cache-blocked loop (the CPU baseline's GEMM)."""
        A = self._f(A)
        B = self._f(B)
        colsA, rows = A.shape
        colsB, rows2 = B.shape
        assert rows == rows2
        Cm = np.empty((colsB, colsA), dtype=self.dtype)
        self._atb_blk(rows, colsA, colsB, _ptr(A), _ptr(B), _ptr(Cm))
        return Cm

    def gram(self, D):
        return self.at_times_b(D, D)

    def correlate(self, D, X):
        return self.at_times_b(D, X)

    def batch_omp(self, DtX, G, T, *, dense=False, c_init=None, xnorm2=None, eps2=None, gaps=False):
        """Batch-OMP.  Returns dict with idx, val, cnt, pending and optionally
        sel (slot map), c (coefficients), gap (T x N tie diagnostic)."""
        DtX = self._f(DtX)
        G = self._f(G)
        N, K = DtX.shape
        out = {
            "idx": np.empty((N, T), np.int32),
            "val": np.empty((N, T), self.dtype),
            "cnt": np.empty((N,), np.int32),
            "pending": np.empty((N,), np.int32),
        }
        sel = c = gap = None
        if dense:
            sel = np.empty((N, K), self.dtype)
            c = self._f(c_init).copy() if c_init is not None else np.zeros((N, T), self.dtype)
            out["sel"], out["c"] = sel, c
        if gaps:
            gap = np.empty((N, T), np.float64)
            out["gap"] = gap
        xn = self._f(xnorm2) if xnorm2 is not None else None
        e2 = self._f(eps2) if eps2 is not None else None
        self._omp(K, N, T, _ptr(DtX), _ptr(G), _ptr(sel), _ptr(c), _ptr(out["idx"]), _ptr(out["val"]),
                  _ptr(out["cnt"]), _ptr(xn), _ptr(e2), _ptr(gap), _ptr(out["pending"]))
        return out

    def perm_shuffle(self, buf, state):
        """In-place SplitMix64 Fisher-Yates of buf (uint32); state is a 1-elem uint64 array."""
        assert buf.dtype == np.uint32 and state.dtype == np.uint64
        self._perm(_ptr(buf), buf.size, _ptr(state))
        return buf

    def perms(self, K, iters, seed):
        buf = np.arange(K, dtype=np.uint32)
        st = np.array([seed], dtype=np.uint64)
        out = []
        for _ in range(iters):
            self.perm_shuffle(buf, st)
            out.append(buf.copy())
        return np.stack(out)

    def du_state(self, N, K):
        st = {
            "unused": np.empty(N, np.int32),
            "usc": np.zeros(1, np.int32),
            "bitmap": np.empty(N, np.int32),
            "replaced": np.empty(K, np.int32),
        }
        self._du_init(N, K, _ptr(st["unused"]), _ptr(st["usc"]), _ptr(st["bitmap"]), _ptr(st["replaced"]))
        return st

    def dict_update(self, X, D, idx, val, cnt, perm, state=None, max_signals=5000):
        """One approximate-K-SVD sweep.  Returns (D_new, val_new, state)."""
        X = self._f(X)
        D = self._f(D).copy()
        val = self._f(val).copy()
        idx = self._i(idx)
        cnt = self._i(cnt)
        N, n = X.shape
        K = D.shape[0]
        T = idx.shape[1]
        perm = np.ascontiguousarray(perm, dtype=np.uint32)
        if state is None:
            state = self.du_state(N, K)
        self._du(n, K, N, T, _ptr(X), _ptr(D), _ptr(idx), _ptr(val), _ptr(cnt), _ptr(perm),
                 _ptr(state["unused"]), _ptr(state["usc"]), _ptr(state["bitmap"]), _ptr(state["replaced"]),
                 int(max_signals))
        return D, val, state

    def error(self, X, D, idx, val, cnt, bitmap=None):
        X = self._f(X)
        D = self._f(D)
        N, n = X.shape
        K = D.shape[0]
        T = idx.shape[1]
        err = np.empty(N, self.dtype)
        bm = self._i(bitmap) if bitmap is not None else None
        rmse = self._err(n, K, N, T, _ptr(X), _ptr(D), _ptr(self._i(idx)), _ptr(self._f(val)),
                         _ptr(self._i(cnt)), _ptr(bm), _ptr(err))
        return rmse, err

    def clear_dict(self, X, D, idx, val, cnt, replaced, err, mu=0.8, use_thresh=4):
        X = self._f(X)
        D = self._f(D).copy()
        err = self._f(err).copy()
        N, n = X.shape
        K = D.shape[0]
        T = idx.shape[1]
        flags = np.empty(K, np.int32)
        self._clear(n, K, N, T, _ptr(X), _ptr(D), _ptr(self._i(idx)), _ptr(self._f(val)), _ptr(self._i(cnt)),
                    _ptr(self._i(replaced)), _ptr(err), float(mu), int(use_thresh), _ptr(flags))
        return D, err, flags

    def ksvd(self, X, Din, T, iters, seed=0, mu=0.8, use_thresh=4, max_signals=5000, eps2=None):
        """K-SVD iterations; eps2 (N,) switches Batch-OMP to the error-constrained stop."""
        X = self._f(X)
        e2 = self._f(eps2) if eps2 is not None else None
        Din = self._f(Din)
        N, n = X.shape
        K = Din.shape[0]
        Dout = np.empty_like(Din)
        idx = np.empty((N, T), np.int32)
        val = np.empty((N, T), self.dtype)
        cnt = np.empty(N, np.int32)
        rmse = np.empty(iters, np.float64)
        self._ksvd(n, K, N, T, _ptr(X), _ptr(Din), _ptr(Dout), iters, int(seed), float(mu), int(use_thresh),
                   int(max_signals), _ptr(idx), _ptr(val), _ptr(cnt), _ptr(rmse), _ptr(e2))
        return {"D": Dout, "idx": idx, "val": val, "cnt": cnt, "rmse": rmse}

    def ls_fixed(self, D, X, idx, cnt):
        D = self._f(D)
        X = self._f(X)
        N, n = X.shape
        K = D.shape[0]
        T = idx.shape[1]
        val = np.empty((N, T), self.dtype)
        self._lsfix(n, K, N, T, _ptr(D), _ptr(X), _ptr(self._i(idx)), _ptr(self._i(cnt)), _ptr(val))
        return val

    def ls_adjoint(self, D, X, idx, cnt, val, gval):
        D = self._f(D)
        X = self._f(X)
        N, n = X.shape
        K = D.shape[0]
        T = idx.shape[1]
        gD = np.empty((K, n), self.dtype)
        gX = np.empty((N, n), self.dtype)
        self._lsadj(n, K, N, T, _ptr(D), _ptr(X), _ptr(self._i(idx)), _ptr(self._i(cnt)), _ptr(self._f(val)),
                    _ptr(self._f(gval)), _ptr(gD), _ptr(gX))
        return gD, gX


_CACHE: dict = {}


def load(prec: str = "f64") -> Oracle:
    if prec not in _CACHE:
        _CACHE[prec] = Oracle(prec)
    return _CACHE[prec]
