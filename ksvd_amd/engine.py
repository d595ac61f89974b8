"""Host-side mirror of the reference's K-SVD interface over the C-ABI.

The reference exposes C++ classes over a global device arena:
    KSVD::ExecuteIterations(iternum), KSVD::isValid()       S/KSVD.cuh:11-29
    OMP::BatchOMP()                                         S/OMP.cuh:12-25
    DictionaryUpdate::updateDictionary(unsigned int* p)     S/DictionaryUpdate.cuh:11-25
    Algorithms::Multiplication::AT_times_B(...)             S/Algorithms.cuh:99
`Engine` keeps those names (and their bool-on-success convention where the
reference has one) while every buffer is a caller-owned torch CUDA tensor in
the reference's column-major layout, seen row-per-column:
    X (N, n), D (K, n), G (K, K), DtX (N, K), SELECTED_ATOMS (N, K), c (N, T),
    compact Gamma: idx (N, T) int32, val (N, T) float32, cnt (N,) int32.
Torch is used for device memory and streams only; all compute is the HIP
library.  Errors raise KsvdError with the library's message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


def _p(t):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if not t.is_cuda:
        raise ValueError("engine buffers must be CUDA (HIP) device tensors")
    if not t.is_contiguous():
        raise ValueError("engine buffers must be contiguous")
    return C.c_void_p(t.data_ptr())


def _check_shape(t, shape, dtype, name):
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype:
        raise ValueError(f"{name}: expected {tuple(shape)} {dtype}, got {tuple(t.shape)} {t.dtype}")


class Engine:
    """One engine handle (ksvd_create) bound to the current torch stream."""

    def __init__(self, n, K, N, T, *, iterations=30, gemm_form="fp32", omp_variant="auto", mu_thresh=0.8,
                 use_thresh=4, max_signals=5000, pivot_fp64=1e-3, perm_seed=0, use_graph=True):
        self.lib = _lib.load()
        self.cfg = _lib.default_config(n=n, K=K, N=N, T=T, iterations=iterations, gemm_form=gemm_form,
                                       omp_variant=omp_variant, mu_thresh=mu_thresh, use_thresh=use_thresh,
                                       max_signals=max_signals, pivot_fp64=pivot_fp64, perm_seed=perm_seed,
                                       use_graph=int(bool(use_graph)))
        self.n, self.K, self.N, self.T = n, K, N, T
        h = C.c_void_p()
        rc = self.lib.ksvd_create(C.byref(h), C.byref(self.cfg))
        self.h = h
        if rc != 0:
            msg = self.lib.ksvd_last_error(h).decode() if h else "ksvd_create failed"
            if h:
                self.lib.ksvd_destroy(h)
                self.h = None
            raise _lib.KsvdError(f"{_lib.STATUS.get(rc, rc)}: {msg}")
        self.set_stream(torch.cuda.current_stream())
        self.valid = True

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.ksvd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream: torch.cuda.Stream):
        self._rc(self.lib.ksvd_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def isValid(self) -> bool:  # KSVD::isValid  S/KSVD.cu:291-293
        return bool(self.h) and self.valid

    def _rc(self, rc):
        if rc != 0:
            self.valid = False
            raise _lib.KsvdError(f"{_lib.STATUS.get(rc, rc)}: {self.lib.ksvd_last_error(self.h).decode()}")

    def sync(self):
        self._rc(self.lib.ksvd_sync(self.h))

    # -- launch shapes ----------------------------------------------------------------
    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        self._rc(self.lib.ksvd_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        """Sets launch-shape fields (omp_kpl, omp_waves, omp_eps_launch,
        gemm_wpe, du_blocks); the others keep their current values."""
        t = _lib.Tuning()
        self._rc(self.lib.ksvd_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self._rc(self.lib.ksvd_set_tuning(self.h, C.byref(t)))
        return self.get_tuning()

    def autotune(self, D, X, eps2=None, force=False) -> dict:
        """ksvd_autotune: measures the candidate launch shapes on D and X and
        keeps the fastest; returns the chosen tuning with its timings."""
        self._rc(self.lib.ksvd_autotune(self.h, _p(D), _p(X), _p(eps2), int(bool(force))))
        t = self.get_tuning()
        t["log"] = self.lib.ksvd_tuning_log(self.h).decode()
        return t

    # -- allocation helpers -------------------------------------------------------
    def _dev(self):
        return torch.device("cuda", torch.cuda.current_device())

    def empty_compact(self):
        d = self._dev()
        return (torch.empty((self.N, self.T), dtype=torch.int32, device=d),
                torch.empty((self.N, self.T), dtype=torch.float32, device=d),
                torch.empty((self.N,), dtype=torch.int32, device=d))

    # -- primitives -----------------------------------------------------------------
    def normcols(self, D_in, D_out=None):
        D_out = torch.empty_like(D_in) if D_out is None else D_out
        _check_shape(D_in, (self.K, self.n), torch.float32, "D")
        self._rc(self.lib.ksvd_normcols(self.h, _p(D_in), _p(D_out)))
        return D_out

    def AT_times_B(self, A, B, out=None):
        """C = A^T B for column-major A (rows x colsA) and B (rows x colsB),
        given as tensors (colsA, rows) and (colsB, rows); returns (colsB, colsA)."""
        colsA, rows = A.shape
        colsB, rows2 = B.shape
        if rows != rows2:
            raise ValueError("AT_times_B: row mismatch")
        out = torch.empty((colsB, colsA), dtype=torch.float32, device=A.device) if out is None else out
        self._rc(self.lib.ksvd_at_times_b(self.h, _p(A), rows, colsA, _p(B), colsB, _p(out)))
        return out

    def gram(self, D, G=None):
        G = torch.empty((self.K, self.K), dtype=torch.float32, device=D.device) if G is None else G
        self._rc(self.lib.ksvd_gram(self.h, _p(D), _p(G)))
        return G

    def correlate(self, D, X, DtX=None):
        _check_shape(X, (self.N, self.n), torch.float32, "X")
        DtX = torch.empty((self.N, self.K), dtype=torch.float32, device=X.device) if DtX is None else DtX
        self._rc(self.lib.ksvd_correlate(self.h, _p(D), _p(X), _p(DtX)))
        return DtX

    # -- Batch-OMP ------------------------------------------------------------------
    def batch_omp(self, DtX, G, selected_atoms=None, c=None):
        """Reference layout: returns (SELECTED_ATOMS (N, K) slot map, c (N, T))."""
        _check_shape(DtX, (self.N, self.K), torch.float32, "DtX")
        if selected_atoms is None:
            selected_atoms = torch.empty((self.N, self.K), dtype=torch.float32, device=DtX.device)
        if c is None:
            c = torch.zeros((self.N, self.T), dtype=torch.float32, device=DtX.device)
        self._rc(self.lib.ksvd_batch_omp(self.h, _p(DtX), _p(G), _p(selected_atoms), _p(c)))
        return selected_atoms, c

    def batch_omp_compact(self, DtX, G, out=None):
        _check_shape(DtX, (self.N, self.K), torch.float32, "DtX")
        idx, val, cnt = self.empty_compact() if out is None else out
        self._rc(self.lib.ksvd_batch_omp_compact(self.h, _p(DtX), _p(G), _p(idx), _p(val), _p(cnt)))
        return idx, val, cnt

    def BatchOMP(self, D, X, G, out=None):
        """OMP::BatchOMP (S/OMP.cu:42-59): DtX = D^T X, then Batch-OMP."""
        idx, val, cnt = self.empty_compact() if out is None else out
        self._rc(self.lib.ksvd_batch_omp_from_signals(self.h, _p(D), _p(X), _p(G), _p(idx), _p(val), _p(cnt)))
        return idx, val, cnt

    def BatchOMP_dense(self, D, X, G, selected_atoms=None, c=None):
        """OMP::BatchOMP with the reference layout (SELECTED_ATOMS slot map, c)."""
        if selected_atoms is None:
            selected_atoms = torch.empty((self.N, self.K), dtype=torch.float32, device=X.device)
        if c is None:
            c = torch.zeros((self.N, self.T), dtype=torch.float32, device=X.device)
        self._rc(self.lib.ksvd_batch_omp_from_signals_dense(self.h, _p(D), _p(X), _p(G), _p(selected_atoms), _p(c)))
        return selected_atoms, c

    def BatchOMP_eps(self, D, X, G, eps2, out=None):
        """Error-constrained Batch-OMP from signals (configured variant)."""
        idx, val, cnt = self.empty_compact() if out is None else out
        self._rc(self.lib.ksvd_batch_omp_eps_from_signals(self.h, _p(D), _p(X), _p(G), _p(eps2), _p(idx), _p(val),
                                                          _p(cnt)))
        return idx, val, cnt

    def BatchOMP_noalloc(self, D, X, G, N=None, out=None):
        """Allocation-free Batch-OMP on the first N signals of X
        (ksvd_batch_omp_from_signals_noalloc): the engine allocates nothing;
        outputs are the caller's (N, T) / (N,) tensors."""
        N = self.N if N is None else int(N)
        if out is None:
            d = X.device
            out = (torch.empty((N, self.T), dtype=torch.int32, device=d),
                   torch.empty((N, self.T), dtype=torch.float32, device=d),
                   torch.empty((N,), dtype=torch.int32, device=d))
        idx, val, cnt = out
        if X.shape[0] < N or idx.shape[0] < N or val.shape[0] < N or cnt.shape[0] < N:
            raise ValueError("BatchOMP_noalloc: X and the outputs need at least N rows")
        self._rc(self.lib.ksvd_batch_omp_from_signals_noalloc(self.h, _p(D), _p(X), _p(G), N, _p(idx), _p(val),
                                                              _p(cnt)))
        return idx, val, cnt

    def omp_stats(self):
        """(candidates re-evaluated in FP32, fallback steps) of the last Batch-OMP call."""
        c, f = C.c_int64(), C.c_int64()
        self._rc(self.lib.ksvd_omp_stats(self.h, C.byref(c), C.byref(f)))
        return c.value, f.value

    def omp_fallbacks(self) -> int:
        return self.omp_stats()[1]

    def batch_omp_eps(self, DtX, G, xnorm2, eps2, out=None):
        """Error-constrained Batch-OMP from DtX; xnorm2 (N,) float64 = ||x_s||^2."""
        _check_shape(xnorm2, (self.N,), torch.float64, "xnorm2")
        idx, val, cnt = self.empty_compact() if out is None else out
        self._rc(self.lib.ksvd_batch_omp_eps(self.h, _p(DtX), _p(G), _p(xnorm2), _p(eps2), _p(idx), _p(val),
                                             _p(cnt)))
        return idx, val, cnt

    # -- dictionary update / error / clear ----------------------------------------------
    def du_init(self):
        self._rc(self.lib.ksvd_du_init(self.h))

    def updateDictionary(self, X, D, idx, val, cnt, p):
        """DictionaryUpdate::updateDictionary(p): p is a host permutation of 0..K-1.
        D and val are updated in place."""
        perm = np.ascontiguousarray(np.asarray(p, dtype=np.uint32))
        if perm.shape != (self.K,):
            raise ValueError("perm must have K entries")
        self._rc(self.lib.ksvd_dictionary_update(self.h, _p(X), _p(D), _p(idx), _p(val), _p(cnt),
                                                 perm.ctypes.data_as(C.c_void_p)))
        return True

    def updateDictionary_dense(self, X, D, selected_atoms, c, p):
        perm = np.ascontiguousarray(np.asarray(p, dtype=np.uint32))
        self._rc(self.lib.ksvd_dictionary_update_dense(self.h, _p(X), _p(D), _p(selected_atoms), _p(c),
                                                       perm.ctypes.data_as(C.c_void_p)))
        return True

    def errorComputation(self, X, D, idx, val, cnt, err=None):
        err = torch.empty((self.N,), dtype=torch.float32, device=X.device) if err is None else err
        rmse = C.c_double()
        self._rc(self.lib.ksvd_error(self.h, _p(X), _p(D), _p(idx), _p(val), _p(cnt), _p(err), C.byref(rmse)))
        return rmse.value, err

    def clearDict(self, X, D, idx, val, cnt, err):
        self._rc(self.lib.ksvd_clear_dict(self.h, _p(X), _p(D), _p(idx), _p(val), _p(cnt), _p(err)))
        return True

    def ExecuteIterations(self, X, D_in, iternum, D_out=None, out=None):
        """KSVD::ExecuteIterations (S/KSVD.cu:55-131).  Returns (D, idx, val, cnt, rmse)."""
        D_out = torch.empty_like(D_in) if D_out is None else D_out
        idx, val, cnt = self.empty_compact() if out is None else out
        rmse = (C.c_double * max(1, iternum))()
        self._rc(self.lib.ksvd_execute_iterations(self.h, _p(X), _p(D_in), _p(D_out), _p(idx), _p(val), _p(cnt),
                                                  int(iternum), rmse))
        return D_out, idx, val, cnt, list(rmse)[:iternum]

    def ExecuteIterations_eps(self, X, D_in, eps2, iternum, D_out=None, out=None):
        """KSVD::ExecuteIterations with error-constrained Batch-OMP (stop at
        ||r_s||^2 <= eps2[s] or T atoms).  Returns (D, idx, val, cnt, rmse)."""
        _check_shape(eps2, (self.N,), torch.float32, "eps2")
        D_out = torch.empty_like(D_in) if D_out is None else D_out
        idx, val, cnt = self.empty_compact() if out is None else out
        rmse = (C.c_double * max(1, iternum))()
        self._rc(self.lib.ksvd_execute_iterations_eps(self.h, _p(X), _p(D_in), _p(D_out), _p(eps2), _p(idx), _p(val),
                                                      _p(cnt), int(iternum), rmse))
        return D_out, idx, val, cnt, list(rmse)[:iternum]

    # -- differentiation ------------------------------------------------------------------
    def omp_adjoint(self, D, X, idx, val, cnt, gval, gD=None, gX=None, Lfac=None):
        """Fixed-support adjoint: (gD, gX).  With Lfac (from
        batch_omp_compact_factors) the stored Cholesky factors are reused,
        otherwise G_II is recomputed per signal."""
        gD = torch.empty_like(D) if gD is None else gD
        gX = torch.empty_like(X) if gX is None else gX
        if Lfac is None:
            self._rc(self.lib.ksvd_omp_adjoint(self.h, _p(D), _p(X), _p(idx), _p(val), _p(cnt), _p(gval), _p(gD),
                                               _p(gX)))
        else:
            _check_shape(Lfac, (self.N, self.T * (self.T + 1) // 2), torch.float32, "Lfac")
            self._rc(self.lib.ksvd_omp_adjoint_factored(self.h, _p(D), _p(X), _p(idx), _p(val), _p(cnt), _p(Lfac),
                                                        _p(gval), _p(gD), _p(gX)))
        return gD, gX

    def batch_omp_compact_factors(self, DtX, G, out=None, Lfac=None):
        """Batch-OMP (compact output) that also returns each signal's packed
        Cholesky factor (N, T(T+1)/2)."""
        _check_shape(DtX, (self.N, self.K), torch.float32, "DtX")
        idx, val, cnt = self.empty_compact() if out is None else out
        if Lfac is None:
            Lfac = torch.empty((self.N, self.T * (self.T + 1) // 2), dtype=torch.float32, device=DtX.device)
        self._rc(self.lib.ksvd_batch_omp_compact_factors(self.h, _p(DtX), _p(G), _p(idx), _p(val), _p(cnt),
                                                         _p(Lfac)))
        return idx, val, cnt, Lfac

    def normcols_adjoint(self, D_in, gD_out, gD_in=None):
        gD_in = torch.empty_like(D_in) if gD_in is None else gD_in
        self._rc(self.lib.ksvd_normcols_adjoint(self.h, _p(D_in), _p(gD_out), _p(gD_in)))
        return gD_in

    def compact_to_dense(self, idx, val, cnt, out=None):
        """Dense Gamma (N, K) (the reference's K x N column-major) from the compact form."""
        out = torch.empty((self.N, self.K), dtype=torch.float32, device=idx.device) if out is None else out
        self._rc(self.lib.ksvd_compact_to_dense(self.h, _p(idx), _p(val), _p(cnt), _p(out)))
        return out

    def compact_gather(self, dense, idx, cnt, out=None):
        """out (N, T) = dense[s, idx[s, j]] for j < cnt[s], else 0."""
        _check_shape(dense, (self.N, self.K), torch.float32, "dense")
        out = torch.empty((self.N, self.T), dtype=torch.float32, device=dense.device) if out is None else out
        self._rc(self.lib.ksvd_compact_gather(self.h, _p(dense), _p(idx), _p(cnt), _p(out)))
        return out


def version() -> str:
    return _lib.load().ksvd_version().decode()
