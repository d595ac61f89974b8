"""ctypes binding of libksvd_amd.so (the C-ABI in include/ksvd_amd.h).

Loading fails loudly when the library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libksvd_amd.so")

FORM_FP32, FORM_BF16, FORM_BF16X3, FORM_BF16X6 = 0, 1, 2, 3
FORMS = {"fp32": FORM_FP32, "bf16": FORM_BF16, "bf16x3": FORM_BF16X3, "bf16x6": FORM_BF16X6}
OMP_AUTO, OMP_WAVE, OMP_QFORM = 0, 1, 4
OMP_VARIANTS = {"auto": OMP_AUTO, "wave": OMP_WAVE, "qform": OMP_QFORM}

KSVD_OK = 0
STATUS = {0: "KSVD_OK", -1: "KSVD_ERR_INVALID", -2: "KSVD_ERR_LIMIT", -3: "KSVD_ERR_HIP", -4: "KSVD_ERR_ALLOC",
          -5: "KSVD_ERR_NULL"}

# Every entry point include/ksvd_amd.h declares (checked by the CPU tests).
EXPORTS = [
    "ksvd_config_default", "ksvd_create", "ksvd_destroy", "ksvd_set_stream", "ksvd_last_error", "ksvd_version",
    "ksvd_sync", "ksvd_normcols", "ksvd_at_times_b", "ksvd_gram", "ksvd_correlate", "ksvd_batch_omp",
    "ksvd_batch_omp_compact", "ksvd_batch_omp_from_signals", "ksvd_batch_omp_from_signals_dense",
    "ksvd_batch_omp_eps", "ksvd_batch_omp_eps_from_signals", "ksvd_batch_omp_from_signals_noalloc", "ksvd_omp_stats", "ksvd_du_init",
    "ksvd_dictionary_update", "ksvd_dictionary_update_dense", "ksvd_error", "ksvd_clear_dict",
    "ksvd_execute_iterations", "ksvd_perm_sequence", "ksvd_omp_adjoint", "ksvd_mat_read", "ksvd_mat_write",
    "ksvd_batch_omp_compact_factors", "ksvd_omp_adjoint_factored", "ksvd_normcols_adjoint", "ksvd_compact_to_dense",
    "ksvd_compact_gather", "ksvd_execute_iterations_eps", "ksvd_get_tuning", "ksvd_set_tuning", "ksvd_autotune",
    "ksvd_tuning_log",
    "ksvd_mat_last_error",
]


class Config(C.Structure):
    """Mirror of struct ksvd_config."""
    _fields_ = [
        ("n", C.c_int32), ("K", C.c_int32), ("N", C.c_int32), ("T", C.c_int32),
        ("iterations", C.c_int32), ("use_thresh", C.c_int32), ("max_signals", C.c_int32),
        ("gemm_form", C.c_int32), ("omp_variant", C.c_int32), ("use_graph", C.c_int32),
        ("mu_thresh", C.c_double), ("pivot_fp64", C.c_double), ("perm_seed", C.c_uint64),
    ]


class Tuning(C.Structure):
    """Mirror of struct ksvd_tuning."""
    _fields_ = [
        ("omp_kpl", C.c_int32), ("omp_waves", C.c_int32), ("omp_eps_launch", C.c_int32), ("gemm_wpe", C.c_int32),
        ("du_blocks", C.c_int32), ("source", C.c_int32),
        ("omp_ms", C.c_double), ("gemm_ms", C.c_double), ("du_ms", C.c_double),
        ("default_omp_ms", C.c_double), ("default_gemm_ms", C.c_double), ("default_du_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KsvdError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KsvdError(f"{LIB_PATH} not found: build it with `python -m ksvd_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P, I, H = C.c_void_p, C.c_int, C.c_void_p

    def sig(name, res, *args):
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = list(args)

    sig("ksvd_config_default", None, C.POINTER(Config))
    sig("ksvd_create", I, C.POINTER(H), C.POINTER(Config))
    sig("ksvd_destroy", I, H)
    sig("ksvd_set_stream", I, H, P)
    sig("ksvd_last_error", C.c_char_p, H)
    sig("ksvd_version", C.c_char_p)
    sig("ksvd_sync", I, H)
    sig("ksvd_normcols", I, H, P, P)
    sig("ksvd_at_times_b", I, H, P, I, I, P, I, P)
    sig("ksvd_gram", I, H, P, P)
    sig("ksvd_correlate", I, H, P, P, P)
    sig("ksvd_batch_omp", I, H, P, P, P, P)
    sig("ksvd_batch_omp_compact", I, H, P, P, P, P, P)
    sig("ksvd_batch_omp_from_signals", I, H, P, P, P, P, P, P)
    sig("ksvd_batch_omp_from_signals_dense", I, H, P, P, P, P, P)
    sig("ksvd_batch_omp_eps", I, H, P, P, P, P, P, P, P)
    sig("ksvd_batch_omp_eps_from_signals", I, H, P, P, P, P, P, P, P)
    sig("ksvd_batch_omp_from_signals_noalloc", I, H, P, P, P, I, P, P, P)
    sig("ksvd_omp_stats", I, H, C.POINTER(C.c_int64), C.POINTER(C.c_int64))
    sig("ksvd_du_init", I, H)
    sig("ksvd_dictionary_update", I, H, P, P, P, P, P, P)
    sig("ksvd_dictionary_update_dense", I, H, P, P, P, P, P)
    sig("ksvd_error", I, H, P, P, P, P, P, P, P)
    sig("ksvd_clear_dict", I, H, P, P, P, P, P, P)
    sig("ksvd_execute_iterations", I, H, P, P, P, P, P, P, I, P)
    sig("ksvd_execute_iterations_eps", I, H, P, P, P, P, P, P, P, I, P)
    sig("ksvd_get_tuning", I, H, C.POINTER(Tuning))
    sig("ksvd_set_tuning", I, H, C.POINTER(Tuning))
    sig("ksvd_autotune", I, H, P, P, P, I)
    sig("ksvd_tuning_log", C.c_char_p, H)
    sig("ksvd_perm_sequence", I, C.c_uint64, I, I, P)
    sig("ksvd_omp_adjoint", I, H, P, P, P, P, P, P, P, P)
    sig("ksvd_batch_omp_compact_factors", I, H, P, P, P, P, P, P)
    sig("ksvd_omp_adjoint_factored", I, H, P, P, P, P, P, P, P, P, P)
    sig("ksvd_normcols_adjoint", I, H, P, P, P)
    sig("ksvd_compact_to_dense", I, H, P, P, P, P)
    sig("ksvd_compact_gather", I, H, P, P, P, P)
    sig("ksvd_mat_read", I, C.c_char_p, C.c_char_p, P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64))
    sig("ksvd_mat_write", I, C.c_char_p, C.c_char_p, P, C.c_int64, C.c_int64, I, I)
    sig("ksvd_mat_last_error", C.c_char_p)
    _lib = lib
    return lib


def perm_sequence(seed: int, K: int, iterations: int):
    """Host-only permutation stream of ksvd_execute_iterations (iterations x K)."""
    import numpy as np
    out = np.empty((iterations, K), dtype=np.uint32)
    rc = load().ksvd_perm_sequence(C.c_uint64(seed), K, iterations, out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise KsvdError(STATUS.get(rc, rc))
    return out


def default_config(**kw) -> Config:
    cfg = Config()
    load().ksvd_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k == "gemm_form" and isinstance(v, str):
            v = FORMS[v]
        if k == "omp_variant" and isinstance(v, str):
            v = OMP_VARIANTS[v]
        setattr(cfg, k, v)
    return cfg


def mat_read(path: str, name: str):
    """Variable `name` of a level-5 MAT-file as a (rows, cols) float64 array
    (host only; ksvd_mat_read)."""
    import numpy as np
    lib = load()
    r, c = C.c_int64(), C.c_int64()
    rc = lib.ksvd_mat_read(path.encode(), name.encode(), None, 0, C.byref(r), C.byref(c))
    if rc != 0:
        raise KsvdError(lib.ksvd_mat_last_error().decode())
    out = np.empty((c.value, r.value), dtype=np.float64)  # column-major storage, transposed view below
    rc = lib.ksvd_mat_read(path.encode(), name.encode(), out.ctypes.data_as(C.c_void_p), out.size, C.byref(r),
                           C.byref(c))
    if rc != 0:
        raise KsvdError(lib.ksvd_mat_last_error().decode())
    return out.T


def mat_write(path: str, variables: dict, global_vars: bool = False):
    """Writes {name: 2-D array} as an uncompressed level-5 MAT-file of doubles
    (host only; ksvd_mat_write)."""
    import numpy as np
    lib = load()
    for i, (name, a) in enumerate(variables.items()):
        a = np.asarray(a, dtype=np.float64)
        if a.ndim != 2:
            raise ValueError(f"{name}: MAT variables here are 2-D")
        col = np.ascontiguousarray(a.T)  # column-major bytes
        rc = lib.ksvd_mat_write(path.encode(), name.encode(), col.ctypes.data_as(C.c_void_p), a.shape[0], a.shape[1],
                                1 if i else 0, 1 if global_vars else 0)
        if rc != 0:
            raise KsvdError(lib.ksvd_mat_last_error().decode())
