"""CPU-only checks of the C-ABI library: it loads, exports exactly what
include/ksvd_amd.h declares, and its host-only pieces behave (no device work)."""
import ctypes as C
import json
import os
import re

import numpy as np

from ksvd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "ksvd_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ksvd_\w+)\s*\(", src)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    nm = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    for name in _declared():
        assert re.search(rf"\bT {name}\b", nm), name


def test_config_defaults_match_reference_parameters():
    cfg = _lib.default_config()
    # S/ProgramParameters.cuh:17-23, S/GlobalDeclarations.cuh:54,63
    assert (cfg.T, cfg.iterations, cfg.use_thresh, cfg.max_signals) == (6, 30, 4, 5000)
    assert cfg.mu_thresh == 0.8 and cfg.gemm_form == 0
    assert C.sizeof(_lib.Config) == 64


def test_version_string():
    assert b"gfx950" in _lib.load().ksvd_version()


def test_perm_sequence_matches_golden_and_oracle(o64):
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "perm_golden.json")))
    for st in g["streams"]:
        got = _lib.perm_sequence(st["seed"], st["K"], st["iters"])
        np.testing.assert_array_equal(got, np.array(st["perms"], dtype=np.uint32))
    np.testing.assert_array_equal(_lib.perm_sequence(77, 1024, 3), o64.perms(1024, 3, 77))


def test_null_and_invalid_arguments_are_reported():
    lib = _lib.load()
    assert lib.ksvd_create(None, None) == -5
    out = np.empty((1, 4), np.uint32)
    assert lib.ksvd_perm_sequence(C.c_uint64(1), 0, 1, out.ctypes.data_as(C.c_void_p)) == -1
    assert lib.ksvd_perm_sequence(C.c_uint64(1), 4, 1, None) == -5
    # a config beyond the kernels' limits is rejected before any device call
    h = C.c_void_p()
    cfg = _lib.default_config(n=64, K=256, N=10, T=65)
    assert lib.ksvd_create(C.byref(h), C.byref(cfg)) == -2
    assert b"T=65" in lib.ksvd_last_error(h)
    lib.ksvd_destroy(h)
    cfg = _lib.default_config(n=64, K=4096, N=10, T=8)
    h = C.c_void_p()
    assert lib.ksvd_create(C.byref(h), C.byref(cfg)) == -2
    lib.ksvd_destroy(h)


def test_register_basis_variant_limits():
    """KSVD_OMP_QFORM holds T x K FP32 values per signal in at most 8 waves'
    VGPRs: T > 32 with K > 1024 is rejected before any device call."""
    lib = _lib.load()
    h = C.c_void_p()
    cfg = _lib.default_config(n=64, K=2048, N=10, T=40, omp_variant="qform")
    assert lib.ksvd_create(C.byref(h), C.byref(cfg)) == -2
    assert b"K=2048 T=40" in lib.ksvd_last_error(h)
    lib.ksvd_destroy(h)
    cfg = _lib.default_config(n=64, K=99, N=10, T=5, omp_variant=5)
    h = C.c_void_p()
    assert lib.ksvd_create(C.byref(h), C.byref(cfg)) == -1
    lib.ksvd_destroy(h)
