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


def test_cpp_facade_compiles_links_and_checks_arguments(tmp_path):
    """include/ksvd_amd.hpp (the reference's class names over the C-ABI) builds
    with a plain host compiler (no HIP headers) against the library; its
    host-side argument checks run without a GPU."""
    import subprocess
    src = tmp_path / "facade.cpp"
    src.write_text(r'''
#include "ksvd_amd.hpp"
#include <type_traits>
int main() {
    ksvd_amd::DeviceMemory m;
    if (m.X || m.D || m.G || m.idx || m.err) return 1;
    // AT_times_B rejects operands of different heights before touching the device
    float a = 0.f;
    if (ksvd_amd::Algorithms::Multiplication::AT_times_B(nullptr, &a, 3, 1, &a, 4, 1, &a)) return 2;
    // the class layout: KSVD is an OMP and a DictionaryUpdate over one CudaAlgorithm
    static_assert(std::is_base_of<ksvd_amd::OMP, ksvd_amd::KSVD>::value, "KSVD : OMP");
    static_assert(std::is_base_of<ksvd_amd::DictionaryUpdate, ksvd_amd::KSVD>::value, "KSVD : DictionaryUpdate");
    static_assert(std::is_base_of<ksvd_amd::CudaAlgorithm, ksvd_amd::KSVD>::value, "KSVD : CudaAlgorithm");
    // the permutations the phase-by-phase loop uses come from the host-only entry point
    uint32_t p[2 * 5];
    if (ksvd_perm_sequence(7, 5, 2, p) != KSVD_OK) return 3;
    int seen = 0;
    for (int k = 0; k < 5; ++k) seen |= 1 << p[k];
    return seen == 31 ? 0 : 4;
}
''')
    exe = tmp_path / "facade"
    lib_dir = os.path.join(ROOT, "ksvd_amd")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), f"-L{lib_dir}", "-lksvd_amd",
                        f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)
    assert "rowsA = 3 and rowsB = 4 differ" in r.stderr
