"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The oracle (oracle/ksvd_oracle.c, FP64 build, serial) driven through every
  entry point on small seeded problems, in sparsity and error mode.
* The MAT-file reader and writer (ksvd_amd/csrc/matio.cpp, the one part of the
  library that parses files a user hands it) on the reference's own MAT
  files, a write/read round trip, and truncated and byte-flipped copies of
  those files: every read either succeeds or returns an error, with no
  sanitizer report.
Both are built here with gcc/g++ from the committed sources (host code only;
the GPU kernels are not involved)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc and g++")

ORACLE_DRIVER = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
typedef double R;
void or_normcols_f64(int n, int K, const R* Din, R* Dout);
void or_at_times_b_f64(int rows, int colsA, int colsB, const R* A, const R* B, R* C);
void or_at_times_b_blocked_f64(int rows, int colsA, int colsB, const R* A, const R* B, R* C);
void or_batch_omp_f64(int K, int N, int T, const R* DtX, const R* G, R* selected, R* c, int32_t* idx, R* val,
                      int32_t* cnt, const R* xnorm2, const R* eps2, double* gap, int32_t* pending);
int or_ksvd_f64(int n, int K, int N, int T, const R* X, const R* Din, R* Dout, int iters, uint64_t seed,
                double mu_thresh, int use_thresh, int max_signals, int32_t* idx, R* val, int32_t* cnt, double* rmse,
                const R* eps2);
void or_ls_fixed_f64(int n, int K, int N, int T, const R* D, const R* X, const int32_t* idx, const int32_t* cnt,
                     R* val);
void or_ls_adjoint_f64(int n, int K, int N, int T, const R* D, const R* X, const int32_t* idx, const int32_t* cnt,
                       const R* val, const R* gval, R* gD, R* gX);

static uint64_t st = 0x9E3779B97F4A7C15ull;
static double gauss(void) {  /* Box-Muller on a 64-bit LCG */
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    double u = ((st >> 11) + 1.0) / 9007199254740994.0;
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    double v = (st >> 11) / 9007199254740992.0;
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}
static void* xmalloc(size_t b) { void* p = malloc(b ? b : 1); if (!p) exit(3); return p; }

static int run(int n, int K, int N, int T, int eps) {
    R *D0 = xmalloc(sizeof(R) * n * K), *D = xmalloc(sizeof(R) * n * K), *X = xmalloc(sizeof(R) * n * N);
    for (int i = 0; i < n * K; ++i) D0[i] = gauss();
    for (int i = 0; i < n * N; ++i) X[i] = gauss();
    for (int t = 0; t < n; ++t) X[t] = 0.0;                 /* a zero signal */
    for (int t = 0; t < n; ++t) X[n + t] = D0[t];           /* a signal on atom 0 (before normalisation) */
    or_normcols_f64(n, K, D0, D);
    R *G = xmalloc(sizeof(R) * K * K), *A0 = xmalloc(sizeof(R) * K * N), *A1 = xmalloc(sizeof(R) * K * N);
    or_at_times_b_f64(n, K, K, D, D, G);
    or_at_times_b_f64(n, K, N, D, X, A0);
    or_at_times_b_blocked_f64(n, K, N, D, X, A1);
    R *sel = xmalloc(sizeof(R) * K * N), *c = xmalloc(sizeof(R) * T * N), *val = xmalloc(sizeof(R) * T * N);
    int32_t *idx = xmalloc(sizeof(int32_t) * T * N), *cnt = xmalloc(sizeof(int32_t) * N);
    double* gap = xmalloc(sizeof(double) * T * N);
    int32_t* pending = xmalloc(sizeof(int32_t) * N);
    R *xn2 = NULL, *e2 = NULL;
    if (eps) {
        xn2 = xmalloc(sizeof(R) * N);
        e2 = xmalloc(sizeof(R) * N);
        for (int s = 0; s < N; ++s) {
            double a = 0;
            for (int t = 0; t < n; ++t) a += X[s * n + t] * X[s * n + t];
            xn2[s] = a;
            e2[s] = 0.05 * a;
        }
    }
    memset(sel, 0, sizeof(R) * K * N);
    memset(c, 0, sizeof(R) * T * N);
    or_batch_omp_f64(K, N, T, A0, G, sel, c, idx, val, cnt, xn2, e2, gap, pending);
    or_ls_fixed_f64(n, K, N, T, D, X, idx, cnt, val);
    R *gval = xmalloc(sizeof(R) * T * N), *gD = xmalloc(sizeof(R) * n * K), *gX = xmalloc(sizeof(R) * n * N);
    for (int i = 0; i < T * N; ++i) gval[i] = gauss();
    or_ls_adjoint_f64(n, K, N, T, D, X, idx, cnt, val, gval, gD, gX);
    R* Dout = xmalloc(sizeof(R) * n * K);
    double rmse[3];
    int rc = or_ksvd_f64(n, K, N, T, X, D0, Dout, 3, 11, 0.8, 4, 5, idx, val, cnt, rmse, e2);
    double chk = 0;
    for (int i = 0; i < n * K; ++i) chk += Dout[i] * Dout[i];
    printf("n=%d K=%d N=%d T=%d eps=%d rc=%d rmse %.6f %.6f %.6f |D|^2 %.3f\n", n, K, N, T, eps, rc, rmse[0],
           rmse[1], rmse[2], chk);
    free(D0); free(D); free(X); free(G); free(A0); free(A1); free(sel); free(c); free(val); free(idx); free(cnt);
    free(gap); free(pending); free(xn2); free(e2); free(gval); free(gD); free(gX); free(Dout);
    return rc;
}

int main(void) {
    /* sparsity and error mode; an unused atom (K > signals' needs) exercises the special case */
    if (run(16, 24, 60, 4, 0) || run(16, 24, 60, 4, 1) || run(9, 40, 12, 3, 0) || run(32, 32, 80, 8, 1)) return 1;
    return 0;
}
"""

MAT_DRIVER = r"""
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>
#include "ksvd_amd.h"

static std::vector<unsigned char> slurp(const char* p) {
    std::vector<unsigned char> b;
    FILE* f = std::fopen(p, "rb");
    if (!f) return b;
    int ch;
    while ((ch = std::fgetc(f)) != EOF) b.push_back((unsigned char)ch);
    std::fclose(f);
    return b;
}
static void spit(const std::string& p, const std::vector<unsigned char>& b) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return;
    if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}
/* reads variable `name` (size query, then the data) and returns its element count, or -1 */
static long read_var(const std::string& path, const char* name) {
    int64_t r = 0, c = 0;
    if (ksvd_mat_read(path.c_str(), name, nullptr, 0, &r, &c) != KSVD_OK) return -1;
    if (r < 0 || c < 0 || r * c > (1 << 26)) return -2;
    std::vector<double> v((size_t)(r * c));
    if (ksvd_mat_read(path.c_str(), name, v.data(), (int64_t)v.size(), &r, &c) != KSVD_OK) return -1;
    return (long)(r * c);
}

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    const std::string dir = argv[3];
    const char* files[2] = {argv[1], argv[2]};
    const char* names[2] = {"D", "CUDA_Gamma"};
    long good[2];
    for (int f = 0; f < 2; ++f) {
        good[f] = read_var(files[f], names[f]);
        std::printf("%s: %ld elements\n", names[f], good[f]);
        if (good[f] <= 0) return 3;
    }
    /* round trip through the writer (compressed flag off: doubles as written) */
    std::vector<double> m(6 * 7);
    for (size_t i = 0; i < m.size(); ++i) m[i] = 0.25 * (double)i - 3.0;
    const std::string w = dir + "/w.mat";
    if (ksvd_mat_write(w.c_str(), "M", m.data(), 6, 7, 0, 1) != KSVD_OK) return 4;
    if (ksvd_mat_write(w.c_str(), "N", m.data(), 7, 6, 1, 0) != KSVD_OK) return 4;
    if (read_var(w, "M") != 42 || read_var(w, "N") != 42 || read_var(w, "Q") != -1) return 5;
    /* damaged copies: every read must end in success or an error, nothing else */
    long ok = 0, err = 0;
    for (int f = 0; f < 2; ++f) {
        const std::vector<unsigned char> b = slurp(files[f]);
        const std::string p = dir + "/bad.mat";
        for (size_t cut = 0; cut < b.size(); cut += (cut < 256 ? 7 : 997)) {
            spit(p, std::vector<unsigned char>(b.begin(), b.begin() + (long)cut));
            (read_var(p, names[f]) >= 0 ? ok : err)++;
        }
        uint64_t s = 88172645463325252ull;
        for (int k = 0; k < 300; ++k) {
            std::vector<unsigned char> c = b;
            for (int j = 0; j < 4; ++j) {
                s ^= s << 13; s ^= s >> 7; s ^= s << 17;
                const size_t at = (size_t)(s % (uint64_t)c.size());
                c[at] = (unsigned char)(c[at] ^ (unsigned char)(1u + (s >> 32) % 255u));
            }
            spit(p, c);
            (read_var(p, names[f]) >= 0 ? ok : err)++;
        }
    }
    std::printf("damaged files: %ld read, %ld rejected\n", ok, err);
    return err > 0 ? 0 : 6;
}
"""


def _build_and_run(cmd, run, tmp_path):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run(run, capture_output=True, text=True, timeout=600, env=ENV, cwd=str(tmp_path))
    report = r.stderr[-6000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "LeakSanitizer" not in r.stderr, report
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], report)
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    drv = tmp_path / "oracle_driver.c"
    drv.write_text(ORACLE_DRIVER)
    exe = tmp_path / "oracle_san"
    out = _build_and_run(["gcc", "-std=c11", *SAN, "-DREAL=double", "-DSUFFIX=_f64", "-Wno-unknown-pragmas",
                          os.path.join(ROOT, "oracle", "ksvd_oracle.c"), str(drv), "-o", str(exe), "-lm"],
                         [str(exe)], tmp_path)
    assert out.count("rc=0") == 4, out


def test_mat_io_under_sanitizers(tmp_path):
    drv = tmp_path / "mat_driver.cpp"
    drv.write_text(MAT_DRIVER)
    exe = tmp_path / "mat_san"
    gold = os.path.join(ROOT, "tests", "golden")
    out = _build_and_run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"),
                          os.path.join(ROOT, "ksvd_amd", "csrc", "matio.cpp"), str(drv), "-o", str(exe), "-lz"],
                         [str(exe), os.path.join(gold, "ref_D.mat"), os.path.join(gold, "ref_temp.mat"),
                          str(tmp_path)], tmp_path)
    assert "D: 2700 elements" in out and "CUDA_Gamma: 32400 elements" in out, out
