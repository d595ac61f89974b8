// ksvd_main.cpp -- command-line driver with the reference program's file and
// console contract (S/main.cu:42-159): reads X and D from level-5 MAT-files,
// runs the K-SVD iterations on one GPU through the C-ABI, prints
//   "Iteration %02d / %d complete, RMSE = %.5f"   (S/ReductionKernels.cu:294,319)
//   "Final Execution Time =>\t\t%.3f ms (%.2f sec)" (S/main.cu:117-119)
// and writes the final dense Gamma (K x N) as the global variable CUDA_Gamma
// (S/main.cu:133-142), optionally the final dictionary as FinalD.  The engine
// is driven through the reference-shaped C++ facade (include/ksvd_amd.hpp):
// KSVD::ExecuteIterations, or with --phases the loop phase by phase
// (normcols, then per iteration the Gram, BatchOMP, initialize,
// updateDictionary, errorComputation, clearDict).
//
//   ksvd_main [--X ./input/X.mat] [--D ./input/D.mat] [--out ./output/temp.mat]
//             [--dict-out path] [--iters 30] [--T 6] [--mu 0.8] [--use-thresh 4]
//             [--seed 0] [--omp auto|wave|qform] [--form fp32|bf16|bf16x3|bf16x6]
//             [--no-graph] [--phases]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ksvd_amd.hpp"

namespace {

struct Options {
    std::string X = "./input/X.mat", D = "./input/D.mat", out = "./output/temp.mat", dict_out;
    int iters = 30, T = 6, use_thresh = 4, omp = KSVD_OMP_AUTO, form = KSVD_FORM_FP32, graph = 1, phases = 0;
    double mu = 0.8;
    unsigned long long seed = 0;
};

void usage() {
    std::printf(
        "usage: ksvd_main [--X path] [--D path] [--out path] [--dict-out path] [--iters N] [--T N]\n"
        "                 [--mu x] [--use-thresh N] [--seed S] [--omp auto|wave|qform]\n"
        "                 [--form fp32|bf16|bf16x3|bf16x6] [--no-graph] [--phases]\n"
        "Reads X (n x N) and D (n x K) from MAT-files (variables X and D), runs the K-SVD\n"
        "iterations and writes the final dense Gamma as CUDA_Gamma.\n");
}

bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", what);
                return nullptr;
            }
            return argv[++i];
        };
        const char* v = nullptr;
        if (a == "-h" || a == "--help") { usage(); std::exit(0); }
        else if (a == "--no-graph") o.graph = 0;
        else if (a == "--phases") o.phases = 1;
        else if (a == "--X") { if (!(v = val("--X"))) return false; o.X = v; }
        else if (a == "--D") { if (!(v = val("--D"))) return false; o.D = v; }
        else if (a == "--out") { if (!(v = val("--out"))) return false; o.out = v; }
        else if (a == "--dict-out") { if (!(v = val("--dict-out"))) return false; o.dict_out = v; }
        else if (a == "--iters") { if (!(v = val("--iters"))) return false; o.iters = std::atoi(v); }
        else if (a == "--T") { if (!(v = val("--T"))) return false; o.T = std::atoi(v); }
        else if (a == "--mu") { if (!(v = val("--mu"))) return false; o.mu = std::atof(v); }
        else if (a == "--use-thresh") { if (!(v = val("--use-thresh"))) return false; o.use_thresh = std::atoi(v); }
        else if (a == "--seed") { if (!(v = val("--seed"))) return false; o.seed = std::strtoull(v, nullptr, 0); }
        else if (a == "--omp") {
            if (!(v = val("--omp"))) return false;
            const std::string s = v;
            if (s == "auto") o.omp = KSVD_OMP_AUTO;
            else if (s == "wave") o.omp = KSVD_OMP_WAVE;
            else if (s == "qform") o.omp = KSVD_OMP_QFORM;
            else { std::fprintf(stderr, "unknown --omp %s\n", v); return false; }
        } else if (a == "--form") {
            if (!(v = val("--form"))) return false;
            const std::string s = v;
            if (s == "fp32") o.form = KSVD_FORM_FP32;
            else if (s == "bf16") o.form = KSVD_FORM_BF16;
            else if (s == "bf16x3") o.form = KSVD_FORM_BF16X3;
            else if (s == "bf16x6") o.form = KSVD_FORM_BF16X6;
            else { std::fprintf(stderr, "unknown --form %s\n", v); return false; }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return false;
        }
    }
    if (o.iters < 1 || o.T < 1) {
        std::fprintf(stderr, "--iters and --T must be positive\n");
        return false;
    }
    return true;
}

bool read_matrix(const std::string& path, const char* name, std::vector<float>& out, int64_t& rows, int64_t& cols) {
    if (ksvd_mat_read(path.c_str(), name, nullptr, 0, &rows, &cols) != KSVD_OK) {
        std::fprintf(stderr, "Error reading %s from %s: %s\n", name, path.c_str(), ksvd_mat_last_error());
        return false;
    }
    std::vector<double> d((size_t)(rows * cols));
    if (ksvd_mat_read(path.c_str(), name, d.data(), (int64_t)d.size(), &rows, &cols) != KSVD_OK) {
        std::fprintf(stderr, "Error reading %s from %s: %s\n", name, path.c_str(), ksvd_mat_last_error());
        return false;
    }
    out.assign(d.begin(), d.end());
    return true;
}

// A HIP runtime call's status, reported on stderr when it failed.
bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
    return false;
}

template <typename T>
bool dev_alloc(T** p, size_t count, const char* what) {
    if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T) + 256) != hipSuccess) {
        std::fprintf(stderr, "Device memory allocation failed (%s)! Aborting...\n", what);
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) {
        usage();
        return 2;
    }
    std::printf("Initializing! Please wait...\n");
    std::vector<float> X, D;
    int64_t n = 0, N = 0, nD = 0, K = 0;
    if (!read_matrix(o.X, "X", X, n, N) || !read_matrix(o.D, "D", D, nD, K)) {
        std::fprintf(stderr, "Host Memory initialization error! Aborting...\n");
        return 255;
    }
    if (nD != n) {
        std::fprintf(stderr, "D has %lld rows but X has %lld! Aborting...\n", (long long)nD, (long long)n);
        return 255;
    }
    ksvd_config cfg;
    ksvd_config_default(&cfg);
    cfg.n = (int)n;
    cfg.K = (int)K;
    cfg.N = (int)N;
    cfg.T = o.T;
    cfg.iterations = o.iters;
    cfg.mu_thresh = o.mu;
    cfg.use_thresh = o.use_thresh;
    cfg.perm_seed = o.seed;
    cfg.omp_variant = o.omp;
    cfg.gemm_form = o.form;
    cfg.use_graph = o.graph;
    float *dX = nullptr, *dD = nullptr, *dDo = nullptr, *dval = nullptr, *dG = nullptr, *derr = nullptr;
    int32_t *didx = nullptr, *dcnt = nullptr;
    const size_t NT = (size_t)N * (size_t)o.T;
    auto release = [&]() {
        (void)hipFree(dX);
        (void)hipFree(dD);
        (void)hipFree(dDo);
        (void)hipFree(didx);
        (void)hipFree(dval);
        (void)hipFree(dcnt);
        (void)hipFree(dG);
        (void)hipFree(derr);
    };
    if (!dev_alloc(&dX, X.size(), "X") || !dev_alloc(&dD, D.size(), "D") || !dev_alloc(&dDo, D.size(), "D out") ||
        !dev_alloc(&didx, NT, "Gamma indices") || !dev_alloc(&dval, NT, "Gamma values") ||
        !dev_alloc(&dcnt, (size_t)N, "Gamma counts") || !dev_alloc(&dG, (size_t)(K * K), "G") ||
        !dev_alloc(&derr, (size_t)N, "errors")) {
        release();
        return 255;
    }
    if (!hip_ok(hipMemcpy(dX, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice), "copy of X") ||
        !hip_ok(hipMemcpy(dD, D.data(), D.size() * sizeof(float), hipMemcpyHostToDevice), "copy of D")) {
        release();
        return 255;
    }
    ksvd_amd::DeviceMemory mem;
    mem.X = dX;
    mem.D_in = dD;
    mem.D = dDo;
    mem.G = dG;
    mem.idx = didx;
    mem.val = dval;
    mem.cnt = dcnt;
    mem.err = derr;
    ksvd_amd::KSVD ksvd(cfg, mem);
    if (!ksvd.isValid()) {
        std::fprintf(stderr, "KSVD initialization error: %s! Aborting...\n", ksvd.lastError());
        release();
        return 255;
    }
    std::printf("Single precision mode.\n");
    std::printf("Dictionary size: %lld x %lld\n", (long long)n, (long long)K);
    std::printf("Input data size: %lld x %lld\n", (long long)n, (long long)N);
    std::printf("Running...\n\n");
    std::fflush(stdout);

    std::vector<double> rmse((size_t)o.iters);
    std::vector<uint32_t> perms;
    if (o.phases) {  // the atom orders ExecuteIterations uses, made on the host beforehand
        perms.resize((size_t)o.iters * (size_t)K);
        ksvd_perm_sequence(o.seed, (int)K, o.iters, perms.data());
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hip_ok(hipEventCreate(&e0), "hipEventCreate") && hip_ok(hipEventCreate(&e1), "hipEventCreate") &&
              hip_ok(hipEventRecord(e0, nullptr), "hipEventRecord");
    if (ok && !o.phases) {
        ok = ksvd.ExecuteIterations((unsigned)o.iters);
        if (ok) rmse = ksvd.rmse();
    } else if (ok) {
        ok = ksvd.normcols();
        for (int it = 0; ok && it < o.iters; ++it)
            ok = ksvd.gram() && ksvd.BatchOMP() && ksvd.initialize() &&
                 ksvd.updateDictionary(reinterpret_cast<const unsigned*>(perms.data() + (size_t)it * (size_t)K)) &&
                 ksvd.errorComputation(&rmse[(size_t)it]) && ksvd.clearDict();
    }
    ok = ok && hip_ok(hipEventRecord(e1, nullptr), "hipEventRecord") && hip_ok(hipEventSynchronize(e1), "hipEventSynchronize");
    if (!ok) {
        std::fprintf(stderr, "An error occured while executing KSVD: %s!...Aborting\n", ksvd.lastError());
        release();
        return 255;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    for (int it = 0; it < o.iters; ++it)
        std::printf("Iteration %02d / %d complete, RMSE = %.5f\n", it + 1, o.iters, rmse[(size_t)it]);
    std::printf("\nFinal Execution Time =>\t\t%.3f ms (%.2f sec)\n", ms, ms / 1000.0);
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    std::printf("Total Available Memory =>\t%.2f MB\n", total_b / 1024.0 / 1024.0);
    std::printf("Total Free Memory =>\t\t%.2f MB\n", free_b / 1024.0 / 1024.0);

    // final dense Gamma, K x N column-major (entry (k, s) at s*K + k)
    std::vector<int32_t> idx(NT), cnt((size_t)N);
    std::vector<float> val(NT);
    if (!hip_ok(hipMemcpy(idx.data(), didx, NT * sizeof(int32_t), hipMemcpyDeviceToHost), "copy of Gamma") ||
        !hip_ok(hipMemcpy(val.data(), dval, NT * sizeof(float), hipMemcpyDeviceToHost), "copy of Gamma") ||
        !hip_ok(hipMemcpy(cnt.data(), dcnt, (size_t)N * sizeof(int32_t), hipMemcpyDeviceToHost), "copy of Gamma")) {
        release();
        return 255;
    }
    std::vector<double> gamma((size_t)K * (size_t)N, 0.0);
    for (int64_t s = 0; s < N; ++s)
        for (int j = 0; j < cnt[(size_t)s]; ++j) {
            const int32_t k = idx[(size_t)s * o.T + j];
            if (k >= 0 && k < K) gamma[(size_t)s * (size_t)K + (size_t)k] = val[(size_t)s * o.T + j];
        }
    int wrc = ksvd_mat_write(o.out.c_str(), "CUDA_Gamma", gamma.data(), K, N, 0, 1);
    if (wrc != KSVD_OK) {
        std::fprintf(stderr, "Cannot write file for temporary array (%s)! Aborting...\n", ksvd_mat_last_error());
        release();
        return 255;
    }
    if (!o.dict_out.empty()) {
        std::vector<float> Dout(D.size());
        if (!hip_ok(hipMemcpy(Dout.data(), dDo, D.size() * sizeof(float), hipMemcpyDeviceToHost), "copy of D")) {
            release();
            return 255;
        }
        std::vector<double> Dd(Dout.begin(), Dout.end());
        if (ksvd_mat_write(o.dict_out.c_str(), "FinalD", Dd.data(), n, K, 0, 0) != KSVD_OK) {
            std::fprintf(stderr, "An error occured while writing file (%s)!...\n", ksvd_mat_last_error());
            release();
            return 255;
        }
    }
    release();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::printf("----------------------------------------------------\n");
    std::printf("# Exiting!\n");
    return 0;
}
