// ksvd_main.cpp -- command-line driver with the reference program's file and
// console contract (S/main.cu:42-159): reads X and D from level-5 MAT-files,
// runs the K-SVD iterations on one GPU through the C-ABI, prints
//   "Iteration %02d / %d complete, RMSE = %.5f"   (S/ReductionKernels.cu:294,319)
//   "Final Execution Time =>\t\t%.3f ms (%.2f sec)" (S/main.cu:117-119)
// and writes the final dense Gamma (K x N) as the global variable CUDA_Gamma
// (S/main.cu:133-142), optionally the final dictionary as FinalD.
//
//   ksvd_main [--X ./input/X.mat] [--D ./input/D.mat] [--out ./output/temp.mat]
//             [--dict-out path] [--iters 30] [--T 6] [--mu 0.8] [--use-thresh 4]
//             [--seed 0] [--omp auto|wave|qform] [--form fp32|bf16|bf16x3|bf16x6]
//             [--no-graph]
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ksvd_amd.h"

namespace {

struct Options {
    std::string X = "./input/X.mat", D = "./input/D.mat", out = "./output/temp.mat", dict_out;
    int iters = 30, T = 6, use_thresh = 4, omp = KSVD_OMP_AUTO, form = KSVD_FORM_FP32, graph = 1;
    double mu = 0.8;
    unsigned long long seed = 0;
};

void usage() {
    std::printf(
        "usage: ksvd_main [--X path] [--D path] [--out path] [--dict-out path] [--iters N] [--T N]\n"
        "                 [--mu x] [--use-thresh N] [--seed S] [--omp auto|wave|qform]\n"
        "                 [--form fp32|bf16|bf16x3|bf16x6] [--no-graph]\n"
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
    ksvd_handle* h = nullptr;
    if (ksvd_create(&h, &cfg) != KSVD_OK) {
        std::fprintf(stderr, "KSVD initialization error: %s! Aborting...\n", h ? ksvd_last_error(h) : "");
        ksvd_destroy(h);
        return 255;
    }
    float *dX = nullptr, *dD = nullptr, *dDo = nullptr, *dval = nullptr;
    int32_t *didx = nullptr, *dcnt = nullptr;
    const size_t NT = (size_t)N * (size_t)o.T;
    if (!dev_alloc(&dX, X.size(), "X") || !dev_alloc(&dD, D.size(), "D") || !dev_alloc(&dDo, D.size(), "D out") ||
        !dev_alloc(&didx, NT, "Gamma indices") || !dev_alloc(&dval, NT, "Gamma values") ||
        !dev_alloc(&dcnt, (size_t)N, "Gamma counts")) {
        ksvd_destroy(h);
        return 255;
    }
    hipMemcpy(dX, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(dD, D.data(), D.size() * sizeof(float), hipMemcpyHostToDevice);
    std::printf("Single precision mode.\n");
    std::printf("Dictionary size: %lld x %lld\n", (long long)n, (long long)K);
    std::printf("Input data size: %lld x %lld\n", (long long)n, (long long)N);
    std::printf("Running...\n\n");
    std::fflush(stdout);

    std::vector<double> rmse((size_t)o.iters);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, nullptr);
    const int rc = ksvd_execute_iterations(h, dX, dD, dDo, didx, dval, dcnt, o.iters, rmse.data());
    hipEventRecord(e1, nullptr);
    hipEventSynchronize(e1);
    if (rc != KSVD_OK) {
        std::fprintf(stderr, "An error occured while executing KSVD: %s!...Aborting\n", ksvd_last_error(h));
        ksvd_destroy(h);
        return 255;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    for (int it = 0; it < o.iters; ++it)
        std::printf("Iteration %02d / %d complete, RMSE = %.5f\n", it + 1, o.iters, rmse[(size_t)it]);
    std::printf("\nFinal Execution Time =>\t\t%.3f ms (%.2f sec)\n", ms, ms / 1000.0);
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    std::printf("Total Available Memory =>\t%.2f MB\n", total_b / 1024.0 / 1024.0);
    std::printf("Total Free Memory =>\t\t%.2f MB\n", free_b / 1024.0 / 1024.0);

    // final dense Gamma, K x N column-major (entry (k, s) at s*K + k)
    std::vector<int32_t> idx(NT), cnt((size_t)N);
    std::vector<float> val(NT);
    hipMemcpy(idx.data(), didx, NT * sizeof(int32_t), hipMemcpyDeviceToHost);
    hipMemcpy(val.data(), dval, NT * sizeof(float), hipMemcpyDeviceToHost);
    hipMemcpy(cnt.data(), dcnt, (size_t)N * sizeof(int32_t), hipMemcpyDeviceToHost);
    std::vector<double> gamma((size_t)K * (size_t)N, 0.0);
    for (int64_t s = 0; s < N; ++s)
        for (int j = 0; j < cnt[(size_t)s]; ++j) {
            const int32_t k = idx[(size_t)s * o.T + j];
            if (k >= 0 && k < K) gamma[(size_t)s * (size_t)K + (size_t)k] = val[(size_t)s * o.T + j];
        }
    int wrc = ksvd_mat_write(o.out.c_str(), "CUDA_Gamma", gamma.data(), K, N, 0, 1);
    if (wrc != KSVD_OK) {
        std::fprintf(stderr, "Cannot write file for temporary array (%s)! Aborting...\n", ksvd_mat_last_error());
        ksvd_destroy(h);
        return 255;
    }
    if (!o.dict_out.empty()) {
        std::vector<float> Dout(D.size());
        hipMemcpy(Dout.data(), dDo, D.size() * sizeof(float), hipMemcpyDeviceToHost);
        std::vector<double> Dd(Dout.begin(), Dout.end());
        if (ksvd_mat_write(o.dict_out.c_str(), "FinalD", Dd.data(), n, K, 0, 0) != KSVD_OK) {
            std::fprintf(stderr, "An error occured while writing file (%s)!...\n", ksvd_mat_last_error());
            ksvd_destroy(h);
            return 255;
        }
    }
    hipFree(dX);
    hipFree(dD);
    hipFree(dDo);
    hipFree(didx);
    hipFree(dval);
    hipFree(dcnt);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ksvd_destroy(h);
    std::printf("----------------------------------------------------\n");
    std::printf("# Exiting!\n");
    return 0;
}
