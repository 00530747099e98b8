// drop_in_caller.cpp — a reference-style caller of the C++ drop-in header.
//
// The call shape is the reference's own (bench/bench_solver.cpp:200-214): device
// buffers for A, eval and evec, a SolveTimer, then
//   cuev::symm_eig_solve<T>(dA, n, d_eval, d_evec, stream, &timer);
// with hipStream_t in place of cudaStream_t. Only include/cuev.h and the C ABI
// library are used. The input is A = (G + G^T)/2 with G from a fixed linear
// congruential sequence; the host checks ||AV - V diag(lam)||_F / ||A||_F,
// ||V^T V - I||_F / n and the ascending order in double, prints one line per
// precision and exits non-zero if a check fails.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "cuev.h"

#define HIP_OK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "%s -> %s\n", #x, hipGetErrorString(e_));                         \
            return 2;                                                                              \
        }                                                                                          \
    } while (0)

static std::vector<double> make_matrix(int n) {
    std::vector<double> g((size_t)n * n), a((size_t)n * n);
    uint64_t x = 88172645463325252ull;
    for (auto &v : g) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        v = (double)(x >> 11) / 9007199254740992.0 - 0.5;
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) a[i + (size_t)j * n] = 0.5 * (g[i + (size_t)j * n] + g[j + (size_t)i * n]);
    return a;
}

template <typename T> static int run(const char *name, int n, double tol) {
    const std::vector<double> a = make_matrix(n);
    std::vector<T> hA(a.begin(), a.end());
    hipStream_t stream;
    HIP_OK(hipStreamCreate(&stream));
    T *dA, *d_eval, *d_evec;
    HIP_OK(hipMalloc(&dA, (size_t)n * n * sizeof(T)));
    HIP_OK(hipMalloc(&d_eval, (size_t)n * sizeof(T)));
    HIP_OK(hipMalloc(&d_evec, (size_t)n * n * sizeof(T)));
    HIP_OK(hipMemcpy(dA, hA.data(), (size_t)n * n * sizeof(T), hipMemcpyHostToDevice));

    cuev::SolveTimer timer;
    HIP_OK(hipStreamSynchronize(stream));
    cuev::symm_eig_solve<T>(dA, n, d_eval, d_evec, stream, &timer);
    HIP_OK(hipStreamSynchronize(stream));

    std::vector<T> w(n), v((size_t)n * n);
    HIP_OK(hipMemcpy(w.data(), d_eval, (size_t)n * sizeof(T), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(v.data(), d_evec, (size_t)n * n * sizeof(T), hipMemcpyDeviceToHost));
    double r2 = 0, o2 = 0, a2 = 0;
    bool ascending = true;
    for (int j = 0; j < n; ++j) {
        if (j > 0 && w[j] < w[j - 1]) ascending = false;
        for (int i = 0; i < n; ++i) {
            double s = 0, g = 0;
            for (int l = 0; l < n; ++l) {
                s += a[i + (size_t)l * n] * (double)v[l + (size_t)j * n];
                g += (double)v[l + (size_t)i * n] * (double)v[l + (size_t)j * n];
            }
            s -= (double)w[j] * (double)v[i + (size_t)j * n];
            g -= (i == j) ? 1.0 : 0.0;
            r2 += s * s;
            o2 += g * g;
            a2 += a[i + (size_t)j * n] * a[i + (size_t)j * n];
        }
    }
    const double res = std::sqrt(r2 / a2), orth = std::sqrt(o2) / n;
    std::printf("drop_in_caller %s n=%d residual=%.3e orthogonality=%.3e ascending=%d total_ms=%.2f\n", name, n,
                res, orth, (int)ascending, timer.total_ms);
    HIP_OK(hipFree(dA));
    HIP_OK(hipFree(d_eval));
    HIP_OK(hipFree(d_evec));
    HIP_OK(hipStreamDestroy(stream));
    return (res <= tol && orth <= tol && ascending && timer.total_ms > 0) ? 0 : 1;
}

int main(int argc, char **argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 300;
    int rc = run<double>("double", n, 1e-13);
    if (rc == 0) rc = run<float>("float", n, 1e-4);
    return rc;
}
