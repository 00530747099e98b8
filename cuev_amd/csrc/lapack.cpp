// lapack.cpp — run-time binding of the host LAPACK used by the tridiagonal stage.
//
// The reference links MKL or LAPACK at build time (CMakeLists.txt:29-45) and
// calls ?steqr/?laed2/?laed4/?lamrg (src/cuda/tridi.cu:38-53). Here the
// tridiagonal eigensolver is LAPACK ?stedc itself, located at run time:
//   1. $CUEV_LAPACK_LIB (a path or soname), if set;
//   2. Intel MKL's single dynamic library (libmkl_rt);
//   3. the OpenBLAS build shipped with SciPy (symbols prefixed "scipy_");
//   4. a system liblapack / libopenblas.
// The first library that exports dstedc (and sstedc) wins; its name is kept for
// reports (cuev_lapack_name()).
#include "lapack.hpp"

#include <dlfcn.h>
#include <glob.h>

#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

namespace cuev {
namespace {

std::mutex g_mu;
bool g_tried = false;
LapackFns g_fns;

bool try_lib(const std::string &path, const char *prefix) {
    void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) return false;
    auto sym = [&](const char *name) -> void * {
        std::string s = std::string(prefix) + name;
        return dlsym(h, s.c_str());
    };
    void *d = sym("dstedc_");
    void *s = sym("sstedc_");
    if (!d || !s) {
        dlclose(h);
        return false;
    }
    g_fns.dstedc = reinterpret_cast<LapackFns::dstedc_t>(d);
    g_fns.sstedc = reinterpret_cast<LapackFns::sstedc_t>(s);
    g_fns.dsyevd = reinterpret_cast<LapackFns::dsyevd_t>(sym("dsyevd_"));
    g_fns.name = path;
    g_fns.handle = h;
    return true;
}

std::vector<std::string> glob_paths(const char *pattern) {
    std::vector<std::string> out;
    glob_t g;
    if (glob(pattern, 0, nullptr, &g) == 0) {
        for (size_t i = 0; i < g.gl_pathc; ++i) out.emplace_back(g.gl_pathv[i]);
    }
    globfree(&g);
    return out;
}

} // namespace

const LapackFns *lapack() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_tried) return g_fns.dstedc ? &g_fns : nullptr;
    g_tried = true;
    if (const char *env = std::getenv("CUEV_LAPACK_LIB")) {
        if (try_lib(env, "") || try_lib(env, "scipy_")) return &g_fns;
    }
    for (const char *p : {"libmkl_rt.so.2", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.2",
                          "/opt/conda/lib/libmkl_rt.so"})
        if (try_lib(p, "")) return &g_fns;
    for (const char *pat : {"/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                            "/usr/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                            "/usr/local/lib/python3*/site-packages/scipy.libs/libscipy_openblas*.so"})
        for (const auto &p : glob_paths(pat))
            if (try_lib(p, "scipy_")) return &g_fns;
    for (const char *p : {"liblapack.so.3", "liblapack.so", "libopenblas.so.0", "libopenblas.so"})
        if (try_lib(p, "")) return &g_fns;
    return nullptr;
}

template <> int stedc<double>(int n, double *d, double *e, double *z, int ldz, double *work,
                              i64 lwork, int *iwork, int liwork) {
    const LapackFns *L = lapack();
    if (!L) CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE, "no host LAPACK exporting dstedc was found");
    if (lwork > 0x7fffffffLL)
        CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "dstedc workspace exceeds the LP64 LAPACK limit (n=%d)", n);
    const int lw = (int)lwork;
    int info = 0;
    const char compz = 'I';
    L->dstedc(&compz, &n, d, e, z, &ldz, work, &lw, iwork, &liwork, &info);
    if (info != 0) CUEV_FAIL(CUEV_STATUS_LAPACK_ERROR, "dstedc returned info = %d (n = %d)", info, n);
    return CUEV_STATUS_SUCCESS;
}

template <> int stedc<float>(int n, float *d, float *e, float *z, int ldz, float *work, i64 lwork,
                             int *iwork, int liwork) {
    const LapackFns *L = lapack();
    if (!L) CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE, "no host LAPACK exporting sstedc was found");
    if (lwork > 0x7fffffffLL)
        CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "sstedc workspace exceeds the LP64 LAPACK limit (n=%d)", n);
    const int lw = (int)lwork;
    int info = 0;
    const char compz = 'I';
    L->sstedc(&compz, &n, d, e, z, &ldz, work, &lw, iwork, &liwork, &info);
    if (info != 0) CUEV_FAIL(CUEV_STATUS_LAPACK_ERROR, "sstedc returned info = %d (n = %d)", info, n);
    return CUEV_STATUS_SUCCESS;
}

i64 stedc_lwork(int n) { return n <= 1 ? 1 : 1 + 4 * (i64)n + (i64)n * n; }
int stedc_liwork(int n) { return n <= 1 ? 1 : 3 + 5 * n; }

} // namespace cuev
