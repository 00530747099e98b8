// lapack.cpp — run-time binding of the host LAPACK used by the tridiagonal stage.
//
// The reference links MKL or LAPACK at build time (CMakeLists.txt:29-45) and
// calls ?steqr/?laed2/?laed4/?lamrg (src/cuda/tridi.cu:38-53). Here the
// tridiagonal eigensolver is LAPACK ?stedc itself, located at run time:
//   1. $CUEV_LAPACK_LIB (a path or soname), if set;
//   2. the OpenBLAS build shipped with SciPy (every symbol prefixed "scipy_");
//   3. Intel MKL's single dynamic library (libmkl_rt);
//   4. a system liblapack / libopenblas.
// The prefixed OpenBLAS comes before MKL on purpose: in a process that has
// loaded PyTorch, torch's bundled libraries export unprefixed dgemm_/dlaed4_,
// which interpose on MKL's internal calls (wrong eigenvectors or a crash); the
// prefixed build cannot be interposed. For n > 46340 the ILP64 build that
// NumPy ships (symbols "scipy_*_64_") supplies dstedc with 64-bit integers.
// The library's name is kept for reports (cuev_lapack_name()).
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
    g_fns.dsygvd = reinterpret_cast<LapackFns::dsygvd_t>(sym("dsygvd_"));
    g_fns.dsteqr = reinterpret_cast<LapackFns::dsteqr_t>(sym("dsteqr_"));
    g_fns.ssteqr = reinterpret_cast<LapackFns::ssteqr_t>(sym("ssteqr_"));
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

void try_ilp64() {
    for (const char *pat : {"/usr/local/lib/python3*/dist-packages/numpy.libs/libscipy_openblas64_*.so",
                            "/usr/lib/python3*/dist-packages/numpy.libs/libscipy_openblas64_*.so",
                            "/usr/local/lib/python3*/site-packages/numpy.libs/libscipy_openblas64_*.so"}) {
        for (const auto &p : glob_paths(pat)) {
            void *h = dlopen(p.c_str(), RTLD_NOW | RTLD_LOCAL);
            if (!h) continue;
            void *d = dlsym(h, "scipy_dstedc_64_");
            void *s = dlsym(h, "scipy_sstedc_64_");
            if (d && s) {
                g_fns.dstedc64 = reinterpret_cast<LapackFns::dstedc64_t>(d);
                g_fns.sstedc64 = reinterpret_cast<LapackFns::sstedc64_t>(s);
                g_fns.name64 = p;
                return;
            }
            dlclose(h);
        }
    }
}

} // namespace

const LapackFns *lapack() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_tried) return g_fns.dstedc ? &g_fns : nullptr;
    g_tried = true;
    if (const char *env = std::getenv("CUEV_LAPACK_LIB")) {
        if (try_lib(env, "") || try_lib(env, "scipy_")) {
            try_ilp64();
            return &g_fns;
        }
    }
    bool ok = false;
    for (const char *pat : {"/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                            "/usr/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                            "/usr/local/lib/python3*/site-packages/scipy.libs/libscipy_openblas*.so"})
        for (const auto &p : glob_paths(pat))
            if (!ok && try_lib(p, "scipy_")) ok = true;
    if (!ok)
        for (const char *p : {"libmkl_rt.so.2", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.2",
                              "/opt/conda/lib/libmkl_rt.so"})
            if (!ok && try_lib(p, "")) ok = true;
    if (!ok)
        for (const char *p : {"liblapack.so.3", "liblapack.so", "libopenblas.so.0", "libopenblas.so"})
            if (!ok && try_lib(p, "")) ok = true;
    try_ilp64();
    return ok ? &g_fns : nullptr;
}

template <typename T, typename F32, typename F64>
static int stedc_impl(F32 f32, F64 f64, const char *what, int n, T *d, T *e, T *z, int ldz, T *work,
                      i64 lwork, int64_t *iwork, i64 liwork) {
    const char compz = 'I';
    if (lwork <= 0x7fffffffLL && liwork <= 0x7fffffffLL && f32) {
        // LP64: 32-bit integer workspace lives in the first half of iwork
        int *iw = reinterpret_cast<int *>(iwork);
        const int lw = (int)lwork, liw = (int)liwork;
        int info = 0;
        f32(&compz, &n, d, e, z, &ldz, work, &lw, iw, &liw, &info);
        if (info != 0) CUEV_FAIL(CUEV_STATUS_LAPACK_ERROR, "%s returned info = %d (n = %d)", what, info, n);
        return CUEV_STATUS_SUCCESS;
    }
    if (!f64)
        CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE,
                  "%s at n = %d needs a 64-bit-integer LAPACK; none was found", what, n);
    const int64_t n64 = n, ldz64 = ldz, lw64 = lwork, liw64 = liwork;
    int64_t info = 0;
    f64(&compz, &n64, d, e, z, &ldz64, work, &lw64, iwork, &liw64, &info);
    if (info != 0)
        CUEV_FAIL(CUEV_STATUS_LAPACK_ERROR, "%s (ILP64) returned info = %lld (n = %d)", what,
                  (long long)info, n);
    return CUEV_STATUS_SUCCESS;
}

template <> int stedc<double>(int n, double *d, double *e, double *z, int ldz, double *work,
                              i64 lwork, int64_t *iwork, i64 liwork) {
    const LapackFns *L = lapack();
    if (!L) CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE, "no host LAPACK exporting dstedc was found");
    return stedc_impl<double>(L->dstedc, L->dstedc64, "dstedc", n, d, e, z, ldz, work, lwork, iwork,
                              liwork);
}

template <> int stedc<float>(int n, float *d, float *e, float *z, int ldz, float *work, i64 lwork,
                             int64_t *iwork, i64 liwork) {
    const LapackFns *L = lapack();
    if (!L) CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE, "no host LAPACK exporting sstedc was found");
    return stedc_impl<float>(L->sstedc, L->sstedc64, "sstedc", n, d, e, z, ldz, work, lwork, iwork,
                             liwork);
}

i64 stedc_lwork(int n) { return n <= 1 ? 1 : 1 + 4 * (i64)n + (i64)n * n; }
i64 stedc_liwork(int n) { return n <= 1 ? 1 : 3 + 5 * (i64)n; }

} // namespace cuev
