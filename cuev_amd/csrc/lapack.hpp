// lapack.hpp — host LAPACK routines used by the tridiagonal stage.
#pragma once

#include <string>

#include "common.hpp"

namespace cuev {

struct LapackFns {
    typedef void (*dstedc_t)(const char *, const int *, double *, double *, double *, const int *,
                             double *, const int *, int *, const int *, int *);
    typedef void (*sstedc_t)(const char *, const int *, float *, float *, float *, const int *,
                             float *, const int *, int *, const int *, int *);
    typedef void (*dsyevd_t)(const char *, const char *, const int *, double *, const int *,
                             double *, double *, const int *, int *, const int *, int *);
    // 64-bit-integer (ILP64) variants, needed once n*n exceeds 2^31 (n > 46340)
    typedef void (*dstedc64_t)(const char *, const int64_t *, double *, double *, double *,
                               const int64_t *, double *, const int64_t *, int64_t *,
                               const int64_t *, int64_t *);
    typedef void (*sstedc64_t)(const char *, const int64_t *, float *, float *, float *,
                               const int64_t *, float *, const int64_t *, int64_t *,
                               const int64_t *, int64_t *);
    typedef void (*dsygvd_t)(const int *, const char *, const char *, const int *, double *,
                             const int *, double *, const int *, double *, double *, const int *,
                             int *, const int *, int *);
    typedef void (*dsteqr_t)(const char *, const int *, double *, double *, double *, const int *,
                             double *, int *);
    typedef void (*ssteqr_t)(const char *, const int *, float *, float *, float *, const int *,
                             float *, int *);
    dsteqr_t dsteqr = nullptr;
    ssteqr_t ssteqr = nullptr;
    dstedc_t dstedc = nullptr;
    sstedc_t sstedc = nullptr;
    dsyevd_t dsyevd = nullptr;
    dsygvd_t dsygvd = nullptr;
    dstedc64_t dstedc64 = nullptr;
    sstedc64_t sstedc64 = nullptr;
    std::string name;
    std::string name64;
    void *handle = nullptr;
};

// Loaded library (nullptr if none was found). Thread-safe.
const LapackFns *lapack();

// ?stedc with compz = 'I' (eigenvectors of the tridiagonal). d: n, e: n-1,
// z: n x n (ld ldz). work/iwork sized by stedc_lwork / stedc_liwork. The LP64
// routine is used while lwork fits in 32 bits, the ILP64 one beyond.
template <typename T>
int stedc(int n, T *d, T *e, T *z, int ldz, T *work, i64 lwork, int64_t *iwork, i64 liwork);
i64 stedc_lwork(int n);
i64 stedc_liwork(int n);

} // namespace cuev
