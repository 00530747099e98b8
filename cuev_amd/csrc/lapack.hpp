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
    dstedc_t dstedc = nullptr;
    sstedc_t sstedc = nullptr;
    dsyevd_t dsyevd = nullptr;
    std::string name;
    void *handle = nullptr;
};

// Loaded library (nullptr if none was found). Thread-safe.
const LapackFns *lapack();

// ?stedc with compz = 'I' (eigenvectors of the tridiagonal). d: n, e: n-1,
// z: n x n (ld ldz). work/iwork sized by stedc_lwork / stedc_liwork.
template <typename T>
int stedc(int n, T *d, T *e, T *z, int ldz, T *work, i64 lwork, int *iwork, int liwork);
i64 stedc_lwork(int n);
int stedc_liwork(int n);

} // namespace cuev
