// workspace.hpp — per-problem device workspace (replaces SolverHandle<T>,
// include/cuda/handle.h:61-100 and src/cuda/handle.cu:16-94).
//
// One hipMalloc'd pool carved into the buffers below. Differences from the
// reference's pool: 64-bit offsets throughout; the device D&C keeps its own
// buffers (dc.hip), so the pool holds none; U uses a shifted layout and M is a
// plain n x n buffer (the back-transform kernels predicate rows >= n instead of
// relying on a padded leading dimension); bc_back's block-reflector operands are
// pre-packed in register order (BtT); separate scratch for split-K GEMM
// partials and the panel-QR grid reduction and gate word.
#pragma once

#include "common.hpp"

namespace cuev {

struct HostBuffers; // pinned host staging for the tridiagonal stage

template <typename T> struct Workspace {
    int n = 0, b = 0, k = 0;
    i64 ld = 0;  // leading dimension of Y, W, Z, M (= n, as the reference's lda)
    i64 ldu = 0; // leading dimension of U: n + 512 rounded up to 8. Rows past each sweep's
                 // reflectors stay zero, so the back-transform's 512-row window needs no masking.
    hipStream_t stream = nullptr;
    int device = 0;

    T *Y = nullptr;   // n x n: DBBR reflectors, panel j at columns j..j+b-1, rows j+b..n-1
    T *W = nullptr;   // n x n: W = Y T per panel (same placement as Y)
    T *Z = nullptr;   // n x k: two-sided companion Z of the current block (ld n)
    T *Zt = nullptr;  // k x n: sbr_back's Y_blk^T M (ld k)
    T *Yh = nullptr;  // n x k: sbr_back_t's block Y with zeros above each panel (ld n)
    T *Dwk = nullptr; // 2k x b: small products Z^T W, Y^T W, W^T P
    T *Tri = nullptr; // b x b: T factor of the last panel
    T *tau = nullptr; // b: Householder scalars of the last panel
    T *B = nullptr;   // 2b x n: packed band
    T *U = nullptr;   // ldu x n (+ slack, zero-filled once): bulge reflectors, shifted layout
    T *M = nullptr;   // n x n: back-transform scratch
    T *d = nullptr;   // n
    T *e = nullptr;   // n
    void *gscratch = nullptr; // split-K partial sums
    size_t gscratch_bytes = 0;
    void *pqscratch = nullptr; // panel-QR reduction buffers
    size_t pq_bytes = 0;
    unsigned *counter = nullptr; // panel-QR grid barrier counter
    int *prog = nullptr;         // bc_chase per-sweep progress (n ints)
    int *err = nullptr;          // device error flag
    // bc_back block reflectors: sweeps grouped kBtGroup at a time; block (g, h) holds
    // the hop-h reflectors of sweeps [16g, 16g+16) as I - V T V^T, its operands packed
    // in the apply kernel's register order (bt_pack_values(b) values), block (g, h)
    // at index bt_off[g] + h.
    T *BtT = nullptr;
    i64 *bt_off = nullptr; // bt_groups + 1 prefix offsets
    int bt_groups = 0;
    i64 bt_blocks = 0;

    void *pool = nullptr;
    size_t pool_bytes = 0;
};

constexpr int kBtGroup = 16; // sweeps per bc_back block reflector
// values per packed bc_back block: 64 lanes x (b/16 + 1) 16-row tiles x 8 (V and -V T)
constexpr int bt_pack_values(int b) { return 64 * 8 * (b / 16 + 1); }

template <typename T> int workspace_alloc(Workspace<T> &ws, int n, int b, int k, hipStream_t s);
template <typename T> void workspace_free(Workspace<T> &ws);

// Read back and clear the device error flag; returns a status.
template <typename T> int workspace_check_error(Workspace<T> &ws);

// Validate (n, b, k) and fill defaults (b = 0 or k = 0 select defaults).
int resolve_blocking(int n, int &b, int &k);

} // namespace cuev
