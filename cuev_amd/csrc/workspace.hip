// workspace.hip — workspace pool allocation.
#include "workspace.hpp"

#include <vector>

#include "gemm.hpp"
#include "panel_qr.hpp"
#include "util.hpp"

namespace cuev {

int resolve_blocking(int n, int &b, int &k) {
    if (n < 1) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "n must be >= 1 (got %d)", n);
    if (b == 0) b = (n >= 4096) ? 64 : 32;
    if (b != 16 && b != 32 && b != 64)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bandwidth must be 16, 32 or 64 (got %d)", b);
    if (k == 0) k = 512;
    if (k < b || k % b != 0)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "block k=%d must be a positive multiple of b=%d", k,
                  b);
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int workspace_alloc(Workspace<T> &ws, int n, int b, int k, hipStream_t s) {
    CUEV_TRY(resolve_blocking(n, b, k));
    ws = Workspace<T>{};
    ws.n = n;
    ws.b = b;
    ws.k = k;
    ws.ld = n;
    ws.ldu = round_up((i64)n + 512, 8);
    ws.stream = s;
    CUEV_HIP(hipGetDevice(&ws.device));

    auto al = [](size_t x) -> size_t { return (x + 255) & ~size_t(255); };
    const size_t S = sizeof(T);
    const size_t nn = (size_t)n * n;
    const size_t nu = (size_t)ws.ldu * n + 2048;
    size_t off = 0;
    const size_t oY = off;
    off += al(nn * S);
    const size_t oW = off;
    off += al(nn * S);
    const size_t oZ = off;
    off += al((size_t)n * k * S);
    const size_t oZt = off;
    off += al((size_t)n * k * S);
    const size_t oYh = off;
    off += al((size_t)n * k * S);
    const size_t oD = off;
    off += al((size_t)2 * k * b * S);
    const size_t oTri = off;
    off += al((size_t)b * b * S);
    const size_t oTau = off;
    off += al((size_t)b * S);
    const size_t oB = off;
    off += al((size_t)2 * b * n * S);
    const size_t oU = off;
    off += al(nu * S);
    const size_t oM = off;
    off += al(nn * S);
    const size_t od = off;
    off += al((size_t)n * S);
    const size_t oe = off;
    off += al((size_t)n * S);
    ws.gscratch_bytes = gemm_scratch_bytes(n, n, S);
    const size_t og = off;
    off += al(ws.gscratch_bytes);
    ws.pq_bytes = panel_qr_scratch_bytes(n, b);
    const size_t opq = off;
    off += al(ws.pq_bytes);
    const size_t oc = off;
    off += al(64);
    const size_t op = off;
    off += al((size_t)n * sizeof(int));
    const size_t oerr = off;
    off += al(64);
    // bc_back block reflectors: group g covers sweeps [16g, 16g+16) (sweeps 0..n-3), and
    // has as many hops as its first sweep, (n - 2 - 16g) / b + 1
    std::vector<i64> bt_off;
    if (n >= 3) {
        ws.bt_groups = (int)ceil_div(n - 2, kBtGroup);
        bt_off.resize(ws.bt_groups + 1);
        bt_off[0] = 0;
        for (int g = 0; g < ws.bt_groups; ++g)
            bt_off[g + 1] = bt_off[g] + (n - 2 - g * kBtGroup) / b + 1;
        ws.bt_blocks = bt_off[ws.bt_groups];
    }
    const size_t oBtT = off;
    off += al((size_t)(ws.bt_blocks + 1) * bt_pack_values(b) * S); // + an all-zero block
    const size_t oBtO = off;
    off += al((bt_off.size() + 1) * sizeof(i64));
    ws.pool_bytes = off;

    hipError_t e = hipMalloc(&ws.pool, ws.pool_bytes);
    if (e != hipSuccess) {
        ws.pool = nullptr;
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "workspace: hipMalloc(%zu bytes) failed: %s",
                  ws.pool_bytes, hipGetErrorString(e));
    }
    auto *base = static_cast<unsigned char *>(ws.pool);
    ws.Y = reinterpret_cast<T *>(base + oY);
    ws.W = reinterpret_cast<T *>(base + oW);
    ws.Z = reinterpret_cast<T *>(base + oZ);
    ws.Zt = reinterpret_cast<T *>(base + oZt);
    ws.Yh = reinterpret_cast<T *>(base + oYh);
    ws.Dwk = reinterpret_cast<T *>(base + oD);
    ws.Tri = reinterpret_cast<T *>(base + oTri);
    ws.tau = reinterpret_cast<T *>(base + oTau);
    ws.B = reinterpret_cast<T *>(base + oB);
    ws.U = reinterpret_cast<T *>(base + oU);
    ws.M = reinterpret_cast<T *>(base + oM);
    ws.d = reinterpret_cast<T *>(base + od);
    ws.e = reinterpret_cast<T *>(base + oe);
    ws.gscratch = base + og;
    ws.pqscratch = base + opq;
    ws.counter = reinterpret_cast<unsigned *>(base + oc);
    ws.prog = reinterpret_cast<int *>(base + op);
    ws.err = reinterpret_cast<int *>(base + oerr);
    ws.BtT = reinterpret_cast<T *>(base + oBtT);
    ws.bt_off = reinterpret_cast<i64 *>(base + oBtO);
    if (!bt_off.empty())
        CUEV_HIP(hipMemcpy(ws.bt_off, bt_off.data(), bt_off.size() * sizeof(i64), hipMemcpyHostToDevice));
    // U slack beyond ld*n is read (masked) by bc_back; give it defined contents.
    CUEV_TRY(fill2d<T>(s, ws.U, (i64)nu, (i64)nu, 1, T(0), T(0)));
    CUEV_TRY(fill_int(s, ws.err, 1, 0));
    CUEV_TRY(fill_int(s, reinterpret_cast<int *>(ws.counter), 2, 0)); // panel QR barrier counters
    return CUEV_STATUS_SUCCESS;
}

template <typename T> void workspace_free(Workspace<T> &ws) {
    if (ws.pool) (void)hipFree(ws.pool);
    ws.pool = nullptr;
}

template <typename T> int workspace_check_error(Workspace<T> &ws) {
    int h = 0;
    CUEV_HIP(hipMemcpyAsync(&h, ws.err, sizeof(int), hipMemcpyDeviceToHost, ws.stream));
    CUEV_HIP(hipStreamSynchronize(ws.stream));
    if (h != DEV_OK) {
        CUEV_TRY(fill_int(ws.stream, ws.err, 1, 0));
        CUEV_FAIL(CUEV_STATUS_WATCHDOG, "a persistent kernel's bounded wait expired (flag %d)", h);
    }
    return CUEV_STATUS_SUCCESS;
}

template int workspace_alloc<float>(Workspace<float> &, int, int, int, hipStream_t);
template int workspace_alloc<double>(Workspace<double> &, int, int, int, hipStream_t);
template void workspace_free<float>(Workspace<float> &);
template void workspace_free<double>(Workspace<double> &);
template int workspace_check_error<float>(Workspace<float> &);
template int workspace_check_error<double>(Workspace<double> &);

} // namespace cuev
