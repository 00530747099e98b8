// common.hpp — shared host/device helpers of the gfx950 eigensolver library.
//
// Error style: every internal routine returns a cuev_status_t; the first
// failure records a message in a thread-local buffer (cuev_last_error_message)
// and is propagated up to the C ABI. Nothing here calls exit() — the
// reference's CHECK macros (include/common.h:9-34) abort the process instead.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <type_traits>

#include "../../include/cuev_c.h"

namespace cuev {

using i64 = int64_t;

namespace detail {
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
} // namespace detail

#define CUEV_FAIL(status, ...)                                                                     \
    do {                                                                                           \
        ::cuev::detail::set_error(__VA_ARGS__);                                                    \
        return (status);                                                                           \
    } while (0)

#define CUEV_HIP(expr)                                                                             \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            CUEV_FAIL(CUEV_STATUS_HIP_ERROR, "%s:%d: %s -> %s", __FILE__, __LINE__, #expr,         \
                      hipGetErrorString(_e));                                                      \
    } while (0)

// Check the launch status of the kernel just enqueued.
#define CUEV_LAUNCH_CHECK() CUEV_HIP(hipGetLastError())

#define CUEV_TRY(expr)                                                                             \
    do {                                                                                           \
        int _s = (expr);                                                                           \
        if (_s != CUEV_STATUS_SUCCESS) return _s;                                                  \
    } while (0)

inline i64 ceil_div(i64 a, i64 b) { return (a + b - 1) / b; }
inline i64 round_up(i64 a, i64 b) { return ceil_div(a, b) * b; }

// Number of compute units of the current device (cached per device).
int device_cu_count();

// Device-side error flag values written by persistent kernels.
enum DeviceError : int { DEV_OK = 0, DEV_WATCHDOG = 1 };

// Bounded-spin budget of persistent kernels, in polls of an agent-scope atomic
// load plus s_sleep(1) (roughly 0.5-2 us each): a few seconds in total. Healthy
// waits (one panel column, two bulge hops) end within micro- to milliseconds.
constexpr long long kSpinLimit = 4000000LL;

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
#if defined(__HIPCC__)

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Sum over the 64 lanes of a wave; every lane receives the total.
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Agent-scope release and acquire across the XCDs' separate L2 caches. The
// release is an L2 write-back (buffer_wbl2 sc1) and the acquire an L2/L1
// invalidate (buffer_inv sc1); both complete asynchronously and count in vmcnt.
// An atomic with release order lowers to the write-back immediately followed by
// the flag store, so the flag could become visible before the write-back has
// finished; and other waves may load right after the acquiring thread passes the
// workgroup barrier. So each is followed by an explicit vmcnt(0) wait: the flag
// goes out only after the write-back, and the barrier opens only after the
// invalidate.
__device__ __forceinline__ void vm_drain() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15): gfx9 encoding, vmcnt in [3:0] and [15:14]
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void fence_release_agent() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    vm_drain();
}
__device__ __forceinline__ void fence_acquire_agent() {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    vm_drain();
}
// Release store of a progress flag / release increment of a counter at agent scope.
__device__ __forceinline__ void flag_store_release(int *p, int v) {
    fence_release_agent();
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned counter_add_release(unsigned *p, unsigned v) {
    fence_release_agent();
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Polling loads: relaxed, so a spin does not invalidate caches on every poll.
// The waiter calls fence_acquire_agent() once after the condition holds.
__device__ __forceinline__ int flag_load_relaxed(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned counter_load_relaxed(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup-wide publish/consume of global data across workgroups.
// publish: every wave waits until its own global stores have been acknowledged
// by the L2 (s_waitcnt vmcnt(0); a plain workgroup barrier waits only for LDS),
// then the workgroup barrier; after it ONE thread performs the agent-scope
// release (flag_store_release / counter_add_release: one L2 write-back that covers
// every wave's stores, waited for, then the flag). consume: after one thread has
// observed the flag with a relaxed poll, it calls fence_acquire_agent (invalidates
// the CU's L1 and the non-coherent L2 lines, waited for) before the barrier
// releases the others.
__device__ __forceinline__ void wave_stores_done() { vm_drain(); }
__device__ __forceinline__ void wg_publish_fence() {
    wave_stores_done();
    __syncthreads();
}

#endif // __HIPCC__

} // namespace cuev
