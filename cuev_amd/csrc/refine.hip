// refine.hip — mixed-precision solve (FP32 stages + FP64 Ogita-Aishima refinement).
#include "refine.hpp"

namespace cuev {

int solve_mixed(double *A, int n, double *eval, double *evec, hipStream_t s, cuev_solve_timer *timer,
                const cuev_options &opt, int b, int k) {
    (void)A; (void)n; (void)eval; (void)evec; (void)s; (void)timer; (void)opt; (void)b; (void)k;
    CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "mixed precision form not built yet");
}

} // namespace cuev
