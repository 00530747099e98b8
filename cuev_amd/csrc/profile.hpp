// profile.hpp — opt-in per-launch timing with HIP events.
//
// When enabled (cuev_profile_enable(1) or CUEV_PROFILE=1 in the environment),
// every instrumented launch site records a pair of events on its stream plus
// the algorithmic work it performs (flops, bytes). cuev_profile_read() waits
// for the recorded events and returns per-tag totals as JSON. Disabled, a
// Scope costs one branch.
#pragma once

#include "common.hpp"

namespace cuev {
namespace prof {

bool enabled();

class Scope {
  public:
    Scope(hipStream_t s, const char *tag, double flops = 0.0, double bytes = 0.0);
    ~Scope();
    Scope(const Scope &) = delete;
    Scope &operator=(const Scope &) = delete;

  private:
    int slot_ = -1;
    hipStream_t s_ = nullptr;
};

} // namespace prof
} // namespace cuev
