// bulge.hpp — band -> tridiagonal stage.
#pragma once

#include "workspace.hpp"

namespace cuev {

// Bp (2b x n packed band, overwritten) -> tridiagonal (d[0:n], e[0:n-1], e[n-1] = 0);
// reflectors stored in ws.U (shifted layout).
template <typename T> int bc_chase(Workspace<T> &ws, T *Bp, T *d, T *e);

} // namespace cuev
