// backtransform.hpp — bc_back (Q_b) and sbr_back (Q_s) application.
#pragma once

#include "workspace.hpp"

namespace cuev {

// M <- Q_b M for an n x n M (ld ldm >= n); U in the workspace's shifted layout.
template <typename T> int bc_back(Workspace<T> &ws, const T *U, T *M, i64 ldm);
// M <- Q_s M for an n x n M (ld ldm >= n); Y, W with leading dimension ws.ld.
template <typename T> int sbr_back(Workspace<T> &ws, const T *Y, const T *W, T *M, i64 ldm);

} // namespace cuev
