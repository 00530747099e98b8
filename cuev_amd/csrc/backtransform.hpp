// backtransform.hpp — bc_back (Q_b) and sbr_back (Q_s) application.
#pragma once

#include "workspace.hpp"

namespace cuev {

// M <- Q_b M for an n x n M (ld ldm >= n); U in the workspace's shifted layout.
template <typename T> int bc_back(Workspace<T> &ws, const T *U, T *M, i64 ldm);
// M <- Q_s M for an n x n M (ld ldm >= n); Y, W with leading dimension ws.ld.
template <typename T> int sbr_back(Workspace<T> &ws, const T *Y, const T *W, T *M, i64 ldm);
// On stream s: M <- Q_b^T M, and M <- Q_s^T (the input M is not read). With them
// X^T = Q_b^T Q_s^T is built in the order Q_s^T first (it needs only the band
// reduction, so it can run while the bulge chasing runs), then Q_b^T.
template <typename T> int bc_back_t(Workspace<T> &ws, hipStream_t s, const T *U, T *M, i64 ldm);
template <typename T>
int sbr_back_t(Workspace<T> &ws, hipStream_t s, const T *Y, const T *W, T *M, i64 ldm);

} // namespace cuev
