// dbbr.hpp — full -> band reduction stage.
#pragma once

#include "workspace.hpp"

namespace cuev {

// Pack the lower band of A (n x n, ld ws.ld) into Bp (2b x n).
template <typename T> int dbbr_pack(Workspace<T> &ws, const T *A, T *Bp, int n, int b);
// Factor one rows x b panel at A (ld ws.ld); Y, W (ld ws.ld) may be null.
template <typename T> int dbbr_panel_qr(Workspace<T> &ws, T *A, T *Y, T *W, int rows, int b);
// A (lower) -> band; reflectors in ws.Y / ws.W; band packed into Bp.
template <typename T> int dbbr_reduce(Workspace<T> &ws, T *A, T *Bp);

} // namespace cuev
