// solver.hpp — stage orchestration entry points.
#pragma once

#include "context.hpp"

namespace cuev {

void options_defaults(cuev_options &o);

// Smallest merge (rows) of the device D&C that waits for a back-transform running
// beside it (the gate of dc_solve).
int kDcGateFrom(int n);

// D&C on device (d, e): eval <- eigenvalues, evec (ld ldv) <- Q_d; host ?stedc
// or the device-resident form, as ctx.dc_mode says.
template <typename T>
int tridi_dc(Context<T> &ctx, T *d, T *e, T *eval, T *evec, i64 ldv, double *dc_ms);
// evec <- Q_s Q_b evec (direct order).
template <typename T>
int back_transform(Context<T> &ctx, const T *Y, const T *W, const T *U, T *evec,
                   cuev_solve_timer *timer);
// Eigenvalues il..iu only (0-based, ascending) into eval[0 .. iu-il]: DBBR, bulge
// chasing, then Sturm multisection on the GPU; no back-transform.
template <typename T>
int solve_values(Context<T> &ctx, T *A, int il, int iu, T *eval, cuev_solve_timer *timer);
// Full solve in the precision T.
template <typename T>
int solve_native(Context<T> &ctx, T *A, T *eval, T *evec, cuev_solve_timer *timer, int bt_mode,
                 int dc_mode, int dc_leaf);

} // namespace cuev
