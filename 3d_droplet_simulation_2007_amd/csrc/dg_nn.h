// dg_nn.h -- hooks of the learned pressure guess into the tolerance-mode
// pressure solve (dg_mg.hip) and the network runtime (dg_nn.hip).
#pragma once

#include "dg_ctx.h"

namespace dg {

// Called by the PCG solve after the initial residual r = b - A x of its
// guess x (zero for the learned guess) and its sums (ds[S_RR] = r.r,
// ds[S_RR+1] = b.b) are on the stream.  Builds the network input on the
// next-coarser level from (coefficients, b, previous pressure); when
// `apply` is set, runs the network, prolongs its prediction d to the finest
// level and, with the energy-optimal step alpha = (d.r)/(d.Ad), replaces x
// by x + alpha d and r by r - alpha A d if that lowers |r| (otherwise x, r
// stay: the fallback is the guess x).  rr_slot receives the new r.r.  When
// capture is on, the input is also written to the sample store.
int guess_solve_begin(Ctx &c, double *x, double *r, double h2i, bool apply, double *rr_slot, bool *r_changed);
// Called after the solve: completes a captured sample with its target (x
// averaged to the coarse level, normalised); a solve that did not converge
// drops it.
int guess_solve_end(Ctx &c, const double *x, double h2i, bool converged);
// whether the solve must call the hooks (learned guess or capture active)
bool guess_wanted(const Ctx &c);
// Evaluation mode (dg_guess_eval): the step's system is solved three times,
// from the zero, the previous-step and the learned guess, each from the same
// previous pressure; the iteration counts are kept and the learned solve's
// result is the step's pressure.
bool guess_eval_enabled();
int guess_eval_solves(Ctx &c, int (*solve)(Ctx &));

}  // namespace dg
