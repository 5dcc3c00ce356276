// gnx_cgstate.h -- per-system state of the FP64 CG state machines
// (gnx_cgm.hip, one lattice per launch; gnx_cgu.hip, systems of uneven size).
#pragma once

#include "gnx_common.h"

namespace gnx {

// per-system state; `active` and `iters` lead, as the replay engine polls them
struct CGMState {
  int active;
  int iters;
  int converged;
  int early;
  double rho;
  double tol;
  long max_iter;
  int start;
  int pad;
  // one-reduction form only: the step length and direction coefficient of
  // the next iteration (its own 16-byte line: the two-reduction kernels
  // never load it)
  double alpha;
  double beta;
};

// field-wise state write (an aggregate copy makes the compiler stage the
// struct through LDS)
GNX_DEVONLY void put_state(CGMState* o, int active, int iters, int conv, int early, double rho,
                           const CGMState& cfg) {
  o->active = active;
  o->iters = iters;
  o->converged = conv;
  o->early = early;
  o->rho = rho;
  o->tol = cfg.tol;
  o->max_iter = cfg.max_iter;
  o->start = cfg.start;
  o->pad = cfg.pad;  // every loaded word stays live: a dead load destination is waited on early
}

GNX_DEVONLY void put_state1(CGMState* o, int active, int iters, int conv, int early, double rho, double alpha,
                            double beta, const CGMState& cfg) {
  put_state(o, active, iters, conv, early, rho, cfg);
  o->alpha = alpha;
  o->beta = beta;
}

}  // namespace gnx
