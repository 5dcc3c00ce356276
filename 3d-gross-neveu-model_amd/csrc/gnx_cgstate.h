// gnx_cgstate.h -- per-system state of the FP64 CG state machines
// (gnx_cgm.hip, one lattice per launch; gnx_cgu.hip, systems of uneven size).
#pragma once

#include "gnx_common.h"

namespace gnx {

// Iterations whose solution and direction updates the deferred form of the
// one-reduction kernels applies together (x += alpha p, p = r + beta p, in
// iteration order), and the ring of residual buffers that keeps their
// residuals: CGD_M + 1 buffers, a power of two that divides the graph chunk.
constexpr int CGD_M = 7;
constexpr int CGD_R = CGD_M + 1;

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
  // deferred form only: the pending iterations' count, step lengths and
  // direction coefficients
  int pend;
  int pad2;
  double hal[CGD_M];
  double hbe[CGD_M];
};

// one system of an uneven batch (gnx_cgu.hip; the mixed solver's uneven
// form in gnx_cgmix.hip): its lattice, block count and offsets into the
// concatenated buffers
struct CGUSys {
  Geom g;
  int nb;     // its blocks (ceil(V / 256), one site per thread)
  long voff;  // offset of its vectors, complex elements
  long soff;  // offset of its sigma mass, doubles
  long poff;  // offset of its partials within a slot
};

// the caller's vectors of one system of an uneven solve
struct CGUPtrs {
  const double* sm;
  const double2* b;
  double2* x;
  const double2* U;  // links (link-coupled solves), or null
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

// the whole state, with the fields of the deferred form
GNX_DEVONLY void put_state_d(CGMState* o, int active, int iters, int conv, int early, double rho, double alpha,
                             double beta, int pend, const double* hal, const double* hbe, const CGMState& cfg) {
  put_state1(o, active, iters, conv, early, rho, alpha, beta, cfg);
  o->pend = pend;
  o->pad2 = cfg.pad2;
#pragma unroll
  for (int j = 0; j < CGD_M; j++) {
    o->hal[j] = hal[j];
    o->hbe[j] = hbe[j];
  }
}

}  // namespace gnx
