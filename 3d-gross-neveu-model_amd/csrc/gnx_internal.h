// gnx_internal.h -- declarations shared by the gnx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gnx.h"
#include "gnx_common.h"
#include "gnx_stencil.h"

namespace gnx {

Geom make_geom(int L, int LT);

hipError_t launch_sigma_mass(const Geom& g, const double* sigma, double* sm, hipStream_t st);
hipError_t launch_evolve_sigma(const Geom& g, double* sigma, double* sm, const double* mom, double eps,
                               hipStream_t st);
hipError_t launch_evolve_momenta(const Geom& g, double* mom, const double* force, double eps, hipStream_t st);
hipError_t launch_dirac(const Geom& g, int op, const double* sm, const double* in, double* out, double* tmp,
                        hipStream_t st);
hipError_t launch_sigma_force(const Geom& g, const double* sigma, double beta_s, int nflav, const double* X0,
                              const double* Y0, const double* X1, const double* Y1, double* force,
                              hipStream_t st);
hipError_t launch_reduce(int n, int kind, const void* a, const void* b, double scale, double* part, int nparts,
                         double* out, hipStream_t st);
hipError_t launch_axpy_real(int V, double alpha, const double* xv, double* y, hipStream_t st);
hipError_t launch_scale_by_field(int V, const double* sm, const double* xv, double* y, hipStream_t st);

// Device-side result block of one CG solve.
struct CGDevResult {
  int converged;
  int early_exit;
  int iterations;
  int reliable_updates;
  double residual;
  double pad[3];
};

// Scratch owned by a context for the persistent CG.
struct CGWork {
  double2* r = nullptr;
  double2* p0 = nullptr;
  double2* p1 = nullptr;
  double2* t = nullptr;
  double* part = nullptr;     // 2 slots x 2 values x max_blocks
  CGDevResult* res = nullptr;
  int max_blocks = 0;
  long vec_elems = 0;         // complex elements per work vector (V * nrhs capacity)
};

hipError_t cg_f64_run(const Geom& g, const double* sm, const double2* b, double2* x, double tol, long max_iter,
                      int start, CGWork& w, int nrhs, hipStream_t st, int device, CGDevResult* host_res);

}  // namespace gnx
