// gnx_internal.h -- declarations shared by the gnx translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

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
                              hipStream_t st, int accumulate = 0, int sigma_term = 1);
hipError_t launch_reduce(int n, int kind, const void* a, const void* b, double scale, double* part, int nparts,
                         double* out, hipStream_t st);
hipError_t launch_mdag_dot_batch(const Geom& g, const double* sm, const double* X, const double* v, int n,
                                 double* part, double* out, hipStream_t st);
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

// U: links in f-order [3V] for the link-coupled operator, nullptr for the live one
hipError_t cgm_f64_run(void** machine_slot, const Geom& g, const double* sm, const double2* U, const double2* b,
                       double2* x, double tol, long max_iter, int start, int nrhs, hipStream_t st,
                       CGDevResult* host_res);
void cgm_release(void* machine);
hipError_t cgmix_run(void** machine_slot, const Geom& g, const double* sm, const double2* b, double2* x, double tol,
                     double delta, long max_iter, int start, int prec, int nrhs, hipStream_t st,
                     CGDevResult* host_res);
void cgmix_release(void* machine);
// tuned launch shape (sites per thread, XCD slab mapping, two systems per
// block, loads-first kernels); the shape index, -1 if none yet
int cgmix_shape_of(void* machine, int nrhs, int prec, int* spt, int* xcd, int* pair, int* early);
int cgm_shape_of(void* machine, int nrhs, int* spt, int* xcd);

void set_err(const char* fmt, ...);

}  // namespace gnx

struct gnx_ctx {
  int device = 0;
  int L = 0, LT = 0;
  long V = 0;
  gnx::Geom g{};
  hipStream_t stream = nullptr;
  double* tmp = nullptr;        // 2V doubles
  double* red_part = nullptr;   // reduction partials
  int red_nparts = 0;
  double* red_out = nullptr;    // 16 device scalars
  double* host_out = nullptr;   // 16 pinned host scalars
  void* cgm = nullptr;          // graph-replayed CG state machine (gnx_cgm.hip)
  void* cgmix = nullptr;        // mixed-precision reliable-update CG (gnx_cgmix.hip)
  double reliable_delta = 0.0;  // residual drop that triggers an FP64 reliable update (0: per precision)
  double* bat_part = nullptr;   // per-vector partials of the batched trace terms
  size_t bat_cap = 0;           // doubles in bat_part
  double* bat_out = nullptr;    // 64 device scalars
};


namespace gnx {
// FP64 CG on the context's engine; host_res[nrhs] filled on return
hipError_t cg_run(gnx_ctx* c, const double* sm, const double2* b, double2* x, double tol, long max_iter, int start,
                  int nrhs, CGDevResult* host_res);
// the same on the link-coupled operator (links f-order [3V])
hipError_t cg_run_link(gnx_ctx* c, const double* sm, const double2* U, const double2* b, double2* x, double tol,
                       long max_iter, int start, int nrhs, CGDevResult* host_res);
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      gnx::set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return GNX_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define CHECK_ARG(cond, msg)     \
  do {                           \
    if (!(cond)) {               \
      gnx::set_err("%s", msg);        \
      return GNX_ERR_ARG;        \
    }                            \
  } while (0)

