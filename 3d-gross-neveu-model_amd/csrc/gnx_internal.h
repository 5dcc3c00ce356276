// gnx_internal.h -- declarations shared by the gnx translation units.
#pragma once

#include <vector>

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/gnx.h"
#include "gnx_common.h"
#include "gnx_stencil.h"

namespace gnx {


hipError_t launch_sigma_mass(const Geom& g, const double* sigma, double* sm, hipStream_t st);
hipError_t launch_evolve_sigma(const Geom& g, double* sigma, double* sm, const double* mom, double eps,
                               hipStream_t st);
hipError_t launch_evolve_momenta(const Geom& g, double* mom, const double* force, double eps, hipStream_t st);
hipError_t launch_dirac(const Geom& g, int op, const double* sm, const double* in, double* out, double* tmp,
                        hipStream_t st, int nvec = 1);
// whether launch_dirac applies M M^dag and M^dag M in one launch on g (and
// accepts batches of them)
bool normal_fused_fits(const Geom& g);
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

// uneven-batch solver (gnx_cgu.hip): system count and the lattice of system k
int multi_count(const gnx_multi* m);
bool multi_system(const gnx_multi* m, int k, int* L, int* LT, int* device);

// what the mixed-precision form of an uneven solve needs of the solver: its
// systems' descriptors (struct CGUSys, gnx_cgstate.h) on the device and the
// host, the concatenated sizes, and the slot of its own buffers
struct CGUSys;
struct MultiView {
  int n = 0, device = 0;
  hipStream_t stream = nullptr;
  const CGUSys* dsys = nullptr;
  const CGUSys* hsys = nullptr;
  long vtotal = 0, stotal = 0, ptotal = 0;
  int nbmax = 0;
  void** mixed = nullptr;
};
bool multi_view(gnx_multi* m, MultiView* v);
// an uneven solve with r, p, t in reduced-precision storage and FP64
// reliable updates (gnx_cgmix.hip); the caller's sm[k], b[k], x[k]
hipError_t cgmix_multi_run(const MultiView& mv, const double* const* sm, const double* const* b, double* const* x,
                           double tol, double delta, long max_iter, int start, int prec, CGDevResult* host_res);
void cgmix_multi_release(void* machine);
// gnx_multi_cg_prec with the reliable-update factor given (the ensembles
// pass their chains' own)
gnx_status multi_cg_mixed(gnx_multi* m, const double* const* sm, const double* const* b, double* const* x,
                          double tol, long max_iter, int start, int prec, double delta, gnx_cg_result* res);
int cgm_shape_of(void* machine, int nrhs, int* spt, int* xcd);
void cgm_set_autotune(void** machine_slot, int on);
// the allocation-free FP64 solve (one system): bytes of the static array it
// needs for g, the array's size, and the solve; pinned: 64 bytes of pinned
// host memory for the polled flags and the result
size_t cgm_static_bytes(const Geom& g);
size_t cgm_static_capacity();
hipError_t cgm_f64_run_static(const Geom& g, int device, const double* sm, const double2* b, double2* x, double tol,
                              long max_iter, int start, hipStream_t st, void* pinned, CGDevResult* host_res);
long cgm_tune_runs(void* machine);
void cgm_replay_stats(void* machine, long* slots, long* polls);
// FP64 CG of small lattices in one launch per solve, one workgroup per
// system (gnx_cgs.hip): whether g fits (sites and iteration cap), the solve
// (results to dres[nrhs] on the device) and its sites per thread
bool cgs_fits(const Geom& g, long max_iter);
hipError_t cgs_run(const Geom& g, const double* sm, const double2* b, double2* x, double tol, long max_iter,
                   int start, int nrhs, CGDevResult* dres, hipStream_t st);
int cgs_sites_per_thread(const Geom& g);
void cgmix_set_autotune(void** machine_slot, int on);
long cgmix_tune_runs(void* machine);
void cgmix_replay_stats(void* machine, long* slots, long* polls);

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
  int cg_pinned = -1;           // GNX_CG_SHAPE or GNX_AUTOTUNE set at the first solve (-1: no solve yet)
  int autotune = 0;             // gnx_set_autotune
  gnx::CGDevResult* cgs_res = nullptr;  // 64 device results of the one-launch small-lattice solves
  unsigned long long cgs_seen = 0;      // bit n-1: a batch of n systems took the one-launch solve
};


namespace gnx {
// FP64 CG on the context's engine; host_res[nrhs] filled on return
hipError_t cg_run(gnx_ctx* c, const double* sm, const double2* b, double2* x, double tol, long max_iter, int start,
                  int nrhs, CGDevResult* host_res);
// residual drop that triggers a mixed solve's FP64 reliable update
double reliable_delta_of(const gnx_ctx* c, int prec);
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

namespace gnx {

// One uneven-batch solve of both flavours of every live chain (systems 2k,
// 2k+1 of m).  A chain that is no longer live has its phi and X cleared, so
// its two systems leave in the prologue; ok[k] and the chain's CG counters
// follow the single-chain solve (hmc_cg, lh_cg).  H: gnx_hmc or
// gnx_link_hmc (fields c, sm, phi, X, iterations, calls).  links: chain k's
// links (3 per site) for solves with the link-coupled operator, or null.
// prec: the solves' arithmetic (mixed precision with the reliable-update
// factor delta; FP64 by default).
template <typename H>
inline gnx_status ens_solve(gnx_multi* m, int n, H* const* hs, const std::vector<char>& live, double tol,
                     long max_iter, int start, std::vector<char>& ok, const double2* const* links = nullptr,
                     int prec = GNX_PREC_F64, double delta = 0.0) {
  std::vector<const double*> sms(2 * n), bs(2 * n), us(2 * n);
  std::vector<double*> xs(2 * n);
  for (int k = 0; k < n; k++) {
    H* h = hs[k];
    gnx_ctx* c = h->c;
    const long V = c->V;
    if (!live[k]) {
      HIP_TRY(hipMemsetAsync(h->phi, 0, sizeof(double2) * 2 * V, c->stream));
      HIP_TRY(hipMemsetAsync(h->X, 0, sizeof(double2) * 2 * V, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    sms[2 * k] = sms[2 * k + 1] = h->sm;
    bs[2 * k] = (const double*)h->phi;
    bs[2 * k + 1] = (const double*)(h->phi + V);
    xs[2 * k] = (double*)h->X;
    xs[2 * k + 1] = (double*)(h->X + V);
    if (links) us[2 * k] = us[2 * k + 1] = (const double*)links[k];
  }
  std::vector<gnx_cg_result> r(2 * n);
  const gnx_status s =
      links ? gnx_multi_link_cg(m, sms.data(), us.data(), bs.data(), xs.data(), tol, max_iter, (gnx_cg_start)start,
                                r.data())
      : prec != GNX_PREC_F64
          ? multi_cg_mixed(m, sms.data(), bs.data(), xs.data(), tol, max_iter, start, prec, delta, r.data())
          : gnx_multi_cg(m, sms.data(), bs.data(), xs.data(), tol, max_iter, (gnx_cg_start)start, r.data());
  if (s != GNX_OK && s != GNX_ERR_NOCONV) return s;
  for (int k = 0; k < n; k++) {
    ok[k] = 0;
    if (!live[k]) continue;
    ok[k] = 1;
    for (int f = 0; f < 2; f++) {
      const gnx_cg_result& q = r[2 * k + f];
      if (q.converged && !q.early_exit) {
        hs[k]->iterations += q.iterations;
        hs[k]->calls += 1;
      }
      if (!q.converged) ok[k] = 0;
    }
  }
  return GNX_OK;
}

}  // namespace gnx
