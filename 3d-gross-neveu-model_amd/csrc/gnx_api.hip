// gnx_api.hip -- the C ABI (include/gnx.h): contexts, argument checks,
// scratch ownership and the host-side sequencing of the hot-path kernels.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gnx_internal.h"

using namespace gnx;

namespace gnx {
static thread_local char g_err[512] = "";

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace gnx

// geometry table export (f-order, reference layout)
__global__ void k_geometry_tables(Geom g, int* fpp, int* fmm, double* eta_pp, double* eta_mm, int* f2lex) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  Site s = decode(g, f);
  NbrF n = neighbours_f(g, s);
  HalfEta e = half_eta(g, s);
  for (int d = 0; d < 3; d++) {
    if (fpp) fpp[3 * f + d] = n.pp[d];
    if (fmm) fmm[3 * f + d] = n.mm[d];
    if (eta_pp) eta_pp[3 * f + d] = 2.0 * e.pp[d];
    if (eta_mm) eta_mm[3 * f + d] = 2.0 * e.mm[d];
  }
  if (f2lex) f2lex[f] = s.lex;
}

// the residual drop that triggers a mixed solve's FP64 reliable update:
// the context's setting, else the default per storage format, measured at
// 64^3 x 32 sources, tol 1e-10: FP32 needs the same iterations for any
// delta, so few updates (0.05); BF16 loses direction quality fast and needs
// frequent true-residual refreshes (0.2: 447 iterations, 19 updates,
// against 549 at 0.1; the rate peaks at 0.15-0.2); FP16, three mantissa
// bits more, as FP32
double gnx::reliable_delta_of(const gnx_ctx* c, int prec) {
  return c->reliable_delta > 0.0 ? c->reliable_delta : (prec == GNX_PREC_BF16 ? 0.2 : 0.05);
}

// Whether the FP64 solves of this context run in one launch per solve (the
// small-lattice kernel, up to 12^3 sites and 100 000 iterations) rather than
// as replayed graphs: under the fixed launch rule only -- a shape pinned by
// GNX_CG_SHAPE, or timing autotune, selects among the graph forms.
static bool cg_small(gnx_ctx* c, long max_iter, int nrhs) {
  if (c->cg_pinned < 0) {
    const char* at = getenv("GNX_AUTOTUNE");
    c->cg_pinned = (getenv("GNX_CG_SHAPE") || (at && atoi(at) != 0)) ? 1 : 0;
  }
  return !c->cg_pinned && !c->autotune && nrhs <= 64 && cgs_fits(c->g, max_iter);
}

hipError_t gnx::cg_run(gnx_ctx* c, const double* sm, const double2* b, double2* x, double tol, long max_iter,
                      int start, int nrhs, CGDevResult* host_res) {
  if (cg_small(c, max_iter, nrhs)) {
    hipError_t e = hipSuccess;
    if (!c->cgs_res) e = hipMalloc(&c->cgs_res, sizeof(CGDevResult) * 64);
    if (e == hipSuccess) e = cgs_run(c->g, sm, b, x, tol, max_iter, start, nrhs, c->cgs_res, c->stream);
    if (e == hipSuccess && host_res)
      e = hipMemcpyAsync(host_res, c->cgs_res, sizeof(CGDevResult) * nrhs, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) c->cgs_seen |= 1ull << (nrhs - 1);
    return e;
  }
  return cgm_f64_run(&c->cgm, c->g, sm, nullptr, b, x, tol, max_iter, start, nrhs, c->stream, host_res);
}

hipError_t gnx::cg_run_link(gnx_ctx* c, const double* sm, const double2* U, const double2* b, double2* x, double tol,
                           long max_iter, int start, int nrhs, CGDevResult* host_res) {
  return cgm_f64_run(&c->cgm, c->g, sm, U, b, x, tol, max_iter, start, nrhs, c->stream, host_res);
}

extern "C" {

int gnx_abi_version(void) { return GNX_ABI_VERSION; }
const char* gnx_last_error(void) { return gnx::g_err; }

gnx_status gnx_create(int lspace, int ltime, int device, gnx_ctx** out) {
  CHECK_ARG(out != nullptr, "gnx_create: out is NULL");
  *out = nullptr;
  CHECK_ARG(lspace >= 4 && ltime >= 4 && lspace % 2 == 0 && ltime % 2 == 0,
            "gnx_create: L and LT must be even and >= 4");
  long V = (long)lspace * lspace * ltime;
  if (V > (1L << 26)) {
    set_err("gnx_create: volume %ld exceeds 2^26 sites", V);
    return GNX_ERR_SIZE;
  }
  if (!geom_decodable(lspace, ltime)) {
    set_err("gnx_create: %d^2 x %d is outside the exact range of the in-kernel site decode", lspace, ltime);
    return GNX_ERR_SIZE;
  }
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  CHECK_ARG(device >= 0 && device < ndev, "gnx_create: no such device");
  HIP_TRY(hipSetDevice(device));
  gnx_ctx* c = new gnx_ctx();
  c->device = device;
  c->L = lspace;
  c->LT = ltime;
  c->V = V;
  c->g = make_geom(lspace, ltime);
  c->red_nparts = 1024;

  hipError_t e;
  if ((e = hipMalloc(&c->tmp, sizeof(double) * 2 * V)) != hipSuccess ||
      (e = hipMalloc(&c->red_part, sizeof(double) * c->red_nparts)) != hipSuccess ||
      (e = hipMalloc(&c->red_out, sizeof(double) * 16)) != hipSuccess ||
      (e = hipHostMalloc(&c->host_out, sizeof(double) * 16, hipHostMallocDefault)) != hipSuccess) {
    set_err("gnx_create: allocation failed: %s", hipGetErrorString(e));
    gnx_destroy(c);
    return GNX_ERR_HIP;
  }
  *out = c;
  return GNX_OK;
}

gnx_status gnx_destroy(gnx_ctx* c) {
  if (!c) return GNX_OK;
  (void)hipSetDevice(c->device);
  (void)hipFree(c->tmp);
  (void)hipFree(c->red_part);
  (void)hipFree(c->red_out);
  if (c->host_out) (void)hipHostFree(c->host_out);
  if (c->bat_part) (void)hipFree(c->bat_part);
  if (c->bat_out) (void)hipFree(c->bat_out);
  if (c->cgs_res) (void)hipFree(c->cgs_res);
  if (c->cgm) cgm_release(c->cgm);
  if (c->cgmix) cgmix_release(c->cgmix);
  delete c;
  return GNX_OK;
}

gnx_status gnx_set_stream(gnx_ctx* c, void* s) {
  CHECK_ARG(c, "null context");
  c->stream = (hipStream_t)s;
  return GNX_OK;
}

gnx_status gnx_synchronize(gnx_ctx* c) {
  CHECK_ARG(c, "null context");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

long gnx_volume(const gnx_ctx* c) { return c ? c->V : -1; }

gnx_status gnx_geometry_tables(gnx_ctx* c, int* fpp, int* fmm, double* eta_pp, double* eta_mm, int* f2lex) {
  CHECK_ARG(c, "null context");
  HIP_TRY(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_geometry_tables, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->g, fpp, fmm,
                     eta_pp, eta_mm, f2lex);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status gnx_sigma_mass(gnx_ctx* c, const double* sigma, double* sm) {
  CHECK_ARG(c && sigma && sm, "gnx_sigma_mass: null argument");
  CHECK_ARG((const void*)sigma != (const void*)sm, "gnx_sigma_mass: sigma and sm alias");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_sigma_mass(c->g, sigma, sm, c->stream));
  return GNX_OK;
}

gnx_status gnx_dirac(gnx_ctx* c, gnx_op op, const double* sm, const double* in, double* out) {
  CHECK_ARG(c && sm && in && out, "gnx_dirac: null argument");
  CHECK_ARG((const void*)in != (const void*)out, "gnx_dirac: in and out alias");
  CHECK_ARG(op >= GNX_OP_M && op <= GNX_OP_SUB_MDAG, "gnx_dirac: bad op");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_dirac(c->g, (int)op, sm, in, out, c->tmp, c->stream));
  return GNX_OK;
}

gnx_status gnx_dirac_batch(gnx_ctx* c, gnx_op op, const double* sm, const double* in, double* out, int nvec) {
  CHECK_ARG(c && sm && in && out, "gnx_dirac_batch: null argument");
  CHECK_ARG(nvec >= 1 && nvec <= 64, "gnx_dirac_batch: nvec must be in [1, 64]");
  CHECK_ARG(op == GNX_OP_M || op == GNX_OP_MDAG || op == GNX_OP_SUB_M || op == GNX_OP_SUB_MDAG ||
                ((op == GNX_OP_MMDAG || op == GNX_OP_MDAGM) && normal_fused_fits(c->g)),
            "gnx_dirac_batch: op must be one operator (M, M^dag, r -= M v, r -= M^dag v), or a normal "
            "operator on a lattice of 4 x 8 x 16 tiles (L divisible by 8, LT by 16)");
  const char* ib = (const char*)in;
  const char* ob = (const char*)out;
  const size_t bytes = sizeof(double2) * (size_t)c->V * nvec;
  CHECK_ARG(ib + bytes <= ob || ob + bytes <= ib, "gnx_dirac_batch: in and out overlap");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_dirac(c->g, (int)op, sm, in, out, c->tmp, c->stream, nvec));
  return GNX_OK;
}

static void fill_result(gnx_cg_result* r, const CGDevResult& d) {
  r->converged = d.converged;
  r->early_exit = d.early_exit;
  r->iterations = d.iterations;
  r->reliable_updates = d.reliable_updates;
  r->residual = d.residual;
}

gnx_status gnx_cg_batch(gnx_ctx* c, const double* sm, const double* b, double* x, int nrhs, double tol,
                        long max_iter, gnx_cg_start start, gnx_prec prec, gnx_cg_result* res) {
  CHECK_ARG(c && sm && b && x && res, "gnx_cg: null argument");
  CHECK_ARG(nrhs >= 1 && nrhs <= 64, "gnx_cg: nrhs must be in [1, 64]");
  CHECK_ARG(start == GNX_START_NULL || start == GNX_START_IMPOSED, "gnx_cg: bad start mode");
  CHECK_ARG(tol >= 0.0, "gnx_cg: negative tolerance");
  CHECK_ARG((const void*)b != (const void*)x, "gnx_cg: b and x alias");
  if (max_iter <= 0) max_iter = c->V + 1;
  HIP_TRY(hipSetDevice(c->device));
  if (prec == GNX_PREC_F32 || prec == GNX_PREC_BF16 || prec == GNX_PREC_F16) {
    CGDevResult hr[64];
    const double delta = reliable_delta_of(c, (int)prec);
    HIP_TRY(cgmix_run(&c->cgmix, c->g, sm, (const double2*)b, (double2*)x, tol, delta, max_iter,
                      (int)start, (int)prec, nrhs, c->stream, hr));
    int all = 1;
    for (int k = 0; k < nrhs; k++) {
      fill_result(&res[k], hr[k]);
      all &= hr[k].converged;
    }
    if (!all) {
      set_err("gnx_cg: not converged within %ld iterations", max_iter);
      return GNX_ERR_NOCONV;
    }
    return GNX_OK;
  }
  CHECK_ARG(prec == GNX_PREC_F64, "gnx_cg: bad precision");
  CGDevResult hr[64];
  HIP_TRY(cg_run(c, sm, (const double2*)b, (double2*)x, tol, max_iter, (int)start, nrhs, hr));
  int all = 1;
  for (int k = 0; k < nrhs; k++) {
    fill_result(&res[k], hr[k]);
    all &= hr[k].converged;
  }
  if (!all) {
    set_err("gnx_cg: not converged within %ld iterations", max_iter);
    return GNX_ERR_NOCONV;
  }
  return GNX_OK;
}

gnx_status gnx_cg(gnx_ctx* c, const double* sm, const double* b, double* x, double tol, long max_iter,
                  gnx_cg_start start, gnx_prec prec, gnx_cg_result* res) {
  return gnx_cg_batch(c, sm, b, x, 1, tol, max_iter, start, prec, res);
}

gnx_status gnx_cg_noalloc(gnx_ctx* c, const double* sm, const double* b, double* x, double tol, long max_iter,
                          gnx_cg_start start, gnx_cg_result* res) {
  CHECK_ARG(c && sm && b && x && res, "gnx_cg_noalloc: null argument");
  CHECK_ARG(start == GNX_START_NULL || start == GNX_START_IMPOSED, "gnx_cg_noalloc: bad start mode");
  CHECK_ARG(tol >= 0.0, "gnx_cg_noalloc: negative tolerance");
  CHECK_ARG((const void*)b != (const void*)x, "gnx_cg_noalloc: b and x alias");
  const size_t need = cgm_static_bytes(c->g), cap = cgm_static_capacity();
  if (need > cap) {
    set_err("gnx_cg_noalloc: %ld sites need %zu bytes of scratch, the static array holds %zu", c->V, need, cap);
    return GNX_ERR_SIZE;
  }
  if (max_iter <= 0) max_iter = c->V + 1;
  HIP_TRY(hipSetDevice(c->device));
  CGDevResult hr;
  if (cg_small(c, max_iter, 1)) {
    // gnx_cg's one-launch solve; its only buffer besides the caller's is the
    // result, in the context's 16 reduction scalars (128 bytes)
    static_assert(sizeof(CGDevResult) <= 16 * sizeof(double), "result block exceeds red_out");
    CGDevResult* dres = (CGDevResult*)c->red_out;
    HIP_TRY(cgs_run(c->g, sm, (const double2*)b, (double2*)x, tol, max_iter, (int)start, 1, dres, c->stream));
    HIP_TRY(hipMemcpyAsync(c->host_out, dres, sizeof(CGDevResult), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(&hr, c->host_out, sizeof(CGDevResult));
    fill_result(res, hr);
    if (!hr.converged) {
      set_err("gnx_cg_noalloc: not converged within %ld iterations", max_iter);
      return GNX_ERR_NOCONV;
    }
    return GNX_OK;
  }
  HIP_TRY(cgm_f64_run_static(c->g, c->device, sm, (const double2*)b, (double2*)x, tol, max_iter, (int)start,
                             c->stream, c->host_out, &hr));
  fill_result(res, hr);
  if (!hr.converged) {
    set_err("gnx_cg_noalloc: not converged within %ld iterations", max_iter);
    return GNX_ERR_NOCONV;
  }
  return GNX_OK;
}

gnx_status gnx_condensate_batch(gnx_ctx* c, const double* sm, const double* noise, int nsrc, double tol,
                                long max_iter, gnx_prec prec, double* X, double* per_source, gnx_cg_result* res) {
  CHECK_ARG(c && sm && noise && X && per_source && res, "gnx_condensate_batch: null argument");
  CHECK_ARG(nsrc >= 1 && nsrc <= 64, "gnx_condensate_batch: nsrc must be in [1, 64]");
  CHECK_ARG((const void*)noise != (const void*)X, "gnx_condensate_batch: noise and X alias");
  HIP_TRY(hipSetDevice(c->device));
  // one batched solve of M M^dag X_k = v_k (the reference inverts each vector
  // in turn from the previous solution, whose prologue exit never applies to a
  // fresh noise vector, so the imposed start gives the same solutions)
  gnx_status st = gnx_cg_batch(c, sm, noise, X, nsrc, tol, max_iter, GNX_START_IMPOSED, prec, res);
  if (st != GNX_OK) return st;
  const size_t need = (size_t)nsrc * (size_t)((c->V + 255) / 256);
  if (c->bat_cap < need) {
    if (c->bat_part) (void)hipFree(c->bat_part);
    c->bat_part = nullptr;
    c->bat_cap = 0;
    HIP_TRY(hipMalloc(&c->bat_part, sizeof(double) * need));
    c->bat_cap = need;
  }
  if (!c->bat_out) HIP_TRY(hipMalloc(&c->bat_out, sizeof(double) * 64));
  HIP_TRY(launch_mdag_dot_batch(c->g, sm, X, noise, nsrc, c->bat_part, c->bat_out, c->stream));
  HIP_TRY(hipMemcpyAsync(per_source, c->bat_out, sizeof(double) * nsrc, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

gnx_status gnx_set_reliable_delta(gnx_ctx* c, double delta) {
  CHECK_ARG(c, "gnx_set_reliable_delta: null context");
  CHECK_ARG(delta >= 0.0 && delta < 1.0, "gnx_set_reliable_delta: delta must be in [0, 1)");
  c->reliable_delta = delta;
  return GNX_OK;
}

gnx_status gnx_cg_shape(gnx_ctx* c, int nrhs, int* spt, int* xcd) {
  CHECK_ARG(c, "gnx_cg_shape: null context");
  if (nrhs >= 1 && nrhs <= 64 && (c->cgs_seen >> (nrhs - 1) & 1) && cg_small(c, 0, nrhs)) {
    if (spt) *spt = cgs_sites_per_thread(c->g);
    if (xcd) *xcd = 16;
    return GNX_OK;
  }
  if (cgm_shape_of(c->cgm, nrhs, spt, xcd) < 0) {
    set_err("gnx_cg_shape: no shape chosen for nrhs=%d yet", nrhs);
    return GNX_ERR_STATE;
  }
  return GNX_OK;
}

gnx_status gnx_set_autotune(gnx_ctx* c, int on) {
  CHECK_ARG(c, "gnx_set_autotune: null context");
  c->autotune = on ? 1 : 0;
  cgm_set_autotune(&c->cgm, on ? 1 : 0);
  cgmix_set_autotune(&c->cgmix, on ? 1 : 0);
  return GNX_OK;
}

gnx_status gnx_replay_stats(gnx_ctx* c, long* slots, long* polls) {
  CHECK_ARG(c && slots && polls, "gnx_replay_stats: null argument");
  long s1, p1, s2, p2;
  cgm_replay_stats(c->cgm, &s1, &p1);
  cgmix_replay_stats(c->cgmix, &s2, &p2);
  *slots = s1 + s2;
  *polls = p1 + p2;
  return GNX_OK;
}

gnx_status gnx_tune_runs(gnx_ctx* c, long* fp64, long* mixed) {
  CHECK_ARG(c && fp64 && mixed, "gnx_tune_runs: null argument");
  *fp64 = cgm_tune_runs(c->cgm);
  *mixed = cgmix_tune_runs(c->cgmix);
  return GNX_OK;
}

gnx_status gnx_cg_mixed_shape(gnx_ctx* c, int nrhs, gnx_prec prec, int* spt, int* xcd, int* pair,
                              int* loads_first) {
  CHECK_ARG(c && spt && xcd && pair && loads_first, "gnx_cg_mixed_shape: null argument");
  if (cgmix_shape_of(c->cgmix, nrhs, (int)prec, spt, xcd, pair, loads_first) < 0) {
    set_err("gnx_cg_mixed_shape: no shape chosen for nrhs=%d yet", nrhs);
    return GNX_ERR_STATE;
  }
  return GNX_OK;
}

gnx_status gnx_sigma_force_from_solutions(gnx_ctx* c, const double* sigma, double beta_s, int nflav,
                                          const double* X0, const double* Y0, const double* X1, const double* Y1,
                                          double* force) {
  CHECK_ARG(c && sigma && X0 && Y0 && X1 && Y1 && force, "gnx_sigma_force: null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_sigma_force(c->g, sigma, beta_s, nflav, X0, Y0, X1, Y1, force, c->stream));
  return GNX_OK;
}

gnx_status gnx_sigma_force_terms(gnx_ctx* c, const double* sigma, double beta_s, int nflav, const double* X0,
                                 const double* Y0, const double* X1, const double* Y1, double* force,
                                 int accumulate, int sigma_term) {
  CHECK_ARG(c && X0 && Y0 && X1 && Y1 && force, "gnx_sigma_force_terms: null argument");
  CHECK_ARG(sigma || !sigma_term, "gnx_sigma_force_terms: the sigma term needs sigma");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_sigma_force(c->g, sigma, beta_s, nflav, X0, Y0, X1, Y1, force, c->stream, accumulate, sigma_term));
  return GNX_OK;
}

gnx_status gnx_sigma_force(gnx_ctx* c, const double* sigma, const double* sm, double beta_s, int nflav,
                           double tol, long max_iter, const double* const phi[2], double* const X[2],
                           double* const Y[2], double* force, gnx_cg_result res[2]) {
  CHECK_ARG(c && sigma && sm && phi && X && Y && force && res, "gnx_sigma_force: null argument");
  CHECK_ARG(phi[0] && phi[1] && X[0] && X[1] && Y[0] && Y[1], "gnx_sigma_force: null vector");
  long V2 = 2 * c->V;
  gnx_status s;
  if (phi[1] == phi[0] + V2 && X[1] == X[0] + V2) {
    s = gnx_cg_batch(c, sm, phi[0], X[0], 2, tol, max_iter, GNX_START_NULL, GNX_PREC_F64, res);
  } else {
    s = gnx_cg_batch(c, sm, phi[0], X[0], 1, tol, max_iter, GNX_START_NULL, GNX_PREC_F64, &res[0]);
    if (s != GNX_OK) return s;
    s = gnx_cg_batch(c, sm, phi[1], X[1], 1, tol, max_iter, GNX_START_NULL, GNX_PREC_F64, &res[1]);
  }
  if (s != GNX_OK) return s;
  for (int f = 0; f < 2; f++) HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, sm, X[f], Y[f], c->tmp, c->stream));
  HIP_TRY(launch_sigma_force(c->g, sigma, beta_s, nflav, X[0], Y[0], X[1], Y[1], force, c->stream));
  return GNX_OK;
}

gnx_status gnx_evolve_momenta(gnx_ctx* c, double* mom, const double* force, double eps) {
  CHECK_ARG(c && mom && force, "gnx_evolve_momenta: null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_evolve_momenta(c->g, mom, force, eps, c->stream));
  return GNX_OK;
}

gnx_status gnx_evolve_sigma(gnx_ctx* c, double* sigma, double* sm, const double* mom, double eps) {
  CHECK_ARG(c && sigma && sm && mom, "gnx_evolve_sigma: null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_evolve_sigma(c->g, sigma, sm, mom, eps, c->stream));
  return GNX_OK;
}

// reductions: results land in c->red_out[slot], copied to host at the end
static gnx_status reduce_to(gnx_ctx* c, int n, int kind, const void* a, const void* b, double scale, int slot) {
  HIP_TRY(launch_reduce(n, kind, a, b, scale, c->red_part, c->red_nparts, c->red_out + slot, c->stream));
  return GNX_OK;
}

static gnx_status fetch(gnx_ctx* c, int n) {
  HIP_TRY(hipMemcpyAsync(c->host_out, c->red_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

gnx_status gnx_energy(gnx_ctx* c, const double* mom, const double* sigma, double beta_s, int nflav,
                      const double* const phi[2], const double* const X[2], double* H) {
  CHECK_ARG(c && mom && sigma && phi && X && H, "gnx_energy: null argument");
  HIP_TRY(hipSetDevice(c->device));
  gnx_status s;
  if ((s = reduce_to(c, (int)c->V, 3, mom, nullptr, 1.0, 0)) != GNX_OK) return s;
  if ((s = reduce_to(c, (int)c->V, 1, phi[0], X[0], 1.0, 1)) != GNX_OK) return s;
  if ((s = reduce_to(c, (int)c->V, 1, phi[1], X[1], 1.0, 2)) != GNX_OK) return s;
  if ((s = reduce_to(c, (int)c->V, 3, sigma, nullptr, 1.0, 3)) != GNX_OK) return s;
  if ((s = fetch(c, 4)) != GNX_OK) return s;
  const double* o = c->host_out;
  *H = o[0] / 2.0 + 0.5 * o[1] + 0.5 * o[2] + (nflav * beta_s / 4.0) * o[3];
  return GNX_OK;
}

gnx_status gnx_dot(gnx_ctx* c, const double* a, const double* b, double* re, double* im) {
  CHECK_ARG(c && a && b, "gnx_dot: null argument");
  HIP_TRY(hipSetDevice(c->device));
  gnx_status s;
  if ((s = reduce_to(c, (int)c->V, 1, a, b, 1.0, 0)) != GNX_OK) return s;
  if ((s = reduce_to(c, (int)c->V, 2, a, b, 1.0, 1)) != GNX_OK) return s;
  if ((s = fetch(c, 2)) != GNX_OK) return s;
  if (re) *re = c->host_out[0];
  if (im) *im = c->host_out[1];
  return GNX_OK;
}

gnx_status gnx_norm2(gnx_ctx* c, const double* a, double* out) {
  CHECK_ARG(c && a && out, "gnx_norm2: null argument");
  HIP_TRY(hipSetDevice(c->device));
  gnx_status s;
  if ((s = reduce_to(c, (int)c->V, 0, a, nullptr, 1.0, 0)) != GNX_OK) return s;
  if ((s = fetch(c, 1)) != GNX_OK) return s;
  *out = c->host_out[0];
  return GNX_OK;
}

gnx_status gnx_scale_by_field(gnx_ctx* c, const double* field, const double* x, double* y) {
  CHECK_ARG(c && field && x && y, "gnx_scale_by_field: null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(launch_scale_by_field((int)c->V, field, x, y, c->stream));
  return GNX_OK;
}

}  // extern "C"
