// gnx_hmc.hip -- device-resident sigma-HMC trajectory (LFMN2VF integrator),
// chiral-condensate estimators and sigma observables.  The host sequences
// kernels on the context stream; every field stays in HBM for the whole
// trajectory; only host draws go up and scalars come down.
#include <complex>
#include <string.h>

#include <vector>

#include "gnx_internal.h"

using namespace gnx;

struct gnx_hmc {
  gnx_ctx* c = nullptr;
  double beta_s = 0.0;
  int nflav = 4;
  double* sigma = nullptr;      // V
  double* sm = nullptr;         // V
  double* sigma_copy = nullptr; // V
  double* mom = nullptr;        // V
  double* force = nullptr;      // V
  double* host_stage = nullptr; // V pinned staging for draws
  double2* chi = nullptr;       // [2][V]
  double2* phi = nullptr;       // [2][V]
  double2* X = nullptr;         // [2][V]  inverted_phi
  double2* Y = nullptr;         // [2][V]  inverted_phi_obs
  double2* vec = nullptr;       // V       condensate source
  double2* ddag = nullptr;      // V       Ddag_inverted_phi_obs
  double2* w = nullptr;         // V       scratch
  double2* phase_t = nullptr;   // LT      exp(i pi x2/L)/V source table
  double2* box_phase = nullptr; // 3*E     exp(i 2pi x_d/L) tables, E = max(L, LT)
  double iterations = 0.0;      // kit->iterations
  int calls = 0;                // kit->CG_call
  int force_prec = GNX_PREC_F64;  // arithmetic of the MEDIUM-tolerance force solves
};

namespace {

__global__ void k_real_to_complex(int V, const double* __restrict__ re, double2* __restrict__ out) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < V) out[f] = make_double2(re[f], 0.0);
}

__global__ void k_fill(int V, double2 v, double2* __restrict__ out) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < V) out[f] = v;
}

// v[f] = table[x2(f)]  (FullVecPhasesForChipOverVolume, condensate.c:84-89, d = 2)
__global__ void k_time_phase(Geom g, const double2* __restrict__ table, double2* __restrict__ out) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  Site s = decode(g, f);
  out[f] = table[s.x2];
}

// CondensateSigmaPi partial sums (observables.c:335-355): per block the 7
// sums  sum temp, sum Re/Im(phase_d[x_d] temp) for d = 0,1,2, temp = sigma/V.
template <int BS>
__global__ __launch_bounds__(BS) void k_condsig_partial(Geom g, const double* __restrict__ sigma,
                                                        const double2* __restrict__ box_phase, int stride,
                                                        double* __restrict__ part) {
  __shared__ double lds[BS / 64];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const double invV = (double)g.V;
  for (int f = blockIdx.x * BS + threadIdx.x; f < g.V; f += gridDim.x * BS) {
    Site s = decode(g, f);
    double temp = sigma[f] / invV;
    acc[0] += temp;
    const int c[3] = {s.x0, s.x1, s.x2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double2 ph = box_phase[d * stride + c[d]];
      acc[1 + 2 * d] += ph.x * temp;
      acc[2 + 2 * d] += ph.y * temp;
    }
  }
#pragma unroll
  for (int k = 0; k < 7; k++) {
    double s = block_sum(acc[k], lds);
    if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = s;
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_condsig_final(int nparts, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  __shared__ double lds[BS / 64];
  for (int k = 0; k < 7; k++) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += BS) acc += part[k * nparts + i];
    double s = block_sum(acc, lds);
    if (threadIdx.x == 0) out[k] = s;
  }
}

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

gnx_status hmc_cg(gnx_hmc* h, const double2* b, double2* x, double tol, long max_iter, int start, int nrhs,
                  bool* ok, int prec = GNX_PREC_F64) {
  gnx_ctx* c = h->c;
  CGDevResult r[2];
  if (prec == GNX_PREC_F64)
    HIP_TRY(cg_run(c, h->sm, b, x, tol, max_iter, start, nrhs, r));
  else
    HIP_TRY(cgmix_run(&c->cgmix, c->g, h->sm, b, x, tol, reliable_delta_of(c, prec), max_iter, start, prec, nrhs,
                      c->stream, r));
  *ok = true;
  for (int k = 0; k < nrhs; k++) {
    if (r[k].converged && !r[k].early_exit) {
      h->iterations += r[k].iterations;
      h->calls += 1;
    }
    if (!r[k].converged) *ok = false;
  }
  return GNX_OK;
}

// the force from the solutions X (moleculardynamics.c:601-619): Y = M^dag X
// per flavour, then the corner sums and the sigma term
gnx_status hmc_force_post(gnx_hmc* h) {
  gnx_ctx* c = h->c;
  const long V = c->V;
  for (int f = 0; f < 2; f++)
    HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)(h->X + f * V), (double*)(h->Y + f * V),
                         c->tmp, c->stream));
  HIP_TRY(launch_sigma_force(c->g, h->sigma, h->beta_s, h->nflav, (const double*)h->X, (const double*)h->Y,
                             (const double*)(h->X + V), (const double*)(h->Y + V), h->force, c->stream));
  return GNX_OK;
}

// ComputeSigmaForce (moleculardynamics.c:631-647) on the device state; the
// solves in prec (FP64 unless the MEDIUM-tolerance solves are mixed)
gnx_status hmc_force(gnx_hmc* h, double tol, long max_iter, bool* ok, int prec = GNX_PREC_F64) {
  gnx_status s = hmc_cg(h, h->phi, h->X, tol, max_iter, GNX_START_NULL, 2, ok, prec);
  if (s != GNX_OK || !*ok) return s;
  return hmc_force_post(h);
}


gnx_status hmc_energy(gnx_hmc* h, double* H) {
  gnx_ctx* c = h->c;
  const long V = c->V;
  const double* phi[2] = {(const double*)h->phi, (const double*)(h->phi + V)};
  const double* X[2] = {(const double*)h->X, (const double*)(h->X + V)};
  return gnx_energy(c, h->mom, h->sigma, h->beta_s, h->nflav, phi, X, H);
}

gnx_status restore_sigma(gnx_hmc* h) {
  gnx_ctx* c = h->c;
  HIP_TRY(hipMemcpyAsync(h->sigma, h->sigma_copy, sizeof(double) * c->V, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(launch_sigma_mass(c->g, h->sigma, h->sm, c->stream));
  return GNX_OK;
}

gnx_status upload_real(gnx_hmc* h, const double* host, double* dev) {
  gnx_ctx* c = h->c;
  memcpy(h->host_stage, host, sizeof(double) * c->V);
  HIP_TRY(hipMemcpyAsync(dev, h->host_stage, sizeof(double) * c->V, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

gnx_status upload_real_as_complex(gnx_hmc* h, const double* host, double2* dev) {
  gnx_ctx* c = h->c;
  gnx_status s = upload_real(h, host, (double*)h->w);  // stage through w (V doubles fit in 2V)
  if (s != GNX_OK) return s;
  hipLaunchKernelGGL(k_real_to_complex, dim3(nblk(c->V, 256)), dim3(256), 0, c->stream, (int)c->V,
                     (const double*)h->w, dev);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status dot_re_im(gnx_hmc* h, const double2* a, const double2* b, double* re, double* im) {
  return gnx_dot(h->c, (const double*)a, (const double*)b, re, im);
}

}  // namespace

extern "C" {

gnx_status gnx_hmc_create(gnx_ctx* c, double beta_s, int nflav, gnx_hmc** out) {
  CHECK_ARG(c && out, "gnx_hmc_create: null argument");
  CHECK_ARG(nflav == 4, "gnx_hmc_create: the pseudo-fermion sector has HALFNFLAV = 2 (nflav must be 4)");
  HIP_TRY(hipSetDevice(c->device));
  gnx_hmc* h = new gnx_hmc();
  h->c = c;
  h->beta_s = beta_s;
  h->nflav = nflav;
  const long V = c->V;
  hipError_t e = hipSuccess;
  auto A = [&](void** p, size_t n) {
    if (e == hipSuccess) e = hipMalloc(p, n);
  };
  A((void**)&h->sigma, sizeof(double) * V);
  A((void**)&h->sm, sizeof(double) * V);
  A((void**)&h->sigma_copy, sizeof(double) * V);
  A((void**)&h->mom, sizeof(double) * V);
  A((void**)&h->force, sizeof(double) * V);
  A((void**)&h->chi, sizeof(double2) * 2 * V);
  A((void**)&h->phi, sizeof(double2) * 2 * V);
  A((void**)&h->X, sizeof(double2) * 2 * V);
  A((void**)&h->Y, sizeof(double2) * 2 * V);
  A((void**)&h->vec, sizeof(double2) * V);
  A((void**)&h->ddag, sizeof(double2) * V);
  A((void**)&h->w, sizeof(double2) * V);
  A((void**)&h->phase_t, sizeof(double2) * c->LT);
  const int E = c->L > c->LT ? c->L : c->LT;
  A((void**)&h->box_phase, sizeof(double2) * 3 * E);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->host_stage, sizeof(double) * V, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemsetAsync(h->sigma, 0, sizeof(double) * V, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->X, 0, sizeof(double2) * 2 * V, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->Y, 0, sizeof(double2) * 2 * V, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->mom, 0, sizeof(double) * V, c->stream);
  if (e != hipSuccess) {
    set_err("gnx_hmc_create: %s", hipGetErrorString(e));
    gnx_hmc_destroy(h);
    return GNX_ERR_HIP;
  }
  // source tables with the reference's expressions (condensate.c:88,
  // observables.c:351; P_MIN_APBC = PI/L, P_MIN = 2 PI/L, macro.h:178-179)
  const double PI = 3.141592653589793238462643383279502884197169399375105820974944;
  std::complex<double> tab[4096];
  int LT = c->LT, L = c->L;
  if (3 * E > 4096) {
    gnx_hmc_destroy(h);
    set_err("gnx_hmc_create: LT too large for phase tables");
    return GNX_ERR_SIZE;
  }
  for (int x = 0; x < LT; x++)
    tab[x] = std::exp(std::complex<double>(0.0, ((PI) / L) * x)) / (double)V;
  e = hipMemcpy(h->phase_t, tab, sizeof(double2) * LT, hipMemcpyHostToDevice);
  for (int d = 0; d < 3; d++)
    for (int x = 0; x < E; x++) tab[d * E + x] = std::exp(std::complex<double>(0.0, ((2 * PI) / L) * x));
  if (e == hipSuccess) e = hipMemcpy(h->box_phase, tab, sizeof(double2) * 3 * E, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    set_err("gnx_hmc_create: %s", hipGetErrorString(e));
    gnx_hmc_destroy(h);
    return GNX_ERR_HIP;
  }
  *out = h;
  return GNX_OK;
}

gnx_status gnx_hmc_destroy(gnx_hmc* h) {
  if (!h) return GNX_OK;
  (void)hipSetDevice(h->c->device);
  void* bufs[] = {h->sigma, h->sm, h->sigma_copy, h->mom, h->force, h->chi, h->phi, h->X, h->Y,
                  h->vec, h->ddag, h->w, h->phase_t, h->box_phase};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h->host_stage) (void)hipHostFree(h->host_stage);
  delete h;
  return GNX_OK;
}

gnx_status gnx_hmc_set_sigma(gnx_hmc* h, const double* sigma_host) {
  CHECK_ARG(h && sigma_host, "gnx_hmc_set_sigma: null argument");
  HIP_TRY(hipSetDevice(h->c->device));
  gnx_status s = upload_real(h, sigma_host, h->sigma);
  if (s != GNX_OK) return s;
  HIP_TRY(launch_sigma_mass(h->c->g, h->sigma, h->sm, h->c->stream));
  return GNX_OK;
}

gnx_status gnx_hmc_get_sigma(gnx_hmc* h, double* sigma_host, double* sm_host) {
  CHECK_ARG(h, "gnx_hmc_get_sigma: null argument");
  HIP_TRY(hipSetDevice(h->c->device));
  const long V = h->c->V;
  if (sigma_host) HIP_TRY(hipMemcpyAsync(sigma_host, h->sigma, sizeof(double) * V, hipMemcpyDeviceToHost, h->c->stream));
  if (sm_host) HIP_TRY(hipMemcpyAsync(sm_host, h->sm, sizeof(double) * V, hipMemcpyDeviceToHost, h->c->stream));
  HIP_TRY(hipStreamSynchronize(h->c->stream));
  return GNX_OK;
}

gnx_status gnx_hmc_device_fields(gnx_hmc* h, double** sigma, double** sm, double** mom, double** force, double** X,
                                 double** Y, double** phi) {
  CHECK_ARG(h, "gnx_hmc_device_fields: null argument");
  if (sigma) *sigma = h->sigma;
  if (sm) *sm = h->sm;
  if (mom) *mom = h->mom;
  if (force) *force = h->force;
  if (X) *X = (double*)h->X;
  if (Y) *Y = (double*)h->Y;
  if (phi) *phi = (double*)h->phi;
  return GNX_OK;
}

gnx_status gnx_hmc_trajectory(gnx_hmc* h, const double* mom_host, const double* chi0_host, const double* chi1_host,
                              double epsilon, int step, double tol_high, double tol_medium, long max_iter,
                              gnx_traj_result* res) {
  CHECK_ARG(h && mom_host && chi0_host && chi1_host && res, "gnx_hmc_trajectory: null argument");
  CHECK_ARG(step >= 2, "gnx_hmc_trajectory: the integrator needs at least 2 steps (info.c:66)");
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  const long V = c->V;
  if (max_iter <= 0) max_iter = V + 1;
  memset(res, 0, sizeof(*res));
  const double it0 = h->iterations;
  const int calls0 = h->calls;
  const double lam = 0.1931833275037836;  // MAGIC_LAMBDA, macro.h:169
  gnx_status s;
  bool ok = true;

  if ((s = upload_real(h, mom_host, h->mom)) != GNX_OK) return s;
  const double* chis[2] = {chi0_host, chi1_host};
  for (int f = 0; f < 2; f++) {
    if ((s = upload_real_as_complex(h, chis[f], h->chi + f * V)) != GNX_OK) return s;
    HIP_TRY(launch_dirac(c->g, GNX_OP_M, h->sm, (const double*)(h->chi + f * V), (double*)(h->phi + f * V), c->tmp,
                         c->stream));
  }
  HIP_TRY(hipMemcpyAsync(h->sigma_copy, h->sigma, sizeof(double) * V, hipMemcpyDeviceToDevice, c->stream));

#define TRY_CG(call)                                  \
  do {                                                \
    if ((s = (call)) != GNX_OK) return s;             \
    if (!ok) {                                        \
      if ((s = restore_sigma(h)) != GNX_OK) return s; \
      res->ok = 0;                                    \
      set_err("gnx_hmc_trajectory: CG did not converge"); \
      return GNX_ERR_NOCONV;                          \
    }                                                 \
  } while (0)
#define P_STEP(coef) HIP_TRY(launch_evolve_momenta(c->g, h->mom, h->force, (coef), c->stream))
#define Q_STEP(coef) HIP_TRY(launch_evolve_sigma(c->g, h->sigma, h->sm, h->mom, (coef), c->stream))

  TRY_CG(hmc_cg(h, h->phi, h->X, tol_high, max_iter, GNX_START_IMPOSED, 2, &ok));
  double H0 = 0.0, H1 = 0.0;
  if ((s = hmc_energy(h, &H0)) != GNX_OK) return s;

  TRY_CG(hmc_force(h, tol_medium, max_iter, &ok, h->force_prec));
  P_STEP((lam * epsilon));
  Q_STEP((epsilon / 2.0));
  TRY_CG(hmc_force(h, tol_medium, max_iter, &ok, h->force_prec));
  P_STEP(((1.0 - (2.0 * lam)) * epsilon));
  Q_STEP((epsilon / 2.0));
  for (int st = 1; st < step - 1; st++) {
    TRY_CG(hmc_force(h, tol_medium, max_iter, &ok, h->force_prec));
    P_STEP((2.0 * lam * epsilon));
    Q_STEP((epsilon / 2.0));
    TRY_CG(hmc_force(h, tol_medium, max_iter, &ok, h->force_prec));
    P_STEP(((1.0 - (2.0 * lam)) * epsilon));
    Q_STEP((epsilon / 2.0));
  }
  TRY_CG(hmc_force(h, tol_medium, max_iter, &ok, h->force_prec));
  P_STEP((2.0 * lam * epsilon));
  Q_STEP((epsilon / 2.0));
  TRY_CG(hmc_force(h, tol_medium, max_iter, &ok, h->force_prec));
  P_STEP(((1.0 - (2.0 * lam)) * epsilon));
  Q_STEP((epsilon / 2.0));
  TRY_CG(hmc_force(h, tol_high, max_iter, &ok));
  P_STEP((lam * epsilon));
#undef TRY_CG
#undef P_STEP
#undef Q_STEP

  if ((s = hmc_energy(h, &H1)) != GNX_OK) return s;
  res->H0 = H0;
  res->H1 = H1;
  res->dH = H1 - H0;
  res->ok = 1;
  res->cg_solves = h->calls - calls0;
  res->cg_iterations = (long)(h->iterations - it0);
  return GNX_OK;
}

gnx_status gnx_hmc_set_force_precision(gnx_hmc* h, gnx_prec prec) {
  CHECK_ARG(h, "gnx_hmc_set_force_precision: null argument");
  CHECK_ARG(prec == GNX_PREC_F64 || prec == GNX_PREC_F32 || prec == GNX_PREC_BF16 || prec == GNX_PREC_F16,
            "gnx_hmc_set_force_precision: bad precision");
  h->force_prec = (int)prec;
  return GNX_OK;
}

gnx_status gnx_hmc_accept(gnx_hmc* h, int accepted) {
  CHECK_ARG(h, "gnx_hmc_accept: null argument");
  HIP_TRY(hipSetDevice(h->c->device));
  if (!accepted) return restore_sigma(h);
  return GNX_OK;
}

// n chains in lockstep (moleculardynamics.c:1234-1434 for each): the steps
// of gnx_hmc_trajectory, chain by chain on each chain's stream, except that
// every solve point of the integrator is one uneven-batch solve over all
// live chains (systems 2k, 2k+1 of m: chain k's two flavours).  The
// MEDIUM-tolerance force solves run in the chains' force precision
// (gnx_hmc_set_force_precision, the same on every chain; mixed precision
// with the first chain's reliable-update factor), the rest in FP64.  A chain
// whose solve fails leaves the lockstep as gnx_hmc_trajectory would (sigma
// restored, ok = 0); the others go on.
gnx_status gnx_hmc_ensemble_trajectory(gnx_multi* m, int n, gnx_hmc* const* hs, const double* const* mom_host,
                                       const double* const* chi0_host, const double* const* chi1_host,
                                       double epsilon, int step, double tol_high, double tol_medium,
                                       long max_iter, gnx_traj_result* res) {
  CHECK_ARG(m && hs && mom_host && chi0_host && chi1_host && res && n >= 1,
            "gnx_hmc_ensemble_trajectory: null argument");
  CHECK_ARG(step >= 2, "gnx_hmc_ensemble_trajectory: the integrator needs at least 2 steps (info.c:66)");
  CHECK_ARG(multi_count(m) == 2 * n, "gnx_hmc_ensemble_trajectory: the solver must hold two systems per chain");
  for (int k = 0; k < n; k++) {
    CHECK_ARG(hs[k] && mom_host[k] && chi0_host[k] && chi1_host[k], "gnx_hmc_ensemble_trajectory: null argument");
    CHECK_ARG(hs[k]->force_prec == hs[0]->force_prec,
              "gnx_hmc_ensemble_trajectory: the chains share their solves, so all need the same force precision");
    for (int f = 0; f < 2; f++) {
      int L = 0, LT = 0, dev = -1;
      multi_system(m, 2 * k + f, &L, &LT, &dev);
      CHECK_ARG(L == hs[k]->c->g.L && LT == hs[k]->c->g.LT && dev == hs[k]->c->device,
                "gnx_hmc_ensemble_trajectory: system 2k, 2k+1 of the solver must have chain k's lattice and device");
    }
  }
  HIP_TRY(hipSetDevice(hs[0]->c->device));
  const double lam = 0.1931833275037836;  // MAGIC_LAMBDA, macro.h:169
  std::vector<char> live(n, 1), ok(n, 1);
  std::vector<double> it0(n), H0(n, 0.0);
  std::vector<int> calls0(n);
  gnx_status s;
  for (int k = 0; k < n; k++) {
    gnx_hmc* h = hs[k];
    gnx_ctx* c = h->c;
    const long V = c->V;
    memset(&res[k], 0, sizeof(res[k]));
    it0[k] = h->iterations;
    calls0[k] = h->calls;
    if ((s = upload_real(h, mom_host[k], h->mom)) != GNX_OK) return s;
    const double* chis[2] = {chi0_host[k], chi1_host[k]};
    for (int f = 0; f < 2; f++) {
      if ((s = upload_real_as_complex(h, chis[f], h->chi + f * V)) != GNX_OK) return s;
      HIP_TRY(launch_dirac(c->g, GNX_OP_M, h->sm, (const double*)(h->chi + f * V), (double*)(h->phi + f * V),
                           c->tmp, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(h->sigma_copy, h->sigma, sizeof(double) * V, hipMemcpyDeviceToDevice, c->stream));
  }
  // a failed solve: the chain leaves as gnx_hmc_trajectory's TRY_CG does
  const int fprec = hs[0]->force_prec;
  const double fdelta = fprec == GNX_PREC_F64 ? 0.0 : reliable_delta_of(hs[0]->c, fprec);
  auto solve = [&](double tol, int start, int prec = GNX_PREC_F64) -> gnx_status {
    gnx_status q = ens_solve(m, n, hs, live, tol, max_iter, start, ok, nullptr, prec, fdelta);
    if (q != GNX_OK) return q;
    for (int k = 0; k < n; k++)
      if (live[k] && !ok[k]) {
        live[k] = 0;
        if ((q = restore_sigma(hs[k])) != GNX_OK) return q;
        res[k].ok = 0;
      }
    return GNX_OK;
  };
  auto force = [&](double tol, int prec = GNX_PREC_F64) -> gnx_status {
    gnx_status q = solve(tol, GNX_START_NULL, prec);
    for (int k = 0; k < n && q == GNX_OK; k++)
      if (live[k]) q = hmc_force_post(hs[k]);
    return q;
  };
  auto pq = [&](double pc, double qc) -> gnx_status {
    for (int k = 0; k < n; k++) {
      if (!live[k]) continue;
      gnx_ctx* c = hs[k]->c;
      HIP_TRY(launch_evolve_momenta(c->g, hs[k]->mom, hs[k]->force, pc, c->stream));
      if (qc != 0.0) HIP_TRY(launch_evolve_sigma(c->g, hs[k]->sigma, hs[k]->sm, hs[k]->mom, qc, c->stream));
    }
    return GNX_OK;
  };
#define ENS_TRY(call)                         \
  do {                                        \
    if ((s = (call)) != GNX_OK) return s;     \
  } while (0)
  ENS_TRY(solve(tol_high, GNX_START_IMPOSED));
  for (int k = 0; k < n; k++)
    if (live[k]) ENS_TRY(hmc_energy(hs[k], &H0[k]));
  ENS_TRY(force(tol_medium, fprec));
  ENS_TRY(pq(lam * epsilon, epsilon / 2.0));
  ENS_TRY(force(tol_medium, fprec));
  ENS_TRY(pq((1.0 - (2.0 * lam)) * epsilon, epsilon / 2.0));
  for (int st = 1; st < step - 1; st++) {
    ENS_TRY(force(tol_medium, fprec));
    ENS_TRY(pq(2.0 * lam * epsilon, epsilon / 2.0));
    ENS_TRY(force(tol_medium, fprec));
    ENS_TRY(pq((1.0 - (2.0 * lam)) * epsilon, epsilon / 2.0));
  }
  ENS_TRY(force(tol_medium, fprec));
  ENS_TRY(pq(2.0 * lam * epsilon, epsilon / 2.0));
  ENS_TRY(force(tol_medium, fprec));
  ENS_TRY(pq((1.0 - (2.0 * lam)) * epsilon, epsilon / 2.0));
  ENS_TRY(force(tol_high));
  ENS_TRY(pq(lam * epsilon, 0.0));
  bool all = true;
  for (int k = 0; k < n; k++) {
    if (!live[k]) {
      all = false;
      continue;
    }
    double H1 = 0.0;
    ENS_TRY(hmc_energy(hs[k], &H1));
    res[k].H0 = H0[k];
    res[k].H1 = H1;
    res[k].dH = H1 - H0[k];
    res[k].ok = 1;
    res[k].cg_solves = hs[k]->calls - calls0[k];
    res[k].cg_iterations = (long)(hs[k]->iterations - it0[k]);
  }
#undef ENS_TRY
  if (!all) {
    set_err("gnx_hmc_ensemble_trajectory: CG did not converge on a chain (its ok = 0)");
    return GNX_ERR_NOCONV;
  }
  return GNX_OK;
}

gnx_status gnx_hmc_condensate(gnx_hmc* h, const double* z2_host, double tol, long max_iter, double* out) {
  CHECK_ARG(h && z2_host && out, "gnx_hmc_condensate: null argument");
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  const long V = c->V;
  if (max_iter <= 0) max_iter = V + 1;
  const double Vd = (double)V;
  double2* x = h->Y;  // inverted_phi_obs[0] aliasing, observables.c:247
  gnx_status s;
  bool ok;
  double re, im;
  double temp[4];
  for (int p = 0; p < 4; p++) {
    if ((s = upload_real_as_complex(h, z2_host + p * V, h->vec)) != GNX_OK) return s;
    if ((s = hmc_cg(h, h->vec, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
    HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
    if ((s = dot_re_im(h, h->vec, h->ddag, &re, &im)) != GNX_OK) return s;
    temp[p] = re / Vd;
  }
  double t0 = temp[0], t1 = temp[1], t2 = temp[2], t3 = temp[3];
  out[0] = (t0 + t1 + t2 + t3) / 4.0;
  out[1] = fabs(out[0]);
  out[2] = ((t0 * t1) + (t0 * t2) + (t0 * t3) + (t1 * t2) + (t1 * t3) + (t2 * t3)) / 6.0;
  out[3] = (t0 * t1 * t2 * t3);

  hipLaunchKernelGGL(k_fill, dim3(nblk(V, 256)), dim3(256), 0, c->stream, (int)V, make_double2(1.0 / V, 0.0),
                     h->vec);
  if ((s = hmc_cg(h, h->vec, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
  HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
  if ((s = dot_re_im(h, h->vec, h->ddag, &re, &im)) != GNX_OK) return s;
  double mu2 = re;
  out[4] = mu2;
  out[5] = mu2 * mu2;
  out[6] = V * mu2;
  out[7] = fabs(out[6]);

  hipLaunchKernelGGL(k_time_phase, dim3(nblk(V, 256)), dim3(256), 0, c->stream, c->g,
                     (const double2*)h->phase_t, h->vec);
  if ((s = hmc_cg(h, h->vec, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
  HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
  if ((s = dot_re_im(h, h->vec, h->ddag, &re, &im)) != GNX_OK) return s;
  out[8] = im * V;

  // M^-2 on Z2 noise (condensate.c:201-212)
  if ((s = upload_real_as_complex(h, z2_host + 4 * V, h->vec)) != GNX_OK) return s;
  if ((s = hmc_cg(h, h->vec, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
  HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
  if ((s = hmc_cg(h, h->ddag, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
  HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
  if ((s = dot_re_im(h, h->vec, h->ddag, &re, &im)) != GNX_OK) return s;
  out[9] = re / Vd;

  // sigma-weighted M^-2 (condensate.c:216-230)
  if ((s = upload_real_as_complex(h, z2_host + 5 * V, h->vec)) != GNX_OK) return s;
  HIP_TRY(launch_scale_by_field((int)V, h->sm, (const double*)h->vec, (double*)h->w, c->stream));
  if ((s = hmc_cg(h, h->w, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
  HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
  HIP_TRY(launch_scale_by_field((int)V, h->sm, (const double*)h->ddag, (double*)h->w, c->stream));
  if ((s = hmc_cg(h, h->w, x, tol, max_iter, GNX_START_NULL, 1, &ok)) != GNX_OK) return s;
  HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)x, (double*)h->ddag, c->tmp, c->stream));
  if ((s = dot_re_im(h, h->vec, h->ddag, &re, &im)) != GNX_OK) return s;
  out[10] = re / Vd;
  return GNX_OK;
}

gnx_status gnx_hmc_condensate_sigma(gnx_hmc* h, double* out) {
  CHECK_ARG(h && out, "gnx_hmc_condensate_sigma: null argument");
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  const long V = c->V;
  int nb = nblk(V, 256);
  if (nb > c->red_nparts / 7) nb = c->red_nparts / 7;
  hipLaunchKernelGGL(k_condsig_partial<256>, dim3(nb), dim3(256), 0, c->stream, c->g, (const double*)h->sigma,
                     (const double2*)h->box_phase, c->L > c->LT ? c->L : c->LT, c->red_part);
  hipLaunchKernelGGL(k_condsig_final<256>, dim3(1), dim3(256), 0, c->stream, nb, (const double*)c->red_part,
                     c->red_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(c->host_out, c->red_out, sizeof(double) * 7, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const double* o = c->host_out;
  double sg = o[0];
  double mu2 = sg * sg;
  out[0] = sg;
  out[1] = fabs(sg);
  out[2] = mu2;
  out[3] = mu2 * mu2;
  out[4] = V * mu2;
  out[5] = V * (o[5] * o[5] + o[6] * o[6]);
  out[6] = V * (o[1] * o[1] + o[2] * o[2]);
  out[7] = V * (o[3] * o[3] + o[4] * o[4]);
  return GNX_OK;
}

gnx_status gnx_hmc_cg_stats(gnx_hmc* h, double* iterations, int* calls) {
  CHECK_ARG(h, "gnx_hmc_cg_stats: null argument");
  if (iterations) *iterations = h->iterations;
  if (calls) *calls = h->calls;
  return GNX_OK;
}

}  // extern "C"
