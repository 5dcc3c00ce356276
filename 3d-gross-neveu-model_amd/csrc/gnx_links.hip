// gnx_links.hip -- the U(1) link sector: the link-coupled operator of the
// reference's commented-out copy_dirac.c:46-123, the link fermion force
// (moleculardynamics.c:164-193), the pure-gauge force and plaquette action
// (linklattice.c:16-172, moleculardynamics.c:127-138), the link energy
// (:268-294) and the device-resident LeapFrog2Link trajectory (:419-542).
//
// Layouts: links on the device are f-order, three per site (U[3 f + d] =
// U_d(f)); the reference keeps them in lex order (lattice.h:15), so the host
// interfaces convert.  Momenta and forces are f-order [V][3], as in the
// reference (moleculardynamics.c:389).
//
// The reference's live operator (dirac.c) ignores the links, so its
// LeapFrog2Link solves with M while its force differentiates M_U.  The
// trajectory here takes `linked`: 0 reproduces that (the drop-in), 1 solves
// with M_U (the consistent link-coupled theory).
#include <string.h>

#include "gnx_internal.h"

using namespace gnx;

namespace {

constexpr int BS = 256;
inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

__global__ void k_links_lex_to_f(Geom g, const double2* __restrict__ ulex, double2* __restrict__ uf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  const Site s = decode(g, f);
#pragma unroll
  for (int d = 0; d < 3; d++) uf[3 * f + d] = ulex[3 * s.lex + d];
}

__global__ void k_links_f_to_lex(Geom g, const double2* __restrict__ uf, double2* __restrict__ ulex) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  const Site s = decode(g, f);
#pragma unroll
  for (int d = 0; d < 3; d++) ulex[3 * s.lex + d] = uf[3 * f + d];
}

// op 0: out = M_U in, 1: out = M_U^+ in, 2: out -= M_U in
__global__ void k_link_dirac(Geom g, int op, const double* __restrict__ sm, const double2* __restrict__ U,
                             const double2* __restrict__ in, double2* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  const double2 r = op == 2 ? out[f] : make_double2(0.0, 0.0);
  out[f] = apply_link(g, decode(g, f), op, sm[f], r, U, [&](int q) { return in[q]; });
}

// f-index of the site at coordinates c shifted by +-1 along d
GNX_DEV int f_shift(const Geom& g, int c0, int c1, int c2, int d, int step) {
  int c[3] = {c0, c1, c2};
  const int E = d == 2 ? g.LT : g.L;
  c[d] += step;
  if (c[d] == E) c[d] = 0;
  if (c[d] < 0) c[d] = E - 1;
  return f_of_coords(g, c[0], c[1], c[2]);
}

// AddForceFromFermionSectorRespectToLink (moleculardynamics.c:181-190) after
// the solves, per f-site and direction, flavour-major accumulation:
//   F += Re(conj(X[x]) i eta_pp(x,mu) U_mu(x) Y[x+mu])
//      + Re(conj(X[x+mu]) i eta_mm(x+mu,mu) conj(U_mu(x)) Y[x])
// eta_mm(x+mu, mu) equals eta_pp(x, mu) (eta_mu does not depend on x_mu and
// the antiperiodic flips match), so both use 2 * half_eta.pp.
__global__ void k_link_force(Geom g, const double2* __restrict__ U, const double2* __restrict__ X0,
                             const double2* __restrict__ Y0, const double2* __restrict__ X1,
                             const double2* __restrict__ Y1, double* __restrict__ force, int accumulate) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  const Site s = decode(g, f);
  const NbrF n = neighbours_f(g, s);
  const HalfEta e = half_eta(g, s);
  const double2* Xs[2] = {X0, X1};
  const double2* Ys[2] = {Y0, Y1};
#pragma unroll
  for (int mu = 0; mu < 3; mu++) {
    const int fp = n.pp[mu];
    const double eta = 2.0 * e.pp[mu];
    const double2 u = U[3 * f + mu];
    double F = accumulate ? force[3 * f + mu] : 0.0;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const double2 xa = Xs[k][f], xb = Xs[k][fp];
      // conj(X) * i = (Im X, Re X)
      double2 t = rmul(eta, make_double2(xa.y, xa.x));
      t = cmul(cmul(t, u), Ys[k][fp]);
      F += t.x;
      t = rmul(eta, make_double2(xb.y, xb.x));
      t = cmul(cmul(t, cconj(u)), Ys[k][f]);
      F += t.x;
    }
    force[3 * f + mu] = F;
  }
}

// ComputeImagAllPlaquettesIncludingLink (linklattice.c:146-172) times beta,
// written to the f-order force (moleculardynamics.c:127-138)
__global__ void k_gauge_force(Geom g, const double2* __restrict__ U, double beta, double* __restrict__ force) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  const Site s = decode(g, f);
#pragma unroll
  for (int mu = 0; mu < 3; mu++) {
    double total = 0.0;
    const int fxm = f_shift(g, s.x0, s.x1, s.x2, mu, 1);
#pragma unroll
    for (int nu = 0; nu < 3; nu++) {
      if (nu == mu) continue;
      const int fxn = f_shift(g, s.x0, s.x1, s.x2, nu, 1);
      double2 t1 = cmul(U[3 * f + mu], U[3 * fxm + nu]);
      double2 t2 = cmul(t1, cconj(U[3 * fxn + mu]));
      t1 = cmul(t2, cconj(U[3 * f + nu]));
      total -= t1.y;
      const int fy = f_shift(g, s.x0, s.x1, s.x2, nu, -1);
      const Site sy = decode(g, fy);
      const int fym = f_shift(g, sy.x0, sy.x1, sy.x2, mu, 1);
      t1 = cmul(U[3 * fy + mu], U[3 * fym + nu]);
      t2 = cmul(t1, cconj(U[3 * f + mu]));
      t1 = cmul(t2, cconj(U[3 * fy + nu]));
      total += t1.y;
    }
    force[3 * f + mu] = beta * total;
  }
}

// per-block partial sums of Re Pi_{mu nu}(x), mu > nu (linklattice.c:16-33, 62-80)
__global__ __launch_bounds__(BS) void k_plaq_partial(Geom g, const double2* __restrict__ U,
                                                     double* __restrict__ part) {
  __shared__ double lds[BS / 64];
  double acc = 0.0;
  for (int f = blockIdx.x * BS + threadIdx.x; f < g.V; f += gridDim.x * BS) {
    const Site s = decode(g, f);
#pragma unroll
    for (int mu = 1; mu < 3; mu++)
#pragma unroll
      for (int nu = 0; nu < mu; nu++) {
        const int fxm = f_shift(g, s.x0, s.x1, s.x2, mu, 1);
        const int fxn = f_shift(g, s.x0, s.x1, s.x2, nu, 1);
        double2 t1 = cmul(U[3 * f + mu], U[3 * fxm + nu]);
        double2 t2 = cmul(t1, cconj(U[3 * fxn + mu]));
        t1 = cmul(t2, cconj(U[3 * f + nu]));
        acc += t1.x;
      }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(BS) void k_sum_final(int n, const double* __restrict__ part, double scale,
                                                  double* __restrict__ out) {
  __shared__ double lds[BS / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += BS) acc += part[i];
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) *out = acc * scale;
}

// EvolveCoordinates (moleculardynamics.c:381-391), Q_CHARGE = 1:
// U = exp(i (arg U + eps P))
__global__ void k_link_evolve(int n, double2* __restrict__ U, const double* __restrict__ mom, double eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 u = U[i];
  const double th = atan2(u.y, u.x) + eps * mom[i];
  double sn, cs;
  sincos(th, &sn, &cs);
  U[i] = make_double2(cs, sn);
}

// EvolveMomenta (moleculardynamics.c:400-407)
__global__ void k_mom_evolve(int n, double* __restrict__ mom, const double* __restrict__ force, double eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mom[i] += force[i] * eps;
}

__global__ void k_real_to_complex_l(int V, const double* __restrict__ re, double2* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < V) out[f] = make_double2(re[f], 0.0);
}

hipError_t link_dirac(gnx_ctx* c, int op, const double* sm, const double2* U, const double2* in, double2* out) {
  hipLaunchKernelGGL(k_link_dirac, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, op, sm, U, in, out);
  return hipGetLastError();
}

// scalar reductions into c->red_out[slot]
hipError_t plaq_sum_to(gnx_ctx* c, const double2* U, int slot) {
  int nb = nblk(c->V, BS);
  if (nb > c->red_nparts) nb = c->red_nparts;
  hipLaunchKernelGGL(k_plaq_partial, dim3(nb), dim3(BS), 0, c->stream, c->g, U, c->red_part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(BS), 0, c->stream, nb, c->red_part, 1.0, c->red_out + slot);
  return hipGetLastError();
}

}  // namespace

struct gnx_link_hmc {
  gnx_ctx* c = nullptr;
  double beta = 0.0;
  int pure_gauge = 1;
  int linked = 1;
  double2* U = nullptr;       // [3V] f-order
  double2* U_copy = nullptr;  // [3V]
  double2* Ulex = nullptr;    // [3V] staging in lex order
  double* sm = nullptr;       // [V]
  double* mom = nullptr;      // [3V]
  double* force = nullptr;    // [3V]
  double* host_stage = nullptr;  // [6V] pinned
  double2* chi = nullptr;     // [2][V]
  double2* phi = nullptr;     // [2][V]
  double2* X = nullptr;       // [2][V]  inverted_phi
  double2* Y = nullptr;       // [2][V]  inverted_phi_obs
  double iterations = 0.0;
  int calls = 0;
};

namespace {

gnx_status lh_upload(gnx_link_hmc* h, const double* host, void* dev, size_t doubles) {
  gnx_ctx* c = h->c;
  memcpy(h->host_stage, host, sizeof(double) * doubles);
  HIP_TRY(hipMemcpyAsync(dev, h->host_stage, sizeof(double) * doubles, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

gnx_status lh_cg(gnx_link_hmc* h, double tol, long max_iter, bool* ok) {
  gnx_ctx* c = h->c;
  CGDevResult r[2];
  if (h->linked)
    HIP_TRY(cg_run_link(c, h->sm, h->U, h->phi, h->X, tol, max_iter, GNX_START_NULL, 2, r));
  else
    HIP_TRY(cg_run(c, h->sm, h->phi, h->X, tol, max_iter, GNX_START_NULL, 2, r));
  *ok = true;
  for (int k = 0; k < 2; k++) {
    if (r[k].converged && !r[k].early_exit) {
      h->iterations += r[k].iterations;
      h->calls += 1;
    }
    if (!r[k].converged) *ok = false;
  }
  return GNX_OK;
}

// ComputeForce (moleculardynamics.c:203-222), in three parts: the gauge
// force, the solve, and the fermion force from the solutions
void lh_force_pre(gnx_link_hmc* h) {
  gnx_ctx* c = h->c;
  if (h->pure_gauge)
    hipLaunchKernelGGL(k_gauge_force, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, h->U, h->beta, h->force);
}

gnx_status lh_force_post(gnx_link_hmc* h) {
  gnx_ctx* c = h->c;
  const long V = c->V;
  for (int f = 0; f < 2; f++) {
    if (h->linked)
      HIP_TRY(link_dirac(c, 1, h->sm, h->U, h->X + f * V, h->Y + f * V));
    else
      HIP_TRY(launch_dirac(c->g, GNX_OP_MDAG, h->sm, (const double*)(h->X + f * V), (double*)(h->Y + f * V),
                           c->tmp, c->stream));
  }
  hipLaunchKernelGGL(k_link_force, dim3(nblk(V, BS)), dim3(BS), 0, c->stream, c->g, (const double2*)h->U,
                     (const double2*)h->X, (const double2*)h->Y, (const double2*)(h->X + V),
                     (const double2*)(h->Y + V), h->force, h->pure_gauge);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status lh_force(gnx_link_hmc* h, double tol, long max_iter, bool* ok) {
  lh_force_pre(h);
  gnx_status s = lh_cg(h, tol, max_iter, ok);
  if (s != GNX_OK || !*ok) return s;
  return lh_force_post(h);
}

// ComputeFullEnergyWithPhiWithoutSigma (moleculardynamics.c:268-294)
gnx_status lh_energy(gnx_link_hmc* h, double* H) {
  gnx_ctx* c = h->c;
  const long V = c->V;
  HIP_TRY(launch_reduce((int)(3 * V), 3, h->mom, nullptr, 0.5, c->red_part, c->red_nparts, c->red_out + 0,
                        c->stream));
  HIP_TRY(plaq_sum_to(c, h->U, 1));
  for (int f = 0; f < 2; f++)
    HIP_TRY(launch_reduce((int)V, 1, h->phi + f * V, h->X + f * V, 1.0, c->red_part, c->red_nparts,
                          c->red_out + 2 + f, c->stream));
  HIP_TRY(hipMemcpyAsync(c->host_out, c->red_out, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const double* o = c->host_out;
  double e = o[0];
  if (h->pure_gauge) e -= h->beta * o[1];
  e += o[2];
  e += o[3];
  *H = e;
  return GNX_OK;
}

}  // namespace

extern "C" {

gnx_status gnx_links_from_lex(gnx_ctx* c, const double* ulex, double* uf) {
  CHECK_ARG(c && ulex && uf && ulex != uf, "gnx_links_from_lex: bad argument");
  HIP_TRY(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_links_lex_to_f, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, (const double2*)ulex,
                     (double2*)uf);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status gnx_links_to_lex(gnx_ctx* c, const double* uf, double* ulex) {
  CHECK_ARG(c && ulex && uf && ulex != uf, "gnx_links_to_lex: bad argument");
  HIP_TRY(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_links_f_to_lex, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, (const double2*)uf,
                     (double2*)ulex);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status gnx_link_dirac(gnx_ctx* c, gnx_op op, const double* sm, const double* uf, const double* in, double* out) {
  CHECK_ARG(c && sm && uf && in && out && in != out, "gnx_link_dirac: bad argument");
  int k;
  switch (op) {
    case GNX_OP_M: k = 0; break;
    case GNX_OP_MDAG: k = 1; break;
    case GNX_OP_SUB_M: k = 2; break;
    default: set_err("gnx_link_dirac: op must be M, MDAG or SUB_M"); return GNX_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(link_dirac(c, k, sm, (const double2*)uf, (const double2*)in, (double2*)out));
  return GNX_OK;
}

gnx_status gnx_link_cg_batch(gnx_ctx* c, const double* sm, const double* uf, const double* b, double* x, int nrhs,
                             double tol, long max_iter, gnx_cg_start start, gnx_cg_result* res) {
  CHECK_ARG(c && sm && uf && b && x && res, "gnx_link_cg_batch: null argument");
  CHECK_ARG(nrhs >= 1 && nrhs <= 64, "gnx_link_cg_batch: nrhs must be in [1, 64]");
  CHECK_ARG(start == GNX_START_NULL || start == GNX_START_IMPOSED, "gnx_link_cg_batch: bad start mode");
  CHECK_ARG((const void*)b != (const void*)x, "gnx_link_cg_batch: b and x alias");
  if (max_iter <= 0) max_iter = c->V + 1;
  HIP_TRY(hipSetDevice(c->device));
  CGDevResult hr[64];
  HIP_TRY(cg_run_link(c, sm, (const double2*)uf, (const double2*)b, (double2*)x, tol, max_iter, (int)start, nrhs,
                      hr));
  int all = 1;
  for (int k = 0; k < nrhs; k++) {
    res[k].converged = hr[k].converged;
    res[k].early_exit = hr[k].early_exit;
    res[k].iterations = hr[k].iterations;
    res[k].reliable_updates = 0;
    res[k].residual = hr[k].residual;
    all &= hr[k].converged;
  }
  if (!all) {
    set_err("gnx_link_cg_batch: not converged within %ld iterations", max_iter);
    return GNX_ERR_NOCONV;
  }
  return GNX_OK;
}

gnx_status gnx_gauge_force(gnx_ctx* c, const double* uf, double beta, double* force) {
  CHECK_ARG(c && uf && force, "gnx_gauge_force: null argument");
  HIP_TRY(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_gauge_force, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, (const double2*)uf, beta,
                     force);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status gnx_link_force_from_solutions(gnx_ctx* c, const double* uf, const double* X0, const double* Y0,
                                         const double* X1, const double* Y1, double* force, int accumulate) {
  CHECK_ARG(c && uf && X0 && Y0 && X1 && Y1 && force, "gnx_link_force_from_solutions: null argument");
  HIP_TRY(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_link_force, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, (const double2*)uf,
                     (const double2*)X0, (const double2*)Y0, (const double2*)X1, (const double2*)Y1, force,
                     accumulate ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return GNX_OK;
}

gnx_status gnx_plaquette_sum(gnx_ctx* c, const double* uf, double* out) {
  CHECK_ARG(c && uf && out, "gnx_plaquette_sum: null argument");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(plaq_sum_to(c, (const double2*)uf, 0));
  HIP_TRY(hipMemcpyAsync(c->host_out, c->red_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *out = c->host_out[0];
  return GNX_OK;
}

gnx_status gnx_link_hmc_destroy(gnx_link_hmc* h) {
  if (!h) return GNX_OK;
  (void)hipSetDevice(h->c->device);
  void* bufs[] = {h->U, h->U_copy, h->Ulex, h->sm, h->mom, h->force, h->chi, h->phi, h->X, h->Y};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h->host_stage) (void)hipHostFree(h->host_stage);
  delete h;
  return GNX_OK;
}

gnx_status gnx_link_hmc_create(gnx_ctx* c, double beta, int pure_gauge, int linked, gnx_link_hmc** out) {
  CHECK_ARG(c && out, "gnx_link_hmc_create: null argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  gnx_link_hmc* h = new gnx_link_hmc();
  h->c = c;
  h->beta = beta;
  h->pure_gauge = pure_gauge ? 1 : 0;
  h->linked = linked ? 1 : 0;
  const size_t V = (size_t)c->V;
  hipError_t e = hipMalloc(&h->U, sizeof(double2) * 3 * V);
  if (e == hipSuccess) e = hipMalloc(&h->U_copy, sizeof(double2) * 3 * V);
  if (e == hipSuccess) e = hipMalloc(&h->Ulex, sizeof(double2) * 3 * V);
  if (e == hipSuccess) e = hipMalloc(&h->sm, sizeof(double) * V);
  if (e == hipSuccess) e = hipMalloc(&h->mom, sizeof(double) * 3 * V);
  if (e == hipSuccess) e = hipMalloc(&h->force, sizeof(double) * 3 * V);
  if (e == hipSuccess) e = hipMalloc(&h->chi, sizeof(double2) * 2 * V);
  if (e == hipSuccess) e = hipMalloc(&h->phi, sizeof(double2) * 2 * V);
  if (e == hipSuccess) e = hipMalloc(&h->X, sizeof(double2) * 2 * V);
  if (e == hipSuccess) e = hipMalloc(&h->Y, sizeof(double2) * 2 * V);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->host_stage, sizeof(double) * 6 * V, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemsetAsync(h->X, 0, sizeof(double2) * 2 * V, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->mom, 0, sizeof(double) * 3 * V, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->force, 0, sizeof(double) * 3 * V, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    set_err("gnx_link_hmc_create: %s", hipGetErrorString(e));
    gnx_link_hmc_destroy(h);
    return GNX_ERR_HIP;
  }
  *out = h;
  return GNX_OK;
}

gnx_status gnx_link_hmc_set_fields(gnx_link_hmc* h, const double* ulex_host, const double* sm_host) {
  CHECK_ARG(h, "gnx_link_hmc_set_fields: null handle");
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  if (ulex_host) {
    gnx_status s = lh_upload(h, ulex_host, h->Ulex, 6 * (size_t)c->V);
    if (s != GNX_OK) return s;
    hipLaunchKernelGGL(k_links_lex_to_f, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g,
                       (const double2*)h->Ulex, h->U);
    HIP_TRY(hipGetLastError());
  }
  if (sm_host) {
    gnx_status s = lh_upload(h, sm_host, h->sm, (size_t)c->V);
    if (s != GNX_OK) return s;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

gnx_status gnx_link_hmc_set_solutions(gnx_link_hmc* h, const double* X0_host, const double* X1_host) {
  CHECK_ARG(h && X0_host && X1_host, "gnx_link_hmc_set_solutions: null argument");
  const size_t V = (size_t)h->c->V;
  HIP_TRY(hipSetDevice(h->c->device));
  gnx_status s = lh_upload(h, X0_host, h->X, 2 * V);
  if (s == GNX_OK) s = lh_upload(h, X1_host, h->X + V, 2 * V);
  return s;
}

gnx_status gnx_link_hmc_get_links(gnx_link_hmc* h, double* ulex_host) {
  CHECK_ARG(h && ulex_host, "gnx_link_hmc_get_links: null argument");
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_links_f_to_lex, dim3(nblk(c->V, BS)), dim3(BS), 0, c->stream, c->g, (const double2*)h->U,
                     h->Ulex);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(h->host_stage, h->Ulex, sizeof(double2) * 3 * c->V, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  memcpy(ulex_host, h->host_stage, sizeof(double2) * 3 * c->V);
  return GNX_OK;
}

gnx_status gnx_link_hmc_device_fields(gnx_link_hmc* h, double** uf, double** sm, double** mom, double** force,
                                      double** X, double** Y, double** phi) {
  CHECK_ARG(h, "gnx_link_hmc_device_fields: null handle");
  if (uf) *uf = (double*)h->U;
  if (sm) *sm = h->sm;
  if (mom) *mom = h->mom;
  if (force) *force = h->force;
  if (X) *X = (double*)h->X;
  if (Y) *Y = (double*)h->Y;
  if (phi) *phi = (double*)h->phi;
  return GNX_OK;
}

gnx_status gnx_link_hmc_compute_force(gnx_link_hmc* h, double tol, long max_iter, int* ok) {
  CHECK_ARG(h && ok, "gnx_link_hmc_compute_force: null argument");
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  if (max_iter <= 0) max_iter = c->V + 1;
  bool good = false;
  gnx_status s = lh_force(h, tol, max_iter, &good);
  if (s != GNX_OK) return s;
  HIP_TRY(hipStreamSynchronize(c->stream));
  *ok = good ? 1 : 0;
  return GNX_OK;
}

gnx_status gnx_link_hmc_trajectory(gnx_link_hmc* h, const double* mom_host, const double* chi0_host,
                                   const double* chi1_host, double eps, int steps, double tol, long max_iter,
                                   gnx_traj_result* res) {
  CHECK_ARG(h && mom_host && chi0_host && chi1_host && res, "gnx_link_hmc_trajectory: null argument");
  CHECK_ARG(steps >= 1, "gnx_link_hmc_trajectory: steps must be >= 1");
  gnx_ctx* c = h->c;
  const long V = c->V;
  HIP_TRY(hipSetDevice(c->device));
  if (max_iter <= 0) max_iter = V + 1;
  memset(res, 0, sizeof(*res));
  const double it0 = h->iterations;
  const int calls0 = h->calls;
  gnx_status s = lh_upload(h, mom_host, h->mom, 3 * (size_t)V);
  const double* chis[2] = {chi0_host, chi1_host};
  for (int f = 0; f < 2 && s == GNX_OK; f++) {
    s = lh_upload(h, chis[f], (double*)h->Ulex, (size_t)V);  // Ulex is free staging here
    if (s != GNX_OK) break;
    hipLaunchKernelGGL(k_real_to_complex_l, dim3(nblk(V, BS)), dim3(BS), 0, c->stream, (int)V,
                       (const double*)h->Ulex, h->chi + f * V);
    // phi = M chi (moleculardynamics.c:434-438): the live M, or M_U
    if (h->linked)
      HIP_TRY(link_dirac(c, 0, h->sm, h->U, h->chi + f * V, h->phi + f * V));
    else
      HIP_TRY(launch_dirac(c->g, GNX_OP_M, h->sm, (const double*)(h->chi + f * V), (double*)(h->phi + f * V),
                           c->tmp, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  if (s != GNX_OK) return s;
  HIP_TRY(hipMemcpyAsync(h->U_copy, h->U, sizeof(double2) * 3 * V, hipMemcpyDeviceToDevice, c->stream));
  bool ok = false;
#define LH_STEP(expr)                                                                                 \
  do {                                                                                                \
    gnx_status s_ = (expr);                                                                           \
    if (s_ != GNX_OK) return s_;                                                                      \
    if (!ok) {                                                                                        \
      HIP_TRY(hipMemcpyAsync(h->U, h->U_copy, sizeof(double2) * 3 * V, hipMemcpyDeviceToDevice, c->stream)); \
      HIP_TRY(hipStreamSynchronize(c->stream));                                                        \
      res->ok = 0;                                                                                    \
      res->cg_iterations = h->iterations - it0;                                                       \
      res->cg_solves = h->calls - calls0;                                                             \
      return GNX_OK;                                                                                  \
    }                                                                                                 \
  } while (0)
  const int n3 = (int)(3 * V);
  const dim3 gm(nblk(n3, BS));
  LH_STEP(lh_cg(h, tol, max_iter, &ok));
  double H0 = 0.0, H1 = 0.0;
  s = lh_energy(h, &H0);
  if (s != GNX_OK) return s;
  LH_STEP(lh_force(h, tol, max_iter, &ok));
  hipLaunchKernelGGL(k_mom_evolve, gm, dim3(BS), 0, c->stream, n3, h->mom, (const double*)h->force, eps / 2.0);
  for (int st = 0; st < steps - 1; st++) {
    hipLaunchKernelGGL(k_link_evolve, gm, dim3(BS), 0, c->stream, n3, h->U, (const double*)h->mom, eps);
    LH_STEP(lh_force(h, tol, max_iter, &ok));
    hipLaunchKernelGGL(k_mom_evolve, gm, dim3(BS), 0, c->stream, n3, h->mom, (const double*)h->force, eps);
  }
  hipLaunchKernelGGL(k_link_evolve, gm, dim3(BS), 0, c->stream, n3, h->U, (const double*)h->mom, eps);
  LH_STEP(lh_force(h, tol, max_iter, &ok));
  hipLaunchKernelGGL(k_mom_evolve, gm, dim3(BS), 0, c->stream, n3, h->mom, (const double*)h->force, eps / 2.0);
  HIP_TRY(hipGetLastError());
  s = lh_energy(h, &H1);
  if (s != GNX_OK) return s;
#undef LH_STEP
  res->H0 = H0;
  res->H1 = H1;
  res->dH = H1 - H0;
  res->ok = 1;
  res->cg_iterations = h->iterations - it0;
  res->cg_solves = h->calls - calls0;
  return GNX_OK;
}

// n chains in lockstep (moleculardynamics.c:419-542 for each): the steps of
// gnx_link_hmc_trajectory chain by chain, except that every solve point is
// one uneven-batch solve over all live chains (systems 2k, 2k+1 of m: chain
// k's two flavours), with the live operator when the chains were created
// with linked = 0 and the link-coupled one (each chain's own links) when
// with linked = 1; all chains of a call alike.  A chain whose solve fails
// leaves as gnx_link_hmc_trajectory would (links restored, ok = 0).
gnx_status gnx_link_hmc_ensemble_trajectory(gnx_multi* m, int n, gnx_link_hmc* const* hs,
                                            const double* const* mom_host, const double* const* chi0_host,
                                            const double* const* chi1_host, double eps, int steps, double tol,
                                            long max_iter, gnx_traj_result* res) {
  CHECK_ARG(m && hs && mom_host && chi0_host && chi1_host && res && n >= 1,
            "gnx_link_hmc_ensemble_trajectory: null argument");
  CHECK_ARG(steps >= 1, "gnx_link_hmc_ensemble_trajectory: steps must be >= 1");
  CHECK_ARG(multi_count(m) == 2 * n, "gnx_link_hmc_ensemble_trajectory: the solver must hold two systems per chain");
  for (int k = 0; k < n; k++) {
    CHECK_ARG(hs[k] && mom_host[k] && chi0_host[k] && chi1_host[k], "gnx_link_hmc_ensemble_trajectory: null argument");
    CHECK_ARG(hs[k]->linked == hs[0]->linked,
              "gnx_link_hmc_ensemble_trajectory: the chains share one solve, so all must have the same `linked`");
    for (int f = 0; f < 2; f++) {
      int L = 0, LT = 0, dev = -1;
      multi_system(m, 2 * k + f, &L, &LT, &dev);
      CHECK_ARG(L == hs[k]->c->g.L && LT == hs[k]->c->g.LT && dev == hs[k]->c->device,
                "gnx_link_hmc_ensemble_trajectory: system 2k, 2k+1 of the solver must have chain k's lattice and device");
    }
  }
  HIP_TRY(hipSetDevice(hs[0]->c->device));
  std::vector<char> live(n, 1), ok(n, 1);
  std::vector<double> it0(n), H0(n, 0.0);
  std::vector<int> calls0(n);
  gnx_status s;
  for (int k = 0; k < n; k++) {
    gnx_link_hmc* h = hs[k];
    gnx_ctx* c = h->c;
    const long V = c->V;
    memset(&res[k], 0, sizeof(res[k]));
    it0[k] = h->iterations;
    calls0[k] = h->calls;
    if ((s = lh_upload(h, mom_host[k], h->mom, 3 * (size_t)V)) != GNX_OK) return s;
    const double* chis[2] = {chi0_host[k], chi1_host[k]};
    for (int f = 0; f < 2; f++) {
      if ((s = lh_upload(h, chis[f], (double*)h->Ulex, (size_t)V)) != GNX_OK) return s;
      hipLaunchKernelGGL(k_real_to_complex_l, dim3(nblk(V, BS)), dim3(BS), 0, c->stream, (int)V,
                         (const double*)h->Ulex, h->chi + f * V);
      // phi = M chi (moleculardynamics.c:434-438): the live M, or M_U
      if (h->linked)
        HIP_TRY(link_dirac(c, 0, h->sm, h->U, h->chi + f * V, h->phi + f * V));
      else
        HIP_TRY(launch_dirac(c->g, GNX_OP_M, h->sm, (const double*)(h->chi + f * V), (double*)(h->phi + f * V),
                             c->tmp, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
    }
    HIP_TRY(hipMemcpyAsync(h->U_copy, h->U, sizeof(double2) * 3 * V, hipMemcpyDeviceToDevice, c->stream));
  }
  std::vector<const double2*> chain_links(n);
  for (int k = 0; k < n; k++) chain_links[k] = hs[k]->U;
  auto solve = [&]() -> gnx_status {
    gnx_status q =
        ens_solve(m, n, hs, live, tol, max_iter, GNX_START_NULL, ok, hs[0]->linked ? chain_links.data() : nullptr);
    if (q != GNX_OK) return q;
    for (int k = 0; k < n; k++)
      if (live[k] && !ok[k]) {
        live[k] = 0;
        gnx_ctx* c = hs[k]->c;
        HIP_TRY(hipMemcpyAsync(hs[k]->U, hs[k]->U_copy, sizeof(double2) * 3 * c->V, hipMemcpyDeviceToDevice,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        res[k].ok = 0;
        res[k].cg_iterations = hs[k]->iterations - it0[k];
        res[k].cg_solves = hs[k]->calls - calls0[k];
      }
    return GNX_OK;
  };
  auto force = [&]() -> gnx_status {
    for (int k = 0; k < n; k++)
      if (live[k]) lh_force_pre(hs[k]);
    gnx_status q = solve();
    for (int k = 0; k < n && q == GNX_OK; k++)
      if (live[k]) q = lh_force_post(hs[k]);
    return q;
  };
  auto mom = [&](double coef) {
    for (int k = 0; k < n; k++)
      if (live[k]) {
        const int n3 = (int)(3 * hs[k]->c->V);
        hipLaunchKernelGGL(k_mom_evolve, dim3(nblk(n3, BS)), dim3(BS), 0, hs[k]->c->stream, n3, hs[k]->mom,
                           (const double*)hs[k]->force, coef);
      }
  };
  auto links = [&](double coef) {
    for (int k = 0; k < n; k++)
      if (live[k]) {
        const int n3 = (int)(3 * hs[k]->c->V);
        hipLaunchKernelGGL(k_link_evolve, dim3(nblk(n3, BS)), dim3(BS), 0, hs[k]->c->stream, n3, hs[k]->U,
                           (const double*)hs[k]->mom, coef);
      }
  };
#define ENS_TRY(call)                     \
  do {                                    \
    if ((s = (call)) != GNX_OK) return s; \
  } while (0)
  ENS_TRY(solve());
  for (int k = 0; k < n; k++)
    if (live[k]) ENS_TRY(lh_energy(hs[k], &H0[k]));
  ENS_TRY(force());
  mom(eps / 2.0);
  for (int st = 0; st < steps - 1; st++) {
    links(eps);
    ENS_TRY(force());
    mom(eps);
  }
  links(eps);
  ENS_TRY(force());
  mom(eps / 2.0);
  HIP_TRY(hipGetLastError());
  for (int k = 0; k < n; k++) {
    if (!live[k]) continue;
    double H1 = 0.0;
    ENS_TRY(lh_energy(hs[k], &H1));
    res[k].H0 = H0[k];
    res[k].H1 = H1;
    res[k].dH = H1 - H0[k];
    res[k].ok = 1;
    res[k].cg_iterations = hs[k]->iterations - it0[k];
    res[k].cg_solves = hs[k]->calls - calls0[k];
  }
#undef ENS_TRY
  return GNX_OK;
}

gnx_status gnx_link_hmc_accept(gnx_link_hmc* h, int accepted) {
  CHECK_ARG(h, "gnx_link_hmc_accept: null handle");
  if (accepted) return GNX_OK;
  gnx_ctx* c = h->c;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(h->U, h->U_copy, sizeof(double2) * 3 * c->V, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GNX_OK;
}

gnx_status gnx_link_hmc_energy(gnx_link_hmc* h, double* H) {
  CHECK_ARG(h && H, "gnx_link_hmc_energy: null argument");
  HIP_TRY(hipSetDevice(h->c->device));
  return lh_energy(h, H);
}

gnx_status gnx_link_hmc_cg_stats(gnx_link_hmc* h, double* iterations, int* calls) {
  CHECK_ARG(h, "gnx_link_hmc_cg_stats: null handle");
  if (iterations) *iterations = h->iterations;
  if (calls) *calls = h->calls;
  return GNX_OK;
}

}  // extern "C"
