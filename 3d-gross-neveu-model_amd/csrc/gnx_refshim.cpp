// gnx_refshim.cpp -- the reference's hot-path functions (include/gnx_refapi.h)
// on top of the gnx C ABI.  Host structs come from the caller; vectors are
// staged to device buffers owned here.  Per-call functions (Dirac, CG, sigma
// mass, force) copy their operands in and their results out; the trajectory
// and the condensate keep everything on the device between their first
// upload and final download.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/gnx.h"
#include "../../include/gnx_refapi.h"

namespace {

struct Shim {
  gnx_ctx* ctx = nullptr;
  gnx_hmc* hmc = nullptr;
  double hmc_beta = -1.0;
  long V = 0;
  int L = 0, LT = 0;
  double (*gauss1)(void) = nullptr;
  double (*casuale)(void) = nullptr;
  void (*gauss2mv)(double*, double*, double, double) = nullptr;
  int pure_gauge = 0;        // PURE_GAUGE, macro.h:47 (off as shipped)
  gnx_link_hmc* lhmc = nullptr;
  double lhmc_beta = 0.0;
  int lhmc_pure = -1;
  double* uf = nullptr;      // device [3V] complex links, f-order
  double* ulex = nullptr;    // device [3V] complex links, lex order
  double* lforce = nullptr;  // device [3V]
  double tol_high = 1e-9;    // HIGH_PRECISION_INVERSION, macro.h:136
  double tol_medium = 1e-3;  // MEDIUM_PRECISION_INVERSION, macro.h:135
  double* sm = nullptr;      // device V
  double* sigma = nullptr;   // device V
  double* force = nullptr;   // device V
  double* in = nullptr;      // device 2V
  double* out = nullptr;     // device 2V
  double* phi = nullptr;     // device [2][2V]
  double* X = nullptr;       // device [2][2V]
  double* Y = nullptr;       // device [2][2V]
};

Shim g_shim;

void report(const char* where, gnx_status s) {
  if (s != GNX_OK) fprintf(stderr, "gnx_ref: %s failed: %s\n", where, gnx_last_error());
}

bool ok_hip(hipError_t e, const char* where) {
  if (e != hipSuccess) {
    fprintf(stderr, "gnx_ref: %s: %s\n", where, hipGetErrorString(e));
    return false;
  }
  return true;
}

size_t vbytes() { return sizeof(double) * 2 * (size_t)g_shim.V; }
size_t fbytes() { return sizeof(double) * (size_t)g_shim.V; }

bool up(void* dev, const void* host, size_t n) { return ok_hip(hipMemcpy(dev, host, n, hipMemcpyHostToDevice), "upload"); }
bool down(void* host, const void* dev, size_t n) {
  return ok_hip(hipMemcpy(host, dev, n, hipMemcpyDeviceToHost), "download");
}

bool ready(const char* fn) {
  if (!g_shim.ctx) {
    fprintf(stderr, "gnx_ref: %s called before GnxRefInit\n", fn);
    return false;
  }
  return true;
}

gnx_hmc* hmc_for(double beta_s) {
  if (g_shim.hmc && g_shim.hmc_beta == beta_s) return g_shim.hmc;
  if (g_shim.hmc) gnx_hmc_destroy(g_shim.hmc);
  g_shim.hmc = nullptr;
  gnx_status s = gnx_hmc_create(g_shim.ctx, beta_s, 4, &g_shim.hmc);
  report("gnx_hmc_create", s);
  g_shim.hmc_beta = beta_s;
  return g_shim.hmc;
}

void apply(gnx_op op, const Lattice* lat, const Fermion* in, Fermion* target, bool keep_target) {
  if (!ready("Dirac")) return;
  if (!up(g_shim.sm, lat->sigmamass, fbytes()) || !up(g_shim.in, in->fermion, vbytes())) return;
  if (keep_target && !up(g_shim.out, target->fermion, vbytes())) return;
  report("gnx_dirac", gnx_dirac(g_shim.ctx, op, g_shim.sm, g_shim.in, g_shim.out));
  report("gnx_synchronize", gnx_synchronize(g_shim.ctx));
  down(target->fermion, g_shim.out, vbytes());
}

bool cg(const Lattice* lat, const Fermion* b, Fermion* x, double tol, Fermionkit* kit, gnx_cg_start start) {
  if (!ready("ConjugateGradient")) return false;
  if (!up(g_shim.sm, lat->sigmamass, fbytes()) || !up(g_shim.in, b->fermion, vbytes()) ||
      !up(g_shim.out, x->fermion, vbytes()))
    return false;
  gnx_cg_result r;
  gnx_status s = gnx_cg(g_shim.ctx, g_shim.sm, g_shim.in, g_shim.out, tol, 0, start, GNX_PREC_F64, &r);
  if (s != GNX_OK && s != GNX_ERR_NOCONV) {
    report("gnx_cg", s);
    return false;
  }
  down(x->fermion, g_shim.out, vbytes());
  // kit counters change only on in-loop convergence (conjugategradient.c:92-93)
  if (r.converged && !r.early_exit) {
    kit->iterations += r.iterations;
    kit->CG_call++;
  }
  return r.converged != 0;
}

// flavour solves + M^dag + fermion force terms, leaving X, Y in kit
bool force_terms(const Lattice* lat, const Info* info, Forcesigma* force, Fermionkit* kit, double tol,
                 bool from_zero) {
  if (!ready("force")) return false;
  const size_t vb = vbytes();
  if (!up(g_shim.sm, lat->sigmamass, fbytes())) return false;
  for (int f = 0; f < 2; f++) {
    if (!up(g_shim.phi + f * 2 * g_shim.V, kit->phi.flav[f].fermion, vb) ||
        !up(g_shim.X + f * 2 * g_shim.V, kit->inverted_phi.flav[f].fermion, vb))
      return false;
  }
  gnx_cg_result r[2];
  gnx_status s = gnx_cg_batch(g_shim.ctx, g_shim.sm, g_shim.phi, g_shim.X, 2, tol, 0, GNX_START_NULL,
                              GNX_PREC_F64, r);
  if (s != GNX_OK && s != GNX_ERR_NOCONV) {
    report("gnx_cg_batch", s);
    return false;
  }
  for (int f = 0; f < 2; f++) {
    down(kit->inverted_phi.flav[f].fermion, g_shim.X + f * 2 * g_shim.V, vb);
    if (r[f].converged && !r[f].early_exit) {
      kit->iterations += r[f].iterations;
      kit->CG_call++;
    }
  }
  if (!r[0].converged || !r[1].converged) return false;
  for (int f = 0; f < 2; f++) {
    report("gnx_dirac", gnx_dirac(g_shim.ctx, GNX_OP_MDAG, g_shim.sm, g_shim.X + f * 2 * g_shim.V,
                                  g_shim.Y + f * 2 * g_shim.V));
  }
  if (!from_zero && !up(g_shim.force, force->force, fbytes())) return false;
  if (from_zero && !up(g_shim.sigma, lat->sigma, fbytes())) return false;
  s = gnx_sigma_force_terms(g_shim.ctx, g_shim.sigma, info->beta_scalar, 4, g_shim.X, g_shim.Y,
                            g_shim.X + 2 * g_shim.V, g_shim.Y + 2 * g_shim.V, g_shim.force, from_zero ? 0 : 1,
                            from_zero ? 1 : 0);
  report("gnx_sigma_force_terms", s);
  report("gnx_synchronize", gnx_synchronize(g_shim.ctx));
  for (int f = 0; f < 2; f++) down(kit->inverted_phi_obs.flav[f].fermion, g_shim.Y + f * 2 * g_shim.V, vb);
  down(force->force, g_shim.force, fbytes());
  return s == GNX_OK;
}

}  // namespace

// lattice volume bound by GnxRefInit, for the checkpoint code (gnx_refio.cpp)
long gnx_ref_bound_volume() { return g_shim.V; }

extern "C" {

int GnxRefInit(int lspace, int ltime, int device, double (*gauss1)(void), double (*casuale)(void)) {
  GnxRefFinalize();
  gnx_status s = gnx_create(lspace, ltime, device, &g_shim.ctx);
  if (s != GNX_OK) {
    report("gnx_create", s);
    return 1;
  }
  g_shim.L = lspace;
  g_shim.LT = ltime;
  g_shim.V = gnx_volume(g_shim.ctx);
  g_shim.gauss1 = gauss1;
  g_shim.casuale = casuale;
  const size_t vb = vbytes(), fb = fbytes();
  hipError_t e = hipMalloc(&g_shim.sm, fb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.sigma, fb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.force, fb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.in, vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.out, vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.phi, 2 * vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.X, 2 * vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.Y, 2 * vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.uf, 3 * vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.ulex, 3 * vb);
  if (e == hipSuccess) e = hipMalloc(&g_shim.lforce, 3 * fb);
  if (!ok_hip(e, "GnxRefInit allocation")) {
    GnxRefFinalize();
    return 2;
  }
  return 0;
}

void GnxRefFinalize(void) {
  if (g_shim.hmc) gnx_hmc_destroy(g_shim.hmc);
  if (g_shim.lhmc) gnx_link_hmc_destroy(g_shim.lhmc);
  double* bufs[] = {g_shim.sm,  g_shim.sigma, g_shim.force, g_shim.in,   g_shim.out,   g_shim.phi,
                    g_shim.X,   g_shim.Y,     g_shim.uf,    g_shim.ulex, g_shim.lforce};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  if (g_shim.ctx) gnx_destroy(g_shim.ctx);
  double (*g1)(void) = g_shim.gauss1;
  double (*cs)(void) = g_shim.casuale;
  void (*g2)(double*, double*, double, double) = g_shim.gauss2mv;
  const int pg = g_shim.pure_gauge;
  g_shim = Shim();
  g_shim.gauss1 = g1;
  g_shim.casuale = cs;
  g_shim.gauss2mv = g2;
  g_shim.pure_gauge = pg;
}

void GnxRefBindGauss2(void (*gauss2mv)(double*, double*, double, double)) { g_shim.gauss2mv = gauss2mv; }

void GnxRefSetPureGauge(int on) { g_shim.pure_gauge = on ? 1 : 0; }

void GnxRefSetTolerances(double high, double medium) {
  g_shim.tol_high = high;
  g_shim.tol_medium = medium;
}

void DiracFermion(const Lattice* lat, const Fermion* fermion, Fermion* target, const Info*, const Fermioncoor*) {
  apply(GNX_OP_M, lat, fermion, target, false);
}

void AdjDiracFermion(const Lattice* lat, const Fermion* fermion, Fermion* target, const Info*, const Fermioncoor*) {
  apply(GNX_OP_MDAG, lat, fermion, target, false);
}

void SubtractDiracFermion(const Lattice* lat, const Fermion* fermion, Fermion* target, const Info*,
                          const Fermioncoor*) {
  apply(GNX_OP_SUB_M, lat, fermion, target, true);
}

void SubtractAdjDiracFermion(const Lattice* lat, const Fermion* fermion, Fermion* target, const Info*,
                             const Fermioncoor*) {
  apply(GNX_OP_SUB_MDAG, lat, fermion, target, true);
}

bool ConjugateGradientFastDiracAdjFastDiracToTheMinusOne(const Lattice* lat, const Fermion* fermion,
                                                         Fermion* target_inverted, const Info*, const Fermioncoor*,
                                                         double inversion_precision, Fermionkit* kit, int) {
  return cg(lat, fermion, target_inverted, inversion_precision, kit, GNX_START_NULL);
}

bool StartX0ImposedConjugateGradientFastDiracAdjFastDiracToTheMinusOne(const Lattice* lat, const Fermion* fermion,
                                                                       Fermion* target_inverted, const Info*,
                                                                       const Fermioncoor*,
                                                                       double inversion_precision, Fermionkit* kit,
                                                                       int) {
  return cg(lat, fermion, target_inverted, inversion_precision, kit, GNX_START_IMPOSED);
}

void ComputeFullSigmaMass(Lattice* lat, const Fermioncoor*, double* sigmamass) {
  if (!ready("ComputeFullSigmaMass")) return;
  if (!up(g_shim.sigma, lat->sigma, fbytes())) return;
  report("gnx_sigma_mass", gnx_sigma_mass(g_shim.ctx, g_shim.sigma, g_shim.sm));
  report("gnx_synchronize", gnx_synchronize(g_shim.ctx));
  down(sigmamass, g_shim.sm, fbytes());
}

bool AddForceFromFermionSectorRespectToSigma(const Lattice* lat, const Info* info, const Fermioncoor*,
                                             Forcesigma* force, Fermionkit* kit, double inversion_precision) {
  return force_terms(lat, info, force, kit, inversion_precision, false);
}

bool ComputeSigmaForce(const Lattice* lat, const Info* info, const Fermioncoor*, Forcesigma* force,
                       Fermionkit* kit, double inversion_precision) {
  // SetForceSigmaToZero + fermion part + sigma part (moleculardynamics.c:631-647)
  return force_terms(lat, info, force, kit, inversion_precision, true);
}

// AddForceFromFermionSectorRespectToLink (moleculardynamics.c:164-193): the
// solves and M^dag use the live operator, the force the links.  The force
// accumulates onto g_shim.lforce, uploaded from the caller unless on_device.
static bool link_fermion_force(const Lattice* lat, Force* force, Fermionkit* kit, bool on_device) {
  const long V = g_shim.V;
  const size_t vb = vbytes();
  if (!up(g_shim.sm, lat->sigmamass, fbytes()) || !up(g_shim.ulex, lat->dlink, 3 * vb)) return false;
  if (!on_device && !up(g_shim.lforce, force->force, 3 * fbytes())) return false;
  for (int f = 0; f < 2; f++)
    if (!up(g_shim.phi + f * 2 * V, kit->phi.flav[f].fermion, vb) ||
        !up(g_shim.X + f * 2 * V, kit->inverted_phi.flav[f].fermion, vb))
      return false;
  gnx_cg_result r[2];
  gnx_status s = gnx_cg_batch(g_shim.ctx, g_shim.sm, g_shim.phi, g_shim.X, 2, g_shim.tol_high, 0, GNX_START_NULL,
                              GNX_PREC_F64, r);
  for (int f = 0; f < 2; f++) {
    down(kit->inverted_phi.flav[f].fermion, g_shim.X + f * 2 * V, vb);
    if (r[f].converged && !r[f].early_exit) {
      kit->iterations += r[f].iterations;
      kit->CG_call++;
    }
  }
  if (s != GNX_OK) {
    if (s != GNX_ERR_NOCONV) report("gnx_cg_batch", s);
    return false;
  }
  for (int f = 0; f < 2; f++)
    report("gnx_dirac", gnx_dirac(g_shim.ctx, GNX_OP_MDAG, g_shim.sm, g_shim.X + f * 2 * V, g_shim.Y + f * 2 * V));
  report("gnx_links_from_lex", gnx_links_from_lex(g_shim.ctx, g_shim.ulex, g_shim.uf));
  s = gnx_link_force_from_solutions(g_shim.ctx, g_shim.uf, g_shim.X, g_shim.Y, g_shim.X + 2 * V, g_shim.Y + 2 * V,
                                    g_shim.lforce, 1);
  if (s != GNX_OK) {
    report("gnx_link_force_from_solutions", s);
    return false;
  }
  report("gnx_synchronize", gnx_synchronize(g_shim.ctx));
  for (int f = 0; f < 2; f++) down(kit->inverted_phi_obs.flav[f].fermion, g_shim.Y + f * 2 * V, vb);
  down(force->force, g_shim.lforce, 3 * fbytes());
  return true;
}

bool AddForceFromFermionSectorRespectToLink(const Lattice* lat, const Info*, const Fermioncoor*, Force* force,
                                            Fermionkit* kit) {
  if (!ready("AddForceFromFermionSectorRespectToLink")) return false;
  return link_fermion_force(lat, force, kit, false);
}

bool ComputeForce(const Lattice* lattice, const Info* info, Force* force, const Fermioncoor*, Fermionkit* kit) {
  if (!ready("ComputeForce")) return false;
  // SetForceToZero, the PURE_GAUGE part, then the fermion part (moleculardynamics.c:203-222)
  if (g_shim.pure_gauge) {
    if (!up(g_shim.ulex, lattice->dlink, 3 * vbytes())) return false;
    report("gnx_links_from_lex", gnx_links_from_lex(g_shim.ctx, g_shim.ulex, g_shim.uf));
    report("gnx_gauge_force", gnx_gauge_force(g_shim.ctx, g_shim.uf, info->beta, g_shim.lforce));
  } else if (!ok_hip(hipMemset(g_shim.lforce, 0, 3 * fbytes()), "ComputeForce")) {
    return false;
  }
  return link_fermion_force(lattice, force, kit, true);
}

static gnx_link_hmc* lhmc_for(double beta) {
  if (g_shim.lhmc && g_shim.lhmc_beta == beta && g_shim.lhmc_pure == g_shim.pure_gauge) return g_shim.lhmc;
  if (g_shim.lhmc) gnx_link_hmc_destroy(g_shim.lhmc);
  g_shim.lhmc = nullptr;
  gnx_status s = gnx_link_hmc_create(g_shim.ctx, beta, g_shim.pure_gauge, 0, &g_shim.lhmc);
  if (s != GNX_OK) {
    report("gnx_link_hmc_create", s);
    return nullptr;
  }
  g_shim.lhmc_beta = beta;
  g_shim.lhmc_pure = g_shim.pure_gauge;
  return g_shim.lhmc;
}

bool LeapFrog2Link(Lattice* lattice, Info* info, Force* force, Momenta* momenta, const Fermioncoor*, Fermionkit* kit,
                   double epsilon, const int step, const char* acceptance_file) {
  if (!ready("LeapFrog2Link")) return false;
  if (!g_shim.gauss1 || !g_shim.casuale || !g_shim.gauss2mv) {
    fprintf(stderr, "gnx_ref: LeapFrog2Link needs Gauss1, Casuale (GnxRefInit) and Gauss2MeanVariance "
                    "(GnxRefBindGauss2)\n");
    return false;
  }
  const long V = g_shim.V;
  FILE* opf = fopen(acceptance_file, "a");
  // draws in the reference's order (moleculardynamics.c:38-46, 428-438)
  for (long r = 0; r < 3 * V / 2; r++) g_shim.gauss2mv(&momenta->momenta[2 * r], &momenta->momenta[2 * r + 1], 0.0, 1.0);
  std::vector<double> chi[2] = {std::vector<double>(V), std::vector<double>(V)};
  for (int f = 0; f < 2; f++)
    for (long r = 0; r < V; r++) {
      chi[f][r] = g_shim.gauss1();
      kit->chi.flav[f].fermion[r] = chi[f][r];
    }
  gnx_link_hmc* h = lhmc_for(info->beta);
  if (!h) {
    if (opf) fclose(opf);
    return false;
  }
  report("gnx_link_hmc_set_fields", gnx_link_hmc_set_fields(h, (const double*)lattice->dlink, lattice->sigmamass));
  report("gnx_link_hmc_set_solutions",
         gnx_link_hmc_set_solutions(h, (const double*)kit->inverted_phi.flav[0].fermion,
                                    (const double*)kit->inverted_phi.flav[1].fermion));
  gnx_traj_result res;
  gnx_status s = gnx_link_hmc_trajectory(h, momenta->momenta, chi[0].data(), chi[1].data(), epsilon, step,
                                         g_shim.tol_high, 0, &res);
  double *duf, *dsm, *dmom, *dforce, *dX, *dY, *dphi;
  gnx_link_hmc_device_fields(h, &duf, &dsm, &dmom, &dforce, &dX, &dY, &dphi);
  kit->iterations += (double)res.cg_iterations;
  kit->CG_call += res.cg_solves;
  const size_t vb = vbytes();
  if (s != GNX_OK || !res.ok) {
    // a failed solve restores the links (moleculardynamics.c:444-449)
    if (s != GNX_OK) report("gnx_link_hmc_trajectory", s);
    gnx_link_hmc_get_links(h, (double*)lattice->dlink);
    if (opf) fclose(opf);
    return false;
  }
  if (opf) fprintf(opf, "# Slink\t%.16f\n", res.H0);
  info->delta_H_link = res.dH;
  if (g_shim.casuale() <= exp(-(res.H1 - res.H0))) {
    info->hybrid_accepted_link++;
    info->hybrid_done_link++;
    info->hybrid_done_acceptance_link = (double)(info->hybrid_accepted_link) / (double)(info->hybrid_done_link);
    if (info->story_link >= 0) info->story_link++;
    else info->story_link = -1;
  } else {
    report("gnx_link_hmc_accept", gnx_link_hmc_accept(h, 0));
    info->hybrid_done_link++;
    info->hybrid_done_acceptance_link = (double)(info->hybrid_accepted_link) / (double)(info->hybrid_done_link);
    if (info->story_link >= 0) info->story_link = -1;
    else info->story_link--;
  }
  gnx_link_hmc_get_links(h, (double*)lattice->dlink);
  down(momenta->momenta, dmom, 3 * fbytes());
  down(force->force, dforce, 3 * fbytes());
  for (int f = 0; f < 2; f++) {
    down(kit->phi.flav[f].fermion, dphi + f * 2 * V, vb);
    down(kit->inverted_phi.flav[f].fermion, dX + f * 2 * V, vb);
    down(kit->inverted_phi_obs.flav[f].fermion, dY + f * 2 * V, vb);
  }
  // with SIGMA_INTERACTION on (the shipped build) no acceptance line follows (:523-526)
  if (opf) fclose(opf);
  return true;
}

bool LeapFrog2MinimumNormSigmaVelocityFirst(Lattice* lattice, Info* info, Forcesigma* force, Momentasigma* momenta,
                                            const Fermioncoor*, Fermionkit* kit, double epsilon, const int step,
                                            const char* acceptance_file) {
  if (!ready("LeapFrog2MinimumNormSigmaVelocityFirst")) return false;
  if (!g_shim.gauss1 || !g_shim.casuale) {
    fprintf(stderr, "gnx_ref: no host RNG bound (GnxRefInit)\n");
    return false;
  }
  const long V = g_shim.V;
  FILE* opf = fopen(acceptance_file, "a");
  // draws in the reference's order (moleculardynamics.c:1241-1250)
  for (long x = 0; x < V; x++) momenta->momenta[x] = g_shim.gauss1();
  std::vector<double> chi[2] = {std::vector<double>(V), std::vector<double>(V)};
  for (int f = 0; f < 2; f++)
    for (long r = 0; r < V; r++) {
      chi[f][r] = g_shim.gauss1();
      kit->chi.flav[f].fermion[r] = chi[f][r];
    }
  gnx_hmc* h = hmc_for(info->beta_scalar);
  if (!h) {
    if (opf) fclose(opf);
    return false;
  }
  report("gnx_hmc_set_sigma", gnx_hmc_set_sigma(h, lattice->sigma));
  gnx_traj_result res;
  gnx_status s = gnx_hmc_trajectory(h, momenta->momenta, chi[0].data(), chi[1].data(), epsilon, step,
                                    g_shim.tol_high, g_shim.tol_medium, 0, &res);
  double *dsig, *dsm, *dmom, *dforce, *dX, *dY, *dphi;
  gnx_hmc_device_fields(h, &dsig, &dsm, &dmom, &dforce, &dX, &dY, &dphi);
  kit->iterations += (double)res.cg_iterations;
  kit->CG_call += res.cg_solves;
  if (s != GNX_OK) {
    // a failed solve restores sigma and its mass (moleculardynamics.c:1262-1283)
    if (s != GNX_ERR_NOCONV) report("gnx_hmc_trajectory", s);
    gnx_hmc_get_sigma(h, lattice->sigma, lattice->sigmamass);
    if (opf) fclose(opf);
    return false;
  }
  const double final_energy = res.H1, starting_energy = res.H0;
  // evolve steps add epsilon/2 to the travelled distance (:832-835)
  for (int k = 0; k < 2 * step; k++) lattice->distance_sigma += (epsilon / 2.0);
  if (g_shim.casuale() <= exp(-(final_energy - starting_energy))) {
    info->hybrid_accepted_sigma++;
    info->hybrid_done_sigma++;
    info->hybrid_done_acceptance_sigma = (double)(info->hybrid_accepted_sigma) / (double)(info->hybrid_done_sigma);
    info->local_accepted_sigma++;
    info->local_done_sigma++;
    info->local_acceptance_sigma = (double)(info->local_accepted_sigma) / (double)(info->local_done_sigma);
    lattice->done_sigma++;
    if (info->story_sigma >= 0) info->story_sigma++;
    else info->story_sigma = -1;
  } else {
    report("gnx_hmc_accept", gnx_hmc_accept(h, 0));
    info->hybrid_done_sigma++;
    info->hybrid_done_acceptance_sigma = (double)(info->hybrid_accepted_sigma) / (double)(info->hybrid_done_sigma);
    info->local_done_sigma++;
    info->local_acceptance_sigma = (double)(info->local_accepted_sigma) / (double)(info->local_done_sigma);
    lattice->done_sigma++;
    if (info->story_sigma >= 0) info->story_sigma = -1;
    else info->story_sigma--;
  }
  // the host structs end as the reference leaves them
  gnx_hmc_get_sigma(h, lattice->sigma, lattice->sigmamass);
  const size_t vb = vbytes();
  down(momenta->momenta, dmom, fbytes());
  down(force->force, dforce, fbytes());
  for (int f = 0; f < 2; f++) {
    down(kit->phi.flav[f].fermion, dphi + f * 2 * V, vb);
    down(kit->inverted_phi.flav[f].fermion, dX + f * 2 * V, vb);
    down(kit->inverted_phi_obs.flav[f].fermion, dY + f * 2 * V, vb);
  }
  if (opf) {
    fprintf(opf, "%.16f\t%.16f\t%.16f\t%.16f\t%d\t%.16f\t%.16f\t%.16f\t%.16f\t%d\t%.16f\t%.16f\t%.16f\n", info->beta,
            info->beta_scalar, info->mass, info->epsilon_link_HMC, info->steps_link_HMC,
            (info->epsilon_link_HMC * info->steps_link_HMC), info->delta_H_link, info->hybrid_done_acceptance_link,
            info->epsilon_sigma_HMC, info->steps_sigma_HMC, (info->epsilon_sigma_HMC * info->steps_sigma_HMC),
            final_energy - starting_energy, info->hybrid_done_acceptance_sigma);
    fclose(opf);
  }
  return true;
}

void ComputeCondensateDirac(const Lattice* lat, double* tr1, double* tr1_abs, double* tr2, double* tr4, double* mu2,
                            double* mu4, double* chi, double* chi_abs, double* chip, double* c12_nosigma,
                            double* c12_withsigma, RandomVec* vec, Fermion* target_inverted, const Info* info,
                            const Fermioncoor*, double inversion_precision, Fermionkit* kit, int) {
  if (!ready("ComputeCondensateDirac")) return;
  const long V = g_shim.V;
  // the six Z2 vectors, drawn in the reference's order (condensate.c:52-60,
  // 130-131, 202, 217); no other draw happens between them
  std::vector<double> z2(6 * V);
  for (long k = 0; k < 6 * V; k++) z2[k] = (g_shim.casuale() <= 0.5) ? 1.0 : -1.0;
  gnx_hmc* h = hmc_for(info->beta_scalar);
  if (!h) return;
  double *dsig, *dsm, *dmom, *dforce, *dX, *dY, *dphi;
  gnx_hmc_device_fields(h, &dsig, &dsm, &dmom, &dforce, &dX, &dY, &dphi);
  // the caller's sigma mass and x_in (target_inverted aliases inverted_phi_obs[0])
  if (!up(dsm, lat->sigmamass, fbytes()) || !up(dY, target_inverted->fermion, vbytes())) return;
  double it0;
  int calls0;
  gnx_hmc_cg_stats(h, &it0, &calls0);
  double out[11];
  gnx_status s = gnx_hmc_condensate(h, z2.data(), inversion_precision, 0, out);
  report("gnx_hmc_condensate", s);
  if (s != GNX_OK) return;
  double it1;
  int calls1;
  gnx_hmc_cg_stats(h, &it1, &calls1);
  kit->iterations += it1 - it0;
  kit->CG_call += calls1 - calls0;
  *tr1 = out[0];
  *tr1_abs = out[1];
  *tr2 = out[2];
  *tr4 = out[3];
  *mu2 = out[4];
  *mu4 = out[5];
  *chi = out[6];
  *chi_abs = out[7];
  *chip = out[8];
  *c12_nosigma = out[9];
  *c12_withsigma = out[10];
  down(target_inverted->fermion, dY, vbytes());
  // the random vector ends holding the sigma-weighted last source (:221)
  for (long x = 0; x < V; x++) vec->vec->fermion[x] = z2[5 * V + x] * lat->sigmamass[x];
}

}  // extern "C"
