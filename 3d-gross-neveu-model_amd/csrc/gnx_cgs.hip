// gnx_cgs.hip -- FP64 CG on A = M M^dag (conjugategradient.c:13-182) for
// small lattices, the whole solve in one launch: one workgroup per system.
//
// On a lattice of a few hundred to a few thousand sites a graph-replayed
// iteration costs two kernel boundaries (~1.7 us each) for under a
// microsecond of work.  Here one 512-thread workgroup runs every iteration
// of one system: the two vectors gathered at neighbours (the direction p and
// t = M^dag p) live in LDS, the residual, the solution and the sigma mass of
// a thread's sites in its registers, and a workgroup barrier stands where
// the graph had a kernel boundary.  The steps are the two-reduction kernels'
// (k_cgm_A / k_cgm_B: the reference's recurrences, conjugategradient.c:58-103,
// with M^dag in its gather form and M in reference order), the starts and
// exits the reference's (prologue early exit of the null start, the imposed
// start's ||b|| < tol exit, absolute tolerance, the iteration cap); the
// global sums are deterministic workgroup sums (DPP within a wave, then the
// waves in order), so every decision is uniform across the workgroup.
#include "gnx_internal.h"
#include "gnx_solver.h"

namespace gnx {

constexpr int CGS_BS = 512;
constexpr int CGS_MAX_V = 1728;  // 12^3: p and t in LDS take 2 V x 16 B = 55 KB
// a solve runs in one launch, so its length is bounded by the iteration cap
// (100 000 iterations of 1728 sites take well under a second); longer caps
// go to the graph-replayed solver, whose launches stay short
constexpr long CGS_MAX_ITER = 100000;

// Workgroup sum in the order of block_sum (the wave sums added in wave
// order, the same bits) with the wave count a constant of this kernel; no
// barrier ahead of the writes (the caller alternates lds slots).  (The
// run-time wave count of block_sum_once costs a branch; 1.68 against 1.48 us
// per iteration at 8^3.)
GNX_DEVONLY double cgs_sum(double v, double* lds) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  return sum_waves_n<CGS_BS / 64>(lds);
}

// Three workgroup barriers per iteration: after the write of p, and one in
// each global sum.  The sums alternate between two LDS slots, so a sum needs
// no barrier ahead of its writes (the other sum's barrier lies between a
// slot's readers and its next writers).  (Forming p = r + beta p_old at the
// neighbours from r and p_old in LDS saves the first barrier but costs more
// than it saves: 2.23 against 1.99 us per iteration at 8^3.)
template <int SPT>
__global__ __launch_bounds__(CGS_BS) void k_cgs(Geom g, const double* __restrict__ sm,
                                                const double2* __restrict__ b, double2* __restrict__ x, double tol,
                                                long max_iter, int start, CGDevResult* __restrict__ res) {
  extern __shared__ double2 shm[];  // t [V], then p [V]
  __shared__ double red[2][CGS_BS / 64];
  double2* T = shm;
  double2* P = shm + g.V;
  const int rhs = blockIdx.x;
  b += (size_t)rhs * g.V;
  x += (size_t)rhs * g.V;
  const bool lead = threadIdx.x == 0;
  // a thread's sites f = i * CGS_BS + thread (-1 past the last), their
  // coordinates and neighbours, sigma mass, residual, solution, direction
  int fs[SPT];
  Site st[SPT];
  SiteCode kc[SPT];
  NbrF nb[SPT];
  double smv[SPT];
  double2 rv[SPT], xv[SPT], pv[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = i * CGS_BS + (int)threadIdx.x;
    fs[i] = f < g.V ? f : -1;
    const int fc = f < g.V ? f : 0;
    st[i] = decode(g, fc);
    kc[i] = site_code(g, st[i]);
    nb[i] = neighbours_f(g, st[i]);
    smv[i] = sm[fc];
    rv[i] = b[fc];
    xv[i] = x[fc];
  }
  if (start == GNX_START_NULL) {
    // prologue (conjugategradient.c:18-33): r = b - M M^dag x_in, with the
    // operators in reference order; x_in kept if |r| < tol
    double2* Xin = P;
#pragma unroll
    for (int i = 0; i < SPT; i++)
      if (fs[i] >= 0) Xin[fs[i]] = xv[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SPT; i++)
      if (fs[i] >= 0) T[fs[i]] = apply_Mdag(g, st[i], smv[i], [&](int q) { return Xin[q]; });
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < SPT; i++)
      if (fs[i] >= 0) acc += cnorm2(apply_sub_M(g, st[i], smv[i], rv[i], [&](int q) { return T[q]; }));
    const double nr = cgs_sum(acc, red[0]);
    if (sqrt(nr) < tol) {
      if (lead) {
        const CGDevResult r = {1, 1, 0, 0, sqrt(nr), {0, 0, 0}};
        res[rhs] = r;
      }
      return;
    }
  }
  // x = 0, r = p = b (conjugategradient.c:31-33, :121-122)
  double rho;
  {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      xv[i] = make_double2(0.0, 0.0);
      pv[i] = rv[i];
      acc += fs[i] >= 0 ? cnorm2(rv[i]) : 0.0;
    }
    rho = cgs_sum(acc, red[1]);
  }
  if (start == GNX_START_IMPOSED && sqrt(rho) < tol) {  // conjugategradient.c:123-128
#pragma unroll
    for (int i = 0; i < SPT; i++)
      if (fs[i] >= 0) x[fs[i]] = xv[i];
    if (lead) {
      const CGDevResult r = {1, 1, 0, 0, sqrt(rho), {0, 0, 0}};
      res[rhs] = r;
    }
    return;
  }
  double rho_prev = rho;
  int conv = 0;
  long it = 0;
  for (;; it++) {
    // the exit test on |r| (absolute tolerance, conjugategradient.c:86-94;
    // the iteration cap, :58), beta and p = r + beta p (:96-102), then
    // t = M^dag p and |t|^2 (:67-69)
    if (it > 0) {
      conv = sqrt(rho) < tol ? 1 : 0;
      if (conv || it >= max_iter) break;
    }
    const double beta = it == 0 ? 0.0 : rho / rho_prev;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      pv[i] = cadd(rv[i], rmul(beta, pv[i]));
      if (fs[i] >= 0) P[fs[i]] = pv[i];
    }
    __syncthreads();
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      if (fs[i] < 0) continue;
      const double2 vpp[3] = {P[nb[i].pp[0]], P[nb[i].pp[1]], P[nb[i].pp[2]]};
      const double2 vmm[3] = {P[nb[i].mm[0]], P[nb[i].mm[1]], P[nb[i].mm[2]]};
      const double2 tt = apply_Mdag_fast_vals(kc[i], smv[i], pv[i], vpp, vmm);
      T[fs[i]] = tt;
      acc += cnorm2(tt);
    }
    const double gamma = cgs_sum(acc, red[0]);  // its barrier publishes t
    // alpha = rho / |t|^2, x += alpha p, r -= alpha M t, |r|^2
    // (conjugategradient.c:73-87)
    const double al = rho / gamma;
    acc = 0.0;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      if (fs[i] < 0) continue;
      const double2 tpp[3] = {T[nb[i].pp[0]], T[nb[i].pp[1]], T[nb[i].pp[2]]};
      const double2 tmm[3] = {T[nb[i].mm[0]], T[nb[i].mm[1]], T[nb[i].mm[2]]};
      const double2 w = apply_M_vals(kc[i], smv[i], T[fs[i]], tpp, tmm);
      xv[i] = cadd(xv[i], rmul(al, pv[i]));
      rv[i] = cadd(rv[i], rmul(-al, w));
      acc += cnorm2(rv[i]);
    }
    rho_prev = rho;
    // its barrier keeps this iteration's reads of t ahead of the next
    // iteration's writes (the reads of p are ahead of the barrier in the
    // sum of |t|^2)
    rho = cgs_sum(acc, red[1]);
  }
#pragma unroll
  for (int i = 0; i < SPT; i++)
    if (fs[i] >= 0) x[fs[i]] = xv[i];
  if (lead) {
    const CGDevResult r = {conv, 0, (int)it, 0, sqrt(rho), {0, 0, 0}};
    res[rhs] = r;
  }
}

bool cgs_fits(const Geom& g, long max_iter) { return g.V <= CGS_MAX_V && max_iter <= CGS_MAX_ITER; }

hipError_t cgs_run(const Geom& g, const double* sm, const double2* b, double2* x, double tol, long max_iter,
                   int start, int nrhs, CGDevResult* dres, hipStream_t st) {
  if (!cgs_fits(g, max_iter) || nrhs < 1) return hipErrorInvalidValue;
  const int spt = (g.V + CGS_BS - 1) / CGS_BS;
  const void* k = spt == 1 ? (const void*)k_cgs<1> : spt == 2 ? (const void*)k_cgs<2> : (const void*)k_cgs<4>;
  const size_t lds = sizeof(double2) * 2 * (size_t)g.V;
  Geom gg = g;
  void* args[] = {(void*)&gg, (void*)&sm, (void*)&b, (void*)&x, (void*)&tol, (void*)&max_iter, (void*)&start,
                  (void*)&dres};
  return hipLaunchKernel(k, dim3(nrhs), dim3(CGS_BS), args, lds, st);
}

int cgs_sites_per_thread(const Geom& g) {
  const int spt = (g.V + CGS_BS - 1) / CGS_BS;
  return spt <= 2 ? spt : 4;
}

}  // namespace gnx
