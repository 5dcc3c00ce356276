// gnx_cgu.hip -- FP64 CG on A = M M^dag (conjugategradient.c:13-182) for n
// independent systems on lattices of different sizes, each with its own sigma
// mass, in one launch per kernel: the graph-replayed two-kernel state machine
// of gnx_cgm.hip with the geometry taken per system from a descriptor table.
//
//   grid = (largest system's block count, n); block (b, k) works on system k
//   when b < nb_k and leaves at once otherwise.  A system that has left the
//   loop leaves at its state read, before any operand load, so systems whose
//   iteration counts differ do not pay for each other's iterations beyond one
//   state read per block.
//
// Per system the steps, the launch partition (one site per thread in the XCD
// slab mapping, 256-thread blocks) and the order of every sum are those of
// gnx_cgm.hip's shape 1, so a system's solution, iteration count and exit
// are bit-identical to its single solve with that shape (while every system
// picks the same head-reduction form: all at most 512 blocks, or all more).  Vectors of all systems are concatenated in
// solver-owned buffers (the caller's sigma mass, b and x_in are copied in,
// the solution copied out); partial sums of system k live at offset poff_k
// of each of the two slots.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "gnx_cgstate.h"
#include "gnx_engine.h"
#include "gnx_internal.h"
#include "gnx_solver.h"

namespace gnx {

constexpr int CGU_BS = ENG_BS;

// Ordinary stores (the single-lattice FP64 kernels write non-temporally): the
// systems of an ensemble are small, their vectors stay in the 256 MB
// Infinity Cache for the next kernel (16 systems of 16^3-48^3: 2000 against
// 1900 RHS-solves/s, lockstep ensembles 81 against 78 trajectories/s).
GNX_DEVONLY void st_cgu(double2* p, double2 v) { *p = v; }

struct CGUFix {
  const CGUSys* sys;  // [n]
  double* sm;         // concatenated sigma masses
  double2* U;         // concatenated links, 3 per site (link-coupled solves), system k at 3 voff_k
  double2* b;         // concatenated right-hand sides
  double2* x;         // concatenated solutions (x_in for the null start)
  double2* r;
  double2* p0;
  double2* p1;
  double2* t;
  double* part;       // [2 slots][ptotal]
  CGMState* state;    // [2][n]
  CGDevResult* res;   // [n]
  long ptotal;
};

// the system of this block; false if the block lies past the system's sites
GNX_DEVONLY bool cgu_sys(const CGUFix& a, CGUSys* d) {
  *d = a.sys[blockIdx.y];
  return (int)blockIdx.x < d->nb;
}

// one site per thread in the XCD slab mapping of the single-lattice solver
// (its shape 1); the grid's x extent is a multiple of 8, so block x of every
// system runs on XCD x mod 8 and each system's slabs stay on their XCDs
GNX_DEVONLY int cgu_site(const CGUSys& d) {
  static_assert(CGU_BS == ENG_BS, "the site map assumes the engine's block size");
  return SiteMap<1, true>(d.nb).f(d.g, 0);
}

// ---- the caller's vectors in and out (one launch each way) ------------------------
// Per system the caller's sigma mass, right-hand side and (null start) x_in
// are gathered into the concatenated buffers, and the solution scattered
// back, by one kernel each instead of one copy per system and vector (three
// copies a system per solve, each its own dispatch).
__global__ __launch_bounds__(CGU_BS) void k_cgu_gather(CGUFix a, const CGUPtrs* __restrict__ ptr, int with_x) {
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int f = (int)blockIdx.x * CGU_BS + (int)threadIdx.x;
  if (f >= d.g.V) return;
  const CGUPtrs p = ptr[blockIdx.y];
  a.sm[d.soff + f] = p.sm[f];
  a.b[d.voff + f] = p.b[f];
  if (with_x) a.x[d.voff + f] = p.x[f];
  if (p.U) {
#pragma unroll
    for (int k = 0; k < 3; k++) a.U[3 * d.voff + 3 * f + k] = p.U[3 * f + k];
  }
}

__global__ __launch_bounds__(CGU_BS) void k_cgu_scatter(CGUFix a, const CGUPtrs* __restrict__ ptr) {
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int f = (int)blockIdx.x * CGU_BS + (int)threadIdx.x;
  if (f >= d.g.V) return;
  ptr[blockIdx.y].x[f] = a.x[d.voff + f];
}

// ---- start kernels (per solve, outside the graph) -------------------------------
// imposed start (conjugategradient.c:121-128): x = 0, r = p0 = b, partial |b|^2 (slot 1)
__global__ __launch_bounds__(CGU_BS) void k_cgu_start_imposed(CGUFix a, double tol, const long* __restrict__ max_iter) {
  __shared__ double lds[CGU_BS / 64];
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int f = cgu_site(d);
  double acc = 0.0;
  if (f >= 0) {
    const double2 bb = *elem(a.b + d.voff, f);
    *elem(a.x + d.voff, f) = make_double2(0.0, 0.0);
    *elem(a.r + d.voff, f) = bb;
    *elem(a.p0 + d.voff, f) = bb;
    acc = cnorm2(bb);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) a.part[a.ptotal + d.poff + partial_slot(blockIdx.x, d.nb)] = acc;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGMState s = {1, 0, 0, 0, 0.0, tol, max_iter[blockIdx.y], GNX_START_IMPOSED, 0};
    a.state[blockIdx.y] = s;
  }
}

// null start, pass 1 (conjugategradient.c:20): t = M^dag x_in
template <bool LINK>
__global__ __launch_bounds__(CGU_BS) void k_cgu_null_p1(CGUFix a) {
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int f = cgu_site(d);
  if (f < 0) return;
  const double2* xv = a.x + d.voff;
  const double smf = *elem(a.sm + d.soff, f);
  if (LINK)
    *elem(a.t + d.voff, f) = apply_link(d.g, decode(d.g, f), 1, smf, make_double2(0.0, 0.0), a.U + 3 * d.voff,
                                        [&](int q) { return *elem(xv, q); });
  else
    *elem(a.t + d.voff, f) = apply_Mdag(d.g, decode(d.g, f), smf, [&](int q) { return *elem(xv, q); });
}

// null start, pass 2 (conjugategradient.c:19-21): r = b - M t; partials |r|^2
// (slot 0) and |b|^2 (slot 1)
template <bool LINK>
__global__ __launch_bounds__(CGU_BS) void k_cgu_null_p2(CGUFix a) {
  __shared__ double lds[CGU_BS / 64];
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int f = cgu_site(d);
  double ar = 0.0, ab = 0.0;
  if (f >= 0) {
    const double2* t = a.t + d.voff;
    const double2 bb = *elem(a.b + d.voff, f);
    const double smf = *elem(a.sm + d.soff, f);
    const double2 rr =
        LINK ? apply_link(d.g, decode(d.g, f), 2, smf, bb, a.U + 3 * d.voff, [&](int q) { return *elem(t, q); })
             : apply_sub_M(d.g, decode(d.g, f), smf, bb, [&](int q) { return *elem(t, q); });
    *elem(a.r + d.voff, f) = rr;
    ar = cnorm2(rr);
    ab = cnorm2(bb);
  }
  ar = block_sum(ar, lds);
  ab = block_sum(ab, lds);
  if (threadIdx.x == 0) {
    a.part[d.poff + partial_slot(blockIdx.x, d.nb)] = ar;
    a.part[a.ptotal + d.poff + partial_slot(blockIdx.x, d.nb)] = ab;
  }
}

// null start, pass 3 (conjugategradient.c:22-33): an early exit keeps x_in;
// otherwise x = 0, r = p0 = b.  The |b|^2 partials stay in slot 1 for A(0).
__global__ __launch_bounds__(CGU_BS) void k_cgu_null_p3(CGUFix a, double tol, const long* __restrict__ max_iter) {
  __shared__ double lds[CGU_BS / 64];
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const double nr = block_sum(partial_share(a.part + d.poff, d.nb), lds);
  const bool early = sqrt(nr) < tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGMState s = {early ? 0 : 1, 0, early ? 1 : 0, early ? 1 : 0, 0.0, tol, max_iter[blockIdx.y], GNX_START_NULL, 0};
    a.state[blockIdx.y] = s;
    if (early) {
      CGDevResult r = {1, 1, 0, 0, sqrt(nr), {0, 0, 0}};
      a.res[blockIdx.y] = r;
    }
  }
  if (early) return;
  const int f = cgu_site(d);
  if (f < 0) return;
  const double2 bb = *elem(a.b + d.voff, f);
  *elem(a.x + d.voff, f) = make_double2(0.0, 0.0);
  *elem(a.r + d.voff, f) = bb;
  *elem(a.p0 + d.voff, f) = bb;
}

// ---- iteration kernels (the steps of k_cgm_A / k_cgm_B) -------------------------
// A: rho from the previous partials, exit test, beta; p_new = r + beta p_old at
// the site and its neighbours; t = M^dag p_new; partial |t|^2
template <bool WH, bool LINK>
__global__ __launch_bounds__(CGU_BS) void k_cgu_A(CGUFix a, int odd, const CGMState* __restrict__ sin,
                                                  CGMState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[CGU_BS / 64], lds2[CGU_BS / 64];
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int k = blockIdx.y;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const CGMState s = sin[k];
  CGMState* out = sout + k;
  if (!s.active) {  // uniform per block: a finished system loads nothing more
    if (lead) put_state(out, 0, s.iters, s.converged, s.early, s.rho, s);
    return;
  }
  const Geom g = d.g;
  const double2* r = a.r + d.voff;
  const double2* pold = (odd ? a.p1 : a.p0) + d.voff;
  double2* pnew = (odd ? a.p0 : a.p1) + d.voff;
  double pl[HEAD_W];
  head_issue<WH>(pin + d.poff, d.nb, pl);
  end_of_loads();
  const int f = cgu_site(d);
  const int fc = f < 0 ? 0 : f;  // past the last site: load site 0, store nothing
  const SiteGeo sgeo = site_geo(g, fc);
  const NbrF n = sgeo.n;
  const int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
  const double smv = *elem(a.sm + d.soff, fc);
  double2 Uo[3], Um[3];
  if (LINK) load_links(a.U + 3 * d.voff, fc, n, Uo, Um);
  double2 rv[7], pv[7];
#pragma unroll
  for (int j = 0; j < 7; j++) {
    rv[j] = *elem(r, q[j]);
    pv[j] = *elem(pold, q[j]);
  }
  end_of_loads();
  const double rho = head_finish<WH>(pin + d.poff, d.nb, pl, lds);
  double beta = 0.0;
  if (s.iters == 0) {
    if (s.start == GNX_START_IMPOSED && sqrt(rho) < s.tol) {  // conjugategradient.c:123-128
      if (lead) {
        put_state(out, 0, 0, 1, 1, rho, s);
        CGDevResult res = {1, 1, 0, 0, sqrt(rho), {0, 0, 0}};
        a.res[k] = res;
      }
      return;
    }
  } else {
    const bool conv = sqrt(rho) < s.tol;
    if (conv || s.iters >= s.max_iter) {
      if (lead) {
        put_state(out, 0, s.iters, conv ? 1 : 0, 0, rho, s);
        CGDevResult res = {conv ? 1 : 0, 0, s.iters, 0, sqrt(rho), {0, 0, 0}};
        a.res[k] = res;
      }
      return;
    }
    beta = rho / s.rho;
  }
  double acc = 0.0;
  if (f >= 0) {
    const SiteCode st = sgeo.k;
    const double2 p0v = cadd(rv[0], rmul(beta, pv[0]));
    double2 ppv[3], pmv[3];
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      ppv[dd] = cadd(rv[1 + dd], rmul(beta, pv[1 + dd]));
      pmv[dd] = cadd(rv[4 + dd], rmul(beta, pv[4 + dd]));
    }
    st_cgu(elem(pnew, f), p0v);
    const double2 tt = LINK ? link_hops<-1>(st, rmul(smv, p0v), ppv, pmv, Uo, Um)
                            : apply_Mdag_fast_vals(st, smv, p0v, ppv, pmv);
    st_cgu(elem(a.t + d.voff, f), tt);
    acc = cnorm2(tt);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[d.poff + partial_slot(blockIdx.x, d.nb)] = acc;
  if (lead) put_state(out, 1, s.iters, 0, 0, rho, s);
}

// B: gamma from A's partials; alpha = rho / gamma; x += alpha p_new;
// r -= alpha M t; partial |r|^2
template <bool WH, bool LINK>
__global__ __launch_bounds__(CGU_BS) void k_cgu_B(CGUFix a, int odd, const CGMState* __restrict__ sin,
                                                  CGMState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[CGU_BS / 64], lds2[CGU_BS / 64];
  CGUSys d;
  if (!cgu_sys(a, &d)) return;
  const int k = blockIdx.y;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const CGMState s = sin[k];
  CGMState* out = sout + k;
  if (!s.active) {
    if (lead) put_state(out, 0, s.iters, s.converged, s.early, s.rho, s);
    return;
  }
  const Geom g = d.g;
  const double2* pnew = (odd ? a.p0 : a.p1) + d.voff;
  const double2* t = a.t + d.voff;
  double2* r = a.r + d.voff;
  double2* x = a.x + d.voff;
  double pl[HEAD_W];
  head_issue<WH>(pin + d.poff, d.nb, pl);
  end_of_loads();
  const int f = cgu_site(d);
  const int fc = f < 0 ? 0 : f;
  const SiteGeo sgeo = site_geo(g, fc);
  const NbrF n = sgeo.n;
  const int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
  const double smv = *elem(a.sm + d.soff, fc);
  double2 Uo[3], Um[3];
  if (LINK) load_links(a.U + 3 * d.voff, fc, n, Uo, Um);
  double2 tv[7];
#pragma unroll
  for (int j = 0; j < 7; j++) tv[j] = *elem(t, q[j]);
  const double2 pv = *elem(pnew, fc);
  const double2 xv = *elem(x, fc);
  const double2 rv = *elem(r, fc);
  end_of_loads();
  const double gamma = head_finish<WH>(pin + d.poff, d.nb, pl, lds);
  const double al = s.rho / gamma;
  double acc = 0.0;
  if (f >= 0) {
    const SiteCode st = sgeo.k;
    const double2 w = LINK ? link_hops<1>(st, rmul(smv, tv[0]), &tv[1], &tv[4], Uo, Um)
                           : apply_M_vals(st, smv, tv[0], &tv[1], &tv[4]);
    st_cgu(elem(x, f), cadd(xv, rmul(al, pv)));
    const double2 rr = cadd(rv, rmul(-al, w));
    st_cgu(elem(r, f), rr);
    acc = cnorm2(rr);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[d.poff + partial_slot(blockIdx.x, d.nb)] = acc;
  if (lead) put_state(out, 1, s.iters + 1, 0, 0, s.rho, s);
}

}  // namespace gnx

// ---- host side ----------------------------------------------------------------
struct gnx_multi {
  int device = 0;
  int n = 0;
  hipStream_t stream = nullptr;
  std::vector<gnx::CGUSys> sys;  // host copy of the descriptors
  long vtotal = 0;               // complex elements over all systems
  long stotal = 0;               // sigma-mass doubles over all systems
  long ptotal = 0;               // partials per slot over all systems
  int nbmax = 0;
  gnx::GraphEngine eng;
  gnx::CGUSys* dsys = nullptr;
  long* dmax_iter = nullptr;     // [n]
  long* hmax_iter = nullptr;     // [n], pinned
  gnx::CGUPtrs* dptr = nullptr;  // [n] the caller's vectors of the current solve
  gnx::CGUPtrs* hptr = nullptr;  // [n], pinned
  double* sm = nullptr;
  double2 *b = nullptr, *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *t = nullptr;
  double2* U = nullptr;  // concatenated links, allocated by the first link-coupled solve
  void* mixed = nullptr;  // the mixed-precision solver's buffers (gnx_cgmix.hip), on first use
  gnx::CGDevResult* res = nullptr;
  gnx::CGDevResult* hres = nullptr;  // pinned
};

namespace gnx {
// systems of an uneven solver and their lattices (for the ensemble drivers)
int multi_count(const gnx_multi* m) { return m ? m->n : 0; }
bool multi_system(const gnx_multi* m, int k, int* L, int* LT, int* device) {
  if (!m || k < 0 || k >= m->n) return false;
  *L = m->sys[k].g.L;
  *LT = m->sys[k].g.LT;
  *device = m->device;
  return true;
}
bool multi_view(gnx_multi* m, MultiView* v) {
  if (!m) return false;
  v->n = m->n;
  v->device = m->device;
  v->stream = m->stream;
  v->dsys = m->dsys;
  v->hsys = m->sys.data();
  v->vtotal = m->vtotal;
  v->stotal = m->stotal;
  v->ptotal = m->ptotal;
  v->nbmax = m->nbmax;
  v->mixed = &m->mixed;
  return true;
}
}  // namespace gnx

namespace gnx {

static void cgu_free(gnx_multi* m) {
  void* dev[] = {m->dsys, m->dmax_iter, m->dptr, m->sm, m->b, m->x, m->r, m->p0, m->p1, m->t, m->U, m->res};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (m->hmax_iter) (void)hipHostFree(m->hmax_iter);
  if (m->hptr) (void)hipHostFree(m->hptr);
  if (m->hres) (void)hipHostFree(m->hres);
  if (m->mixed) cgmix_multi_release(m->mixed);
  m->mixed = nullptr;
}

static CGUFix cgu_fix(gnx_multi* m) {
  CGUFix a;
  a.sys = m->dsys;
  a.sm = m->sm;
  a.U = m->U;
  a.b = m->b;
  a.x = m->x;
  a.r = m->r;
  a.p0 = m->p0;
  a.p1 = m->p1;
  a.t = m->t;
  a.part = m->eng.dev_part();
  a.state = (CGMState*)m->eng.dev_state();
  a.res = m->res;
  a.ptotal = m->ptotal;
  return a;
}

// the chunk and tail graphs: (A, B) pairs over all systems, on M M^dag or
// on M_U M_U^dag (link)
static hipError_t cgu_graph(gnx_multi* m, bool link, EngGraph* exec) {
  CGUFix a = cgu_fix(m);
  const int n = m->n;
  dim3 grid((m->nbmax + NUM_XCD - 1) / NUM_XCD * NUM_XCD, n), blk(CGU_BS);
  const CGMState* s0 = a.state;
  const CGMState* s1 = a.state + n;
  CGMState* w0 = a.state;
  CGMState* w1 = a.state + n;
  const double* q0 = a.part;
  const double* q1 = a.part + a.ptotal;
  double* o0 = a.part;
  double* o1 = a.part + a.ptotal;
  auto body = [&](hipStream_t cs, int iters) {
    for (int i = 0; i < iters; i++) {
      int odd = i & 1;
      // A: state[0] -> state[1], reduces slot 1, writes slot 0; B the reverse
      void* argsA[] = {(void*)&a, (void*)&odd, (void*)&s0, (void*)&w1, (void*)&q1, (void*)&o0};
      void* argsB[] = {(void*)&a, (void*)&odd, (void*)&s1, (void*)&w0, (void*)&q0, (void*)&o1};
      // the per-wave head when every system has at most 512 partials (the
      // form a single-lattice solve of each system then also uses)
      const bool wh = head_wave(m->nbmax);
      const void* A = link ? (wh ? (const void*)k_cgu_A<true, true> : (const void*)k_cgu_A<false, true>)
                           : (wh ? (const void*)k_cgu_A<true, false> : (const void*)k_cgu_A<false, false>);
      const void* B = link ? (wh ? (const void*)k_cgu_B<true, true> : (const void*)k_cgu_B<false, true>)
                           : (wh ? (const void*)k_cgu_B<true, false> : (const void*)k_cgu_B<false, false>);
      (void)hipLaunchKernel(A, grid, blk, argsA, 0, cs);
      (void)hipLaunchKernel(B, grid, blk, argsB, 0, cs);
    }
  };
  return m->eng.graph(link ? 1 : 0, n, body, exec);
}

}  // namespace gnx

extern "C" {

gnx_status gnx_multi_create(int n, const int* lspace, const int* ltime, int device, gnx_multi** out) {
  using namespace gnx;
  CHECK_ARG(out != nullptr && lspace != nullptr && ltime != nullptr, "gnx_multi_create: null argument");
  *out = nullptr;
  CHECK_ARG(n >= 1 && n <= 1024, "gnx_multi_create: n must be in 1..1024");
  for (int k = 0; k < n; k++) {
    CHECK_ARG(lspace[k] >= 4 && ltime[k] >= 4 && lspace[k] % 2 == 0 && ltime[k] % 2 == 0,
              "gnx_multi_create: every L and LT must be even and >= 4");
    if ((long)lspace[k] * lspace[k] * ltime[k] > (1L << 26)) {
      set_err("gnx_multi_create: system %d exceeds 2^26 sites", k);
      return GNX_ERR_SIZE;
    }
    if (!geom_decodable(lspace[k], ltime[k])) {
      set_err("gnx_multi_create: system %d (%d^2 x %d) is outside the exact range of the in-kernel site decode", k,
              lspace[k], ltime[k]);
      return GNX_ERR_SIZE;
    }
  }
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  CHECK_ARG(device >= 0 && device < ndev, "gnx_multi_create: no such device");
  HIP_TRY(hipSetDevice(device));
  gnx_multi* m = new gnx_multi();
  m->device = device;
  m->n = n;
  m->sys.resize(n);
  for (int k = 0; k < n; k++) {
    CGUSys& d = m->sys[k];
    d.g = make_geom(lspace[k], ltime[k]);
    d.nb = (d.g.V + CGU_BS - 1) / CGU_BS;
    d.voff = m->vtotal;
    d.soff = m->stotal;
    d.poff = m->ptotal;
    m->vtotal += d.g.V;
    m->stotal += d.g.V;
    m->ptotal += d.nb;
    if (d.nb > m->nbmax) m->nbmax = d.nb;
  }
  const size_t vb = sizeof(double2) * (size_t)m->vtotal;
  hipError_t e = m->eng.ensure(n, 1, sizeof(CGMState), 2 * (size_t)m->ptotal);
  if (e == hipSuccess) e = hipMalloc(&m->dsys, sizeof(CGUSys) * n);
  if (e == hipSuccess) e = hipMalloc(&m->dmax_iter, sizeof(long) * n);
  if (e == hipSuccess) e = hipHostMalloc(&m->hmax_iter, sizeof(long) * n, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&m->dptr, sizeof(gnx::CGUPtrs) * n);
  if (e == hipSuccess) e = hipHostMalloc(&m->hptr, sizeof(gnx::CGUPtrs) * n, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&m->sm, sizeof(double) * (size_t)m->stotal);
  if (e == hipSuccess) e = hipMalloc(&m->b, vb);
  if (e == hipSuccess) e = hipMalloc(&m->x, vb);
  if (e == hipSuccess) e = hipMalloc(&m->r, vb);
  if (e == hipSuccess) e = hipMalloc(&m->p0, vb);
  if (e == hipSuccess) e = hipMalloc(&m->p1, vb);
  if (e == hipSuccess) e = hipMalloc(&m->t, vb);
  if (e == hipSuccess) e = hipMalloc(&m->res, sizeof(CGDevResult) * n);
  if (e == hipSuccess) e = hipHostMalloc(&m->hres, sizeof(CGDevResult) * n, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemcpy(m->dsys, m->sys.data(), sizeof(CGUSys) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_err("gnx_multi_create: allocation failed: %s", hipGetErrorString(e));
    cgu_free(m);
    delete m;
    return GNX_ERR_HIP;
  }
  *out = m;
  return GNX_OK;
}

gnx_status gnx_multi_destroy(gnx_multi* m) {
  if (!m) return GNX_OK;
  (void)hipSetDevice(m->device);
  (void)hipDeviceSynchronize();
  gnx::cgu_free(m);
  delete m;
  return GNX_OK;
}

gnx_status gnx_multi_set_stream(gnx_multi* m, void* stream) {
  CHECK_ARG(m != nullptr, "gnx_multi_set_stream: null handle");
  m->stream = (hipStream_t)stream;
  return GNX_OK;
}

}  // extern "C"

namespace gnx {

// one uneven-batch solve; U = null: M M^dag, else M_U M_U^dag with system
// k's links U[k] (3 per site, f-order)
static gnx_status multi_cg(gnx_multi* m, const double* const* sm, const double* const* Ul, const double* const* b,
                           double* const* x, double tol, long max_iter, gnx_cg_start start, gnx_cg_result* res) {
  CHECK_ARG(m && sm && b && x, "gnx_multi_cg: null argument");
  CHECK_ARG(tol >= 0.0, "gnx_multi_cg: negative tolerance");
  CHECK_ARG(start == GNX_START_NULL || start == GNX_START_IMPOSED, "gnx_multi_cg: bad start");
  for (int k = 0; k < m->n; k++) {
    CHECK_ARG(sm[k] && b[k] && x[k], "gnx_multi_cg: null vector");
    CHECK_ARG((const void*)b[k] != (const void*)x[k], "gnx_multi_cg: b and x alias");
    CHECK_ARG(!Ul || Ul[k], "gnx_multi_link_cg: null links");
  }
  const bool link = Ul != nullptr;
  HIP_TRY(hipSetDevice(m->device));
  if (link && !m->U) {
    // the links' buffer; captured graphs point at the old arguments
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMalloc(&m->U, sizeof(double2) * 3 * (size_t)m->vtotal));
    m->eng.invalidate_graphs();
  }
  hipStream_t st = m->stream;
  // the pinned tables below are rewritten here: no copy of an earlier call
  // (one that returned early on an error) may still be reading them
  HIP_TRY(hipStreamSynchronize(st));
  const int n = m->n;
  long cap_max = 0;
  for (int k = 0; k < n; k++) {
    const CGUSys& d = m->sys[k];
    // the reference's cap V+1 (macro.h:134) of each system when max_iter <= 0
    m->hmax_iter[k] = max_iter > 0 ? max_iter : (long)d.g.V + 1;
    if (m->hmax_iter[k] > cap_max) cap_max = m->hmax_iter[k];
    m->hptr[k] = CGUPtrs{sm[k], (const double2*)b[k], (double2*)x[k], link ? (const double2*)Ul[k] : nullptr};
  }
  HIP_TRY(hipMemcpyAsync(m->dmax_iter, m->hmax_iter, sizeof(long) * n, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(m->dptr, m->hptr, sizeof(CGUPtrs) * n, hipMemcpyHostToDevice, st));
  CGUFix a = cgu_fix(m);
  dim3 grid((m->nbmax + NUM_XCD - 1) / NUM_XCD * NUM_XCD, n), blk(CGU_BS);
  {
    const CGUPtrs* dp = m->dptr;
    int with_x = start == GNX_START_NULL ? 1 : 0;
    void* ag[] = {(void*)&a, (void*)&dp, (void*)&with_x};
    HIP_TRY(hipLaunchKernel((const void*)k_cgu_gather, grid, blk, ag, 0, st));
  }
  const long* mi = m->dmax_iter;
  if (start == GNX_START_NULL) {
    void* a1[] = {(void*)&a};
    HIP_TRY(hipLaunchKernel(link ? (const void*)k_cgu_null_p1<true> : (const void*)k_cgu_null_p1<false>, grid, blk,
                            a1, 0, st));
    HIP_TRY(hipLaunchKernel(link ? (const void*)k_cgu_null_p2<true> : (const void*)k_cgu_null_p2<false>, grid, blk,
                            a1, 0, st));
    void* a3[] = {(void*)&a, (void*)&tol, (void*)&mi};
    HIP_TRY(hipLaunchKernel((const void*)k_cgu_null_p3, grid, blk, a3, 0, st));
  } else {
    void* a0[] = {(void*)&a, (void*)&tol, (void*)&mi};
    HIP_TRY(hipLaunchKernel((const void*)k_cgu_start_imposed, grid, blk, a0, 0, st));
  }
  EngGraph exec;
  HIP_TRY(cgu_graph(m, link, &exec));
  {
    // the solutions scattered to the callers' vectors and the results
    // copied after each burst (the gathered x_in were read by the start
    // kernels)
    const CGUPtrs* dp = m->dptr;
    const FinishFn out = [&](hipStream_t s) {
      void* as[] = {(void*)&a, (void*)&dp};
      hipError_t f = hipLaunchKernel((const void*)k_cgu_scatter, grid, blk, as, 0, s);
      if (f == hipSuccess) f = hipMemcpyAsync(m->hres, m->res, sizeof(CGDevResult) * n, hipMemcpyDeviceToHost, s);
      return f;
    };
    HIP_TRY(m->eng.run(exec, n, cap_max, st, nullptr, eng_key(link ? 1 : 0, n, tol, start), nullptr, &out));
  }
  int all = 1;
  for (int k = 0; k < n; k++) {
    if (res) {
      res[k].converged = m->hres[k].converged;
      res[k].early_exit = m->hres[k].early_exit;
      res[k].iterations = m->hres[k].iterations;
      res[k].reliable_updates = 0;
      res[k].residual = m->hres[k].residual;
    }
    all &= m->hres[k].converged;
  }
  if (!all) {
    set_err("gnx_multi_cg: a system reached its iteration cap");
    return GNX_ERR_NOCONV;
  }
  return GNX_OK;
}

}  // namespace gnx

extern "C" {

gnx_status gnx_multi_cg(gnx_multi* m, const double* const* sm, const double* const* b, double* const* x, double tol,
                        long max_iter, gnx_cg_start start, gnx_cg_result* res) {
  return gnx::multi_cg(m, sm, nullptr, b, x, tol, max_iter, start, res);
}

gnx_status gnx_multi_link_cg(gnx_multi* m, const double* const* sm, const double* const* U, const double* const* b,
                             double* const* x, double tol, long max_iter, gnx_cg_start start, gnx_cg_result* res) {
  CHECK_ARG(U != nullptr, "gnx_multi_link_cg: null links");
  return gnx::multi_cg(m, sm, U, b, x, tol, max_iter, start, res);
}

gnx_status gnx_multi_cg_prec(gnx_multi* m, const double* const* sm, const double* const* b, double* const* x,
                             double tol, long max_iter, gnx_cg_start start, gnx_prec prec, gnx_cg_result* res) {
  // the reliable-update factor of a context that was never set (gnx_set_reliable_delta)
  return gnx::multi_cg_mixed(m, sm, b, x, tol, max_iter, start, (int)prec, prec == GNX_PREC_BF16 ? 0.2 : 0.05, res);
}

}  // extern "C"

gnx_status gnx::multi_cg_mixed(gnx_multi* m, const double* const* sm, const double* const* b, double* const* x,
                               double tol, long max_iter, int start, int prec, double delta, gnx_cg_result* res) {
  using namespace gnx;
  CHECK_ARG(prec == GNX_PREC_F64 || prec == GNX_PREC_F32 || prec == GNX_PREC_BF16 || prec == GNX_PREC_F16,
            "gnx_multi_cg_prec: bad precision");
  if (prec == GNX_PREC_F64) return multi_cg(m, sm, nullptr, b, x, tol, max_iter, (gnx_cg_start)start, res);
  CHECK_ARG(delta > 0.0 && delta < 1.0, "gnx_multi_cg_prec: the reliable-update factor must be in (0, 1)");
  CHECK_ARG(m && sm && b && x, "gnx_multi_cg_prec: null argument");
  CHECK_ARG(tol >= 0.0, "gnx_multi_cg_prec: negative tolerance");
  CHECK_ARG(start == GNX_START_NULL || start == GNX_START_IMPOSED, "gnx_multi_cg_prec: bad start");
  for (int k = 0; k < m->n; k++) {
    CHECK_ARG(sm[k] && b[k] && x[k], "gnx_multi_cg_prec: null vector");
    CHECK_ARG((const void*)b[k] != (const void*)x[k], "gnx_multi_cg_prec: b and x alias");
  }
  HIP_TRY(hipSetDevice(m->device));
  MultiView mv;
  multi_view(m, &mv);
  HIP_TRY(cgmix_multi_run(mv, sm, b, x, tol, delta, max_iter, start, prec, m->hres));
  int all = 1;
  for (int k = 0; k < m->n; k++) {
    if (res) {
      res[k].converged = m->hres[k].converged;
      res[k].early_exit = m->hres[k].early_exit;
      res[k].iterations = m->hres[k].iterations;
      res[k].reliable_updates = m->hres[k].reliable_updates;
      res[k].residual = m->hres[k].residual;
    }
    all &= m->hres[k].converged;
  }
  if (!all) {
    set_err("gnx_multi_cg_prec: a system reached its iteration cap");
    return GNX_ERR_NOCONV;
  }
  return GNX_OK;
}
