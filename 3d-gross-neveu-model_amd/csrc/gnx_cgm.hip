// gnx_cgm.hip -- FP64 conjugate gradient on A = M M^dag (conjugategradient.c:13-182)
// as a device-side state machine of two kernels per iteration, replayed from
// a captured hipGraph.
//
// Why kernels and not a persistent kernel: on gfx950 a grid-wide barrier
// inside a kernel costs 10-80 us (cooperative grid sync or an atomic counter,
// growing with the block count), a kernel node inside a hipGraph about
// 1.7 us, and even the neighbour-face and partial-sum traffic of one
// persistent iteration alone, with per-tile flags and no barrier or L2
// fence, costs 8.3 us at 64^3 (the halo-exchange probes of gnx_probes.hip).
// Standard CG needs two global reductions per iteration, so each
// iteration is two kernels whose boundaries are the barriers:
//
//   A(k): rho_k from the partials of the previous B (or of the start kernel);
//         exit test sqrt(rho_k) < tol (absolute, conjugategradient.c:87-89);
//         beta = rho_k / rho_{k-1} (0 at k = 0);
//         p_new = r + beta p_old at the site and, recomputed, at its six
//         neighbours; t = M^dag p_new; partial |t|^2.
//   B(k): gamma from A's partials; alpha = rho_k / gamma; x += alpha p_new;
//         r += (-alpha) M t; partial |r|^2.
//
// The per-site updates are the reference's (x += alpha p, r -= alpha M t,
// p = r + beta p, M in reference order); M^dag is applied in its gather form
// (sm v minus the hops, no reordering selects) and the global sums are taken
// in a fixed tree order, so solutions agree with the reference to the
// solver tolerance.  Every block reduces the previous kernel's
// per-block partials in the same order, so all blocks hold bit-identical
// scalars.  Per-system state is double-buffered: A reads state[0] and its
// block 0 writes state[1]; B the reverse; no block reads what another block
// of the same kernel writes.  Systems of a batch (blockIdx.y) are
// independent: each keeps its alpha, beta and exit iteration.
//
// Latency: each kernel issues every load it needs before its first wait --
// the state and the partials first, then the stencil operands, all vector
// loads -- and reduces the partials while the operands are in flight.  Kernel arguments are passed by value and point at
// buffers this solver owns (the caller's sigma mass and x are copied in and
// the solution copied out), and the parity of p_old/p_new is baked into each
// graph node (one slot pair = one iteration, a chunk holds an even count).
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "gnx_cgstate.h"
#include "gnx_engine.h"
#include "gnx_internal.h"
#include "gnx_solver.h"
#include "gnx_tile.h"

namespace gnx {

constexpr int CGM_BS = ENG_BS;

// kernel arguments: fixed per solver instance, passed by value
struct CGMFix {
  Geom g;
  double* sm;        // copy of the caller's sigma mass
  double2* U;        // copy of the caller's links (link-coupled operator only)
  double2* x;        // solutions [nrhs][V]
  double2* r;
  double2* p0;
  double2* p1;
  double2* t;
  double* part;      // [2 slots][nrhs][nb]
  CGMState* state;   // [2][nrhs]
  CGDevResult* res;  // [nrhs]
  double2* rring;    // deferred form: CGD_R residual buffers [CGD_R][nrhs][V], the first one r
  size_t rstride;    // elements between two of them
  double2* t1;       // fused form: the second q buffer (q_k in t or t1 by the parity of k)
  double2* s1;       // fused form: the second s buffer (s_k in p1 or s1)
  int nrhs;
  int nb;            // blocks per system of every kernel of this shape
};

// residual buffer of ring slot j (the deferred form)
GNX_DEV double2* ring(const CGMFix& a, int j) { return a.rring + (size_t)j * a.rstride; }

GNX_DEV double* slot(const CGMFix& a, int s, int rhs) {
  return a.part + ((size_t)s * a.nrhs + rhs) * a.nb;
}

// the one-reduction form's partials: (|r|^2, |q|^2) pairs, [nrhs][nb], in
// the space of the two slots
GNX_DEV double2* pair_slots(const CGMFix& a, int rhs) {
  return (double2*)a.part + (size_t)rhs * a.nb;
}

// ---- start kernels (launched per solve, outside the graph) ------------------
// imposed start (conjugategradient.c:121-128): x = 0, r = p0 = b, partial |b|^2
template <int SPT, bool XCD>
__global__ __launch_bounds__(CGM_BS) void k_cgm_start_imposed(CGMFix a, const double* __restrict__ sm_in,
                                                               const double2* __restrict__ b, double tol,
                                                               long max_iter) {
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const SiteMap<SPT, XCD> map(a.nb);
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(a.g, i);
    if (f < 0) break;
    double2 bb = b[off + f];
    a.x[off + f] = make_double2(0.0, 0.0);
    a.r[off + f] = bb;
    a.p0[off + f] = bb;
    if (rhs == 0) a.sm[f] = sm_in[f];
    acc += cnorm2(bb);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) slot(a, 1, rhs)[partial_slot(blockIdx.x, a.nb)] = acc;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGMState s = {1, 0, 0, 0, 0.0, tol, max_iter, GNX_START_IMPOSED, 0};
    a.state[rhs] = s;
  }
}

// null start, pass 1 (conjugategradient.c:20): t = M^dag x_in; x = x_in; sm copied
template <int SPT, bool XCD, bool LINK>
__global__ __launch_bounds__(CGM_BS) void k_cgm_null_p1(CGMFix a, const double* __restrict__ sm_in,
                                                         const double2* __restrict__ xin) {
  const size_t off = (size_t)blockIdx.y * a.g.V;
  const SiteMap<SPT, XCD> map(a.nb);
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(a.g, i);
    if (f < 0) break;
    const double2* xv = xin + off;
    if (LINK)
      a.t[off + f] = apply_link(a.g, decode(a.g, f), 1, sm_in[f], make_double2(0.0, 0.0), a.U,
                                [&](int q) { return xv[q]; });
    else
      a.t[off + f] = apply_Mdag(a.g, decode(a.g, f), sm_in[f], [&](int q) { return xv[q]; });
    a.x[off + f] = xv[f];
    if (blockIdx.y == 0) a.sm[f] = sm_in[f];
  }
}

// null start, pass 2 (conjugategradient.c:19-21): r = b - M t; partials |r|^2
// (slot 0) and |b|^2 (slot 1)
template <int SPT, bool XCD, bool LINK>
__global__ __launch_bounds__(CGM_BS) void k_cgm_null_p2(CGMFix a, const double2* __restrict__ b) {
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const SiteMap<SPT, XCD> map(a.nb);
  double ar = 0.0, ab = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(a.g, i);
    if (f < 0) break;
    const double2* t = a.t + off;
    double2 bb = b[off + f];
    double2 rr = LINK ? apply_link(a.g, decode(a.g, f), 2, a.sm[f], bb, a.U, [&](int q) { return t[q]; })
                      : apply_sub_M(a.g, decode(a.g, f), a.sm[f], bb, [&](int q) { return t[q]; });
    a.r[off + f] = rr;
    ar += cnorm2(rr);
    ab += cnorm2(bb);
  }
  ar = block_sum(ar, lds);
  ab = block_sum(ab, lds);
  if (threadIdx.x == 0) {
    slot(a, 0, rhs)[partial_slot(blockIdx.x, a.nb)] = ar;
    slot(a, 1, rhs)[partial_slot(blockIdx.x, a.nb)] = ab;
  }
}

// null start, pass 3 (conjugategradient.c:22-33): an early exit keeps x_in;
// otherwise x = 0, r = p0 = b.  The |b|^2 partials stay in slot 1 for A(0).
template <int SPT, bool XCD>
__global__ __launch_bounds__(CGM_BS) void k_cgm_null_p3(CGMFix a, const double2* __restrict__ b, double tol,
                                                         long max_iter) {
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const double nr = block_sum(partial_share(slot(a, 0, rhs), a.nb), lds);
  const bool early = sqrt(nr) < tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGMState s = {early ? 0 : 1, 0, early ? 1 : 0, early ? 1 : 0, 0.0, tol, max_iter, GNX_START_NULL, 0};
    a.state[rhs] = s;
    if (early) {
      CGDevResult r = {1, 1, 0, 0, sqrt(nr), {0, 0, 0}};
      a.res[rhs] = r;
    }
  }
  if (early) return;
  const SiteMap<SPT, XCD> map(a.nb);
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(a.g, i);
    if (f < 0) break;
    double2 bb = b[off + f];
    a.x[off + f] = make_double2(0.0, 0.0);
    a.r[off + f] = bb;
    a.p0[off + f] = bb;
  }
}

// ---- iteration kernels -------------------------------------------------------
// odd = iteration parity: p_old = odd ? p1 : p0, p_new the other one
// sin/sout: this kernel's input and output state halves; pin/pout: the slot
// it reduces and the slot it writes (separate restrict pointers, so the
// compiler may keep the reads ahead of the writes)
template <int SPT, bool XCD, bool LINK, bool WH>
__global__ __launch_bounds__(CGM_BS) void k_cgm_A(CGMFix a, int odd, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout, const double* __restrict__ pin,
                                                   double* __restrict__ pout) {
  __shared__ double lds[CGM_BS / 64], lds2[CGM_BS / 64];  // head and tail reductions
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double2* r = a.r + off;
  const double2* pold = (odd ? a.p1 : a.p0) + off;
  double2* pnew = (odd ? a.p0 : a.p1) + off;
  const SiteMap<SPT, XCD> map(a.nb);
  // loads in the order they are needed: state and partials, then the
  // operands; the reduction waits for the first group only (loads return in
  // issue order) and runs while the operands are in flight
  const CGMState s = load_state(sin + rhs);
  double pl[HEAD_W];
  head_issue<WH>(pin + (size_t)rhs * a.nb, a.nb, pl);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  double smv[SPT];
  double2 rv[SPT][7], pv[SPT][7];
  double2 Uo[LINK ? SPT : 1][3], Um[LINK ? SPT : 1][3];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = map.f(g, i);
    // a thread past the last site loads site 0 and stores nothing: the loads
    // stay one branch-free block, so none of them waits for another
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fc);
    if (LINK) load_links(a.U, fc, n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) {
      rv[i][k] = *elem(r, q[k]);
      pv[i][k] = *elem(pold, q[k]);
    }
  }
  end_of_loads();
  const double rho = head_finish<WH>(pin + (size_t)rhs * a.nb, a.nb, pl, lds);
  CGMState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_state(out, 0, s.iters, s.converged, s.early, s.rho, s);
    return;
  }
  double beta = 0.0;
  if (s.iters == 0) {
    // ||b|| < tol returns before the loop only in the imposed start
    // (conjugategradient.c:123-128); the null start tested its prologue
    // residual instead (:22)
    if (s.start == GNX_START_IMPOSED && sqrt(rho) < s.tol) {
      if (lead) {
        put_state(out, 0, 0, 1, 1, rho, s);
        CGDevResult res = {1, 1, 0, 0, sqrt(rho), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
  } else {
    const bool conv = sqrt(rho) < s.tol;
    if (conv || s.iters >= s.max_iter) {
      if (lead) {
        put_state(out, 0, s.iters, conv ? 1 : 0, 0, rho, s);
        CGDevResult res = {conv ? 1 : 0, 0, s.iters, 0, sqrt(rho), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    beta = rho / s.rho;
  }
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const SiteCode st = kc[i];
    // p_new at the site and its neighbours, from the preloaded r and p_old
    const double2 p0v = cadd(rv[i][0], rmul(beta, pv[i][0]));
    double2 ppv[3], pmv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      ppv[d] = cadd(rv[i][1 + d], rmul(beta, pv[i][1 + d]));
      pmv[d] = cadd(rv[i][4 + d], rmul(beta, pv[i][4 + d]));
    }
    st_nt(elem(pnew, fs[i]), p0v);
    double2 tt = LINK ? link_hops<-1>(st, rmul(smv[i], p0v), ppv, pmv, Uo[LINK ? i : 0], Um[LINK ? i : 0])
                      : apply_Mdag_fast_vals(st, smv[i], p0v, ppv, pmv);
    st_nt(elem(a.t + off, fs[i]), tt);
    acc += cnorm2(tt);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[(size_t)rhs * a.nb + partial_slot(blockIdx.x, a.nb)] = acc;
  if (lead) put_state(out, 1, s.iters, 0, 0, rho, s);
}

template <int SPT, bool XCD, bool LINK, bool WH>
__global__ __launch_bounds__(CGM_BS) void k_cgm_B(CGMFix a, int odd, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout, const double* __restrict__ pin,
                                                   double* __restrict__ pout) {
  __shared__ double lds[CGM_BS / 64], lds2[CGM_BS / 64];  // head and tail reductions
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double2* pnew = (odd ? a.p0 : a.p1) + off;
  const double2* t = a.t + off;
  double2* r = a.r + off;
  double2* x = a.x + off;
  const SiteMap<SPT, XCD> map(a.nb);
  // loads in the order they are needed (as in A)
  const CGMState s = load_state(sin + rhs);
  double pl[HEAD_W];
  head_issue<WH>(pin + (size_t)rhs * a.nb, a.nb, pl);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  double smv[SPT];
  double2 tv[SPT][7], pv[SPT], xv[SPT], rv[SPT];
  double2 Uo[LINK ? SPT : 1][3], Um[LINK ? SPT : 1][3];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = map.f(g, i);
    // a thread past the last site loads site 0 and stores nothing: the loads
    // stay one branch-free block, so none of them waits for another
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fc);
    if (LINK) load_links(a.U, fc, n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) tv[i][k] = *elem(t, q[k]);
    pv[i] = *elem(pnew, fc);
    xv[i] = *elem(x, fc);
    rv[i] = *elem(r, fc);
  }
  end_of_loads();
  const double gamma = head_finish<WH>(pin + (size_t)rhs * a.nb, a.nb, pl, lds);
  CGMState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_state(out, 0, s.iters, s.converged, s.early, s.rho, s);
    return;
  }
  const double al = s.rho / gamma;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    // M t from the preloaded operands, the order of apply_M / apply_link
    const SiteCode st = kc[i];
    const double2 m0 = rmul(smv[i], tv[i][0]);
    const double2 w = LINK ? link_hops<1>(st, m0, &tv[i][1], &tv[i][4], Uo[LINK ? i : 0], Um[LINK ? i : 0])
                           : apply_M_vals(st, smv[i], tv[i][0], &tv[i][1], &tv[i][4]);
    st_nt(elem(x, fs[i]), cadd(xv[i], rmul(al, pv[i])));
    double2 rr = cadd(rv[i], rmul(-al, w));
    st_nt(elem(r, fs[i]), rr);
    acc += cnorm2(rr);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[(size_t)rhs * a.nb + partial_slot(blockIdx.x, a.nb)] = acc;
  if (lead) put_state(out, 1, s.iters + 1, 0, 0, s.rho, s);
}

// ---- one-reduction form (Chronopoulos-Gear CG) ------------------------------
// The same iterates as the kernels above in exact arithmetic, with both dot
// products of an iteration taken at one point.  With A = M M^dag, s = A p
// kept as a vector and q = M^dag r:
//
//   X(k): q = M^dag r; partials (|r|^2, |q|^2 = (r, A r)) as one pair per
//         block -- a stencil with no reduction ahead of it.
//   Y(k): gamma_k = |r|^2 and delta_k = (r, A r) from X's partials; exit
//         test sqrt(gamma_k) < tol (or the iteration cap); beta_k = gamma_k /
//         gamma_{k-1}, alpha_k = gamma_k / (delta_k - beta_k gamma_k /
//         alpha_{k-1}) (beta_0 = 0, alpha_0 = gamma_0 / delta_0); then, at
//         the site only: s = M q + beta s, p = r + beta p, x += alpha p,
//         r -= alpha s.
//
// One head reduction per iteration instead of two, 18 operand loads per site
// instead of 24, and no vector is read at a neighbour by the kernel that
// writes it, so r, s, p, x are updated in place.  Per-site arithmetic: x
// and p as in conjugategradient.c (x += alpha p, p = r + beta p); r by the
// recurrence r -= alpha s instead of r -= alpha M t.  Buffers: q in t, s in
// p1.  The start kernels are the two-reduction form's (x, r = b); their
// partials are not read.  X reads state[0] and writes state[1], Y the
// reverse.

template <int SPT, bool XCD, bool LINK, bool WH>
__global__ __launch_bounds__(CGM_BS) void k_cg1_X(CGMFix a, int odd, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout) {
  __shared__ double lds[2 * CGM_BS / 64];
  (void)odd;
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double2* r = a.r + off;
  double2* qo = a.t + off;
  const SiteMap<SPT, XCD> map(a.nb);
  const CGMState s = load_state(sin + rhs);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  double smv[SPT];
  double2 rv[SPT][7];
  double2 Uo[LINK ? SPT : 1][3], Um[LINK ? SPT : 1][3];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = map.f(g, i);
    // a thread past the last site loads site 0 and stores nothing
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fc);
    if (LINK) load_links(a.U, fc, n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) rv[i][k] = *elem(r, q[k]);
  }
  end_of_loads();
  CGMState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_state1(out, 0, s.iters, s.converged, s.early, s.rho, s.alpha, s.beta, s);
    return;
  }
  double ag = 0.0, ad = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const SiteCode st = kc[i];
    const double2 qq = LINK ? link_hops<-1>(st, rmul(smv[i], rv[i][0]), &rv[i][1], &rv[i][4], Uo[LINK ? i : 0],
                                            Um[LINK ? i : 0])
                            : apply_Mdag_fast_vals(st, smv[i], rv[i][0], &rv[i][1], &rv[i][4]);
    st_nt(elem(qo, fs[i]), qq);
    ag += cnorm2(rv[i][0]);
    ad += cnorm2(qq);
  }
  const double2 t = block_sum2(ag, ad, lds);
  if (threadIdx.x == 0) pair_slots(a, rhs)[partial_slot(blockIdx.x, a.nb)] = t;
  if (lead) put_state1(out, 1, s.iters, 0, 0, s.rho, s.alpha, s.beta, s);
}

template <int SPT, bool XCD, bool LINK, bool WH>
__global__ __launch_bounds__(CGM_BS) void k_cg1_Y(CGMFix a, int odd, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout) {
  __shared__ double lds[2 * CGM_BS / 64];
  (void)odd;
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double2* qv = a.t + off;
  double2* r = a.r + off;
  double2* x = a.x + off;
  double2* p = a.p0 + off;
  double2* sv = a.p1 + off;
  const double2* pp = pair_slots(a, rhs);
  const SiteMap<SPT, XCD> map(a.nb);
  // loads in the order they are needed: state and partials, then the operands
  const CGMState s = load_state(sin + rhs);
  double2 pv2[HEAD_W];
  head_issue2<WH>(pp, a.nb, pv2);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  double smv[SPT];
  double2 tv[SPT][7], rv[SPT], xv[SPT], pv[SPT], so[SPT];
  double2 Uo[LINK ? SPT : 1][3], Um[LINK ? SPT : 1][3];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = map.f(g, i);
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fc);
    if (LINK) load_links(a.U, fc, n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) tv[i][k] = *elem(qv, q[k]);
    rv[i] = *elem(r, fc);
    xv[i] = *elem(x, fc);
    pv[i] = *elem(p, fc);
    so[i] = *elem(sv, fc);
  }
  end_of_loads();
  const double2 gd = head_finish2<WH>(pp, a.nb, pv2, lds);
  const double gam = gd.x, del = gd.y;
  CGMState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_state1(out, 0, s.iters, s.converged, s.early, s.rho, s.alpha, s.beta, s);
    return;
  }
  const int it = s.iters;
  double al, be;
  if (it == 0) {
    // ||b|| < tol returns before the loop only in the imposed start
    // (conjugategradient.c:123-128); the null start tested its prologue
    // residual instead (:22)
    if (s.start == GNX_START_IMPOSED && sqrt(gam) < s.tol) {
      if (lead) {
        put_state1(out, 0, 0, 1, 1, gam, 0.0, 0.0, s);
        CGDevResult res = {1, 1, 0, 0, sqrt(gam), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    al = gam / del;
    be = 0.0;
  } else {
    const bool conv = sqrt(gam) < s.tol;
    if (conv || it >= s.max_iter) {
      if (lead) {
        put_state1(out, 0, it, conv ? 1 : 0, 0, gam, s.alpha, s.beta, s);
        CGDevResult res = {conv ? 1 : 0, 0, it, 0, sqrt(gam), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    be = gam / s.rho;
    al = gam / (del - be * gam / s.alpha);
  }
  // the first iteration has no previous direction: p = r, s = A r
  const bool first = it == 0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const SiteCode st = kc[i];
    const double2 m0 = rmul(smv[i], tv[i][0]);
    const double2 w = LINK ? link_hops<1>(st, m0, &tv[i][1], &tv[i][4], Uo[LINK ? i : 0], Um[LINK ? i : 0])
                           : apply_M_vals(st, smv[i], tv[i][0], &tv[i][1], &tv[i][4]);
    const double2 sn = first ? w : cadd(w, rmul(be, so[i]));
    const double2 pn = first ? rv[i] : cadd(rv[i], rmul(be, pv[i]));
    st_nt(elem(sv, fs[i]), sn);
    st_nt(elem(p, fs[i]), pn);
    st_nt(elem(x, fs[i]), cadd(xv[i], rmul(al, pn)));
    st_nt(elem(r, fs[i]), cadd(rv[i], rmul(-al, sn)));
  }
  if (lead) put_state1(out, 1, it + 1, 0, 0, gam, al, be, s);
}

// ---- one-reduction form with deferred solution updates ------------------------------
// The same iterates as k_cg1_X / k_cg1_Y, bit for bit, with fewer bytes
// written per iteration.  The direction p and the solution x enter no other
// update: each iteration only needs x += alpha_k p_k, p_k = r_k + beta_k
// p_{k-1} at the site.  So Y records (alpha_k, beta_k) in the state instead,
// keeps every residual r_k in a ring of CGD_R buffers (Y(k) reads slot k
// and writes r_{k+1} to slot k+1), and every CGD_M iterations -- and at the
// exit -- applies the pending updates at the site in iteration order, the
// same operations as k_cg1_Y performs one iteration at a time.  Per
// iteration Y writes s and r (and p and x once per CGD_M iterations) instead
// of s, p, x and r: the kernel boundary's write-back of dirty L2 lines is
// the larger part of an iteration's memory time here (a 64 B/site write
// stream costs as much as a 152 B/site read-and-write one).

// x += alpha_j p_j, p_j = r_j + beta_j p_{j-1} (p = r at iteration 0) for the
// n pending iterations j0 .. j0 + n - 1, at this thread's sites
GNX_DEVONLY void cgd_flush_site(const CGMFix& a, size_t off, int f, int j0, int n, const double* hal,
                                const double* hbe) {
  double2 rj[CGD_M];
#pragma unroll
  for (int j = 0; j < CGD_M; j++)
    if (j < n) rj[j] = *elem(ring(a, (j0 + j) & (CGD_R - 1)) + off, f);
  double2 p = *elem(a.p0 + off, f), x = *elem(a.x + off, f);
#pragma unroll
  for (int j = 0; j < CGD_M; j++) {
    if (j >= n) break;
    p = j0 + j == 0 ? rj[j] : cadd(rj[j], rmul(hbe[j], p));
    x = cadd(x, rmul(hal[j], p));
  }
  st_nt(elem(a.p0 + off, f), p);
  st_nt(elem(a.x + off, f), x);
}

template <int SPT, bool XCD>
GNX_DEVONLY void cgd_flush(const CGMFix& a, size_t off, const SiteMap<SPT, XCD>& map, int j0, int n, const double* hal,
                           const double* hbe) {
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(a.g, i);
    if (f < 0) continue;
    cgd_flush_site(a, off, f, j0, n, hal, hbe);
  }
}

template <int SPT, bool XCD, bool LINK, bool WH>
__global__ __launch_bounds__(CGM_BS) void k_cgd_X(CGMFix a, int rs, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout) {
  __shared__ double lds[2 * CGM_BS / 64];
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double2* r = ring(a, rs) + off;
  double2* qo = a.t + off;
  const SiteMap<SPT, XCD> map(a.nb);
  const CGMState s = load_state(sin + rhs);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  double smv[SPT];
  double2 rv[SPT][7];
  double2 Uo[LINK ? SPT : 1][3], Um[LINK ? SPT : 1][3];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = map.f(g, i);
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fc);
    if (LINK) load_links(a.U, fc, n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) rv[i][k] = *elem(r, q[k]);
  }
  end_of_loads();
  CGMState* out = sout + rhs;
  const CGMState* sp = sin + rhs;
  if (!s.active) {
    if (lead) put_state_d(out, 0, s.iters, s.converged, s.early, s.rho, s.alpha, s.beta, s.pend, sp->hal, sp->hbe, s);
    return;
  }
  double ag = 0.0, ad = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const SiteCode st = kc[i];
    const double2 qq = LINK ? link_hops<-1>(st, rmul(smv[i], rv[i][0]), &rv[i][1], &rv[i][4], Uo[LINK ? i : 0],
                                            Um[LINK ? i : 0])
                            : apply_Mdag_fast_vals(st, smv[i], rv[i][0], &rv[i][1], &rv[i][4]);
    st_nt(elem(qo, fs[i]), qq);
    ag += cnorm2(rv[i][0]);
    ad += cnorm2(qq);
  }
  const double2 t = block_sum2(ag, ad, lds);
  if (threadIdx.x == 0) pair_slots(a, rhs)[partial_slot(blockIdx.x, a.nb)] = t;
  if (lead) put_state_d(out, 1, s.iters, 0, 0, s.rho, s.alpha, s.beta, s.pend, sp->hal, sp->hbe, s);
}

template <int SPT, bool XCD, bool LINK, bool WH>
__global__ __launch_bounds__(CGM_BS) void k_cgd_Y(CGMFix a, int rs, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout) {
  __shared__ double lds[2 * CGM_BS / 64];
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const double2* qv = a.t + off;
  const double2* rk = ring(a, rs) + off;
  double2* rn = ring(a, (rs + 1) & (CGD_R - 1)) + off;
  double2* sv = a.p1 + off;
  const double2* pp = pair_slots(a, rhs);
  const SiteMap<SPT, XCD> map(a.nb);
  // loads in the order they are needed: state and partials, then the operands
  const CGMState s = load_state(sin + rhs);
  double2 pv2[HEAD_W];
  head_issue2<WH>(pp, a.nb, pv2);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  double smv[SPT];
  double2 tv[SPT][7], rv[SPT], so[SPT];
  double2 Uo[LINK ? SPT : 1][3], Um[LINK ? SPT : 1][3];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = map.f(g, i);
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fc);
    if (LINK) load_links(a.U, fc, n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) tv[i][k] = *elem(qv, q[k]);
    rv[i] = *elem(rk, fc);
    so[i] = *elem(sv, fc);
  }
  end_of_loads();
  const double2 gd = head_finish2<WH>(pp, a.nb, pv2, lds);
  const double gam = gd.x, del = gd.y;
  CGMState* out = sout + rhs;
  const CGMState* sp = sin + rhs;
  if (!s.active) {
    if (lead) put_state_d(out, 0, s.iters, s.converged, s.early, s.rho, s.alpha, s.beta, s.pend, sp->hal, sp->hbe, s);
    return;
  }
  const int it = s.iters;
  double al, be;
  if (it == 0) {
    // ||b|| < tol returns before the loop only in the imposed start
    // (conjugategradient.c:123-128); the null start tested its prologue
    // residual instead (:22)
    if (s.start == GNX_START_IMPOSED && sqrt(gam) < s.tol) {
      if (lead) {
        put_state_d(out, 0, 0, 1, 1, gam, 0.0, 0.0, 0, sp->hal, sp->hbe, s);
        CGDevResult res = {1, 1, 0, 0, sqrt(gam), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    al = gam / del;
    be = 0.0;
  } else {
    const bool conv = sqrt(gam) < s.tol;
    if (conv || it >= s.max_iter) {
      // the pending updates, then the exit
      cgd_flush(a, off, map, it - s.pend, s.pend, sp->hal, sp->hbe);
      if (lead) {
        put_state_d(out, 0, it, conv ? 1 : 0, 0, gam, s.alpha, s.beta, 0, sp->hal, sp->hbe, s);
        CGDevResult res = {conv ? 1 : 0, 0, it, 0, sqrt(gam), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    be = gam / s.rho;
    al = gam / (del - be * gam / s.alpha);
  }
  // the first iteration has no previous direction: s = A r
  const bool first = it == 0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const SiteCode st = kc[i];
    const double2 m0 = rmul(smv[i], tv[i][0]);
    const double2 w = LINK ? link_hops<1>(st, m0, &tv[i][1], &tv[i][4], Uo[LINK ? i : 0], Um[LINK ? i : 0])
                           : apply_M_vals(st, smv[i], tv[i][0], &tv[i][1], &tv[i][4]);
    const double2 sn = first ? w : cadd(w, rmul(be, so[i]));
    st_nt(elem(sv, fs[i]), sn);
    st_nt(elem(rn, fs[i]), cadd(rv[i], rmul(-al, sn)));
  }
  // record this iteration's update; apply the pending ones every CGD_M
  double hal[CGD_M], hbe[CGD_M];
#pragma unroll
  for (int j = 0; j < CGD_M; j++) {
    hal[j] = j == s.pend ? al : sp->hal[j];
    hbe[j] = j == s.pend ? be : sp->hbe[j];
  }
  int pend = s.pend + 1;
  if (pend == CGD_M) {
    cgd_flush(a, off, map, it + 1 - CGD_M, CGD_M, hal, hbe);
    pend = 0;
  }
  if (lead) put_state_d(out, 1, it + 1, 0, 0, gam, al, be, pend, hal, hbe, s);
}


// ---- fused form: one kernel per iteration ----------------------------------------
// Y(k) and X(k+1) of the one-reduction form in one kernel, so that an
// iteration crosses one kernel boundary instead of two.  X(k+1) needs
// r_{k+1} at the six neighbours of a site, which other blocks of the same
// kernel compute; so each block computes s_k and r_{k+1} on its tile and on
// the tile's one-site halo (a halo site again, with its owner's operations
// on its owner's operands, so the values are the owner's bit for bit),
// keeps r_{k+1} of tile and halo in LDS, and applies M^dag from there:
//
//   Z(k): gamma_k = |r_k|^2, delta_k = |q_k|^2 from the partials of the
//         previous kernel; exit test, alpha_k and beta_k as in k_cg1_Y;
//         on tile and halo: s_k = M q_k + beta_k s_{k-1},
//                           r_{k+1} = r_k - alpha_k s_k  (into LDS);
//         on the tile: p and x as in k_cg1_Y, or deferred as in k_cgd_Y;
//                      q_{k+1} = M^dag r_{k+1} from LDS;
//                      partials (|r_{k+1}|^2, |q_{k+1}|^2).
//
// A T0 x T1 x T2 tile (x0, x1, x2) has 2 (T1 T2 + T0 T2 + T0 T1) halo
// sites, 448 for 4 x 8 x 16 (0.875 per tile site), each one more M q: seven
// loads, mostly L2 hits on the neighbouring tiles' lines.  q, s and r are
// read at halo sites that their owners rewrite in the same kernel, so each
// is double-buffered by the parity of k (q_k in t / t1, s_k in p1 / s1, r_k
// in ring slot k mod 8 with deferral, k mod 2 without), and so are the
// partials.  The tiles divide the lattice (L % T0 = L % T1 = LT % T2 = 0,
// T2 even: a tile row along x2 starts on an even x2 and holds T2/2 sites of
// each parity, contiguous in f-order).  Blocks take tiles in XCD slabs along
// x0.  The prologue runs X(0) (k_cgd_X) and Z(0) before the replayed chunks.
// the fused form's partial pairs of the parity-j kernel, [2][nrhs][nb]
GNX_DEV double2* fpart(const CGMFix& a, int j, int rhs) {
  return (double2*)a.part + ((size_t)j * a.nrhs + rhs) * a.nb;
}

template <int T0, int T1, int T2, bool LINK, bool WH, bool DEF>
__global__ __launch_bounds__(CGM_BS) void k_cgf_Z(CGMFix a, int k8, const CGMState* __restrict__ sin,
                                                   CGMState* __restrict__ sout) {
  using TL = FTile<T0, T1, T2>;
  constexpr int R = TL::ROUNDS, SPT = TL::SPT;
  __shared__ double2 box[TL::NBOX];
  __shared__ double lds[2 * CGM_BS / 64];
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const int kp = k8 & 1;
  const double2* qk = (kp ? a.t1 : a.t) + off;
  double2* qn = (kp ? a.t : a.t1) + off;
  const double2* sold = (kp ? a.p1 : a.s1) + off;
  double2* snew = (kp ? a.s1 : a.p1) + off;
  const double2* rk = ring(a, DEF ? k8 : kp) + off;
  double2* rn = ring(a, DEF ? (k8 + 1) & (CGD_R - 1) : kp ^ 1) + off;
  double2* pd = a.p0 + off;
  double2* xd = a.x + off;
  const double2* pin = fpart(a, kp, rhs);
  int o0, o1, o2;
  ftile_origin<T0, T1, T2>(g, a.nb, (int)blockIdx.x, &o0, &o1, &o2);
  // loads in the order they are needed: state and partials, then the operands
  const CGMState s = load_state(sin + rhs);
  double2 pv2[HEAD_W];
  head_issue2<WH>(pin, a.nb, pv2);
  end_of_loads();
  int fs[R], bx[R];
  SiteCode kc[R];
  double smv[R];
  double2 tv[R][7], rv[R], so[R];
  double2 pv[DEF ? 1 : SPT], xv[DEF ? 1 : SPT];
  double2 Uo[LINK ? R : 1][3], Um[LINK ? R : 1][3];
#pragma unroll
  for (int i = 0; i < R; i++) {
    // an item past the last one loads the tile's first site and stores nothing
    const int e = i * CGM_BS + (int)threadIdx.x;
    const Site st = ftile_site<T0, T1, T2>(g, o0, o1, o2, e < TL::NEXT ? e : 0, &bx[i]);
    fs[i] = f_of_lex(g, st.lex, st.P);
    kc[i] = site_code(g, st);
    NbrF n = neighbours_f(g, st);
    int q[7] = {fs[i], n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm, fs[i]);
    if (LINK) load_links(a.U, fs[i], n, Uo[LINK ? i : 0], Um[LINK ? i : 0]);
#pragma unroll
    for (int k = 0; k < 7; k++) tv[i][k] = *elem(qk, q[k]);
    rv[i] = *elem(rk, fs[i]);
    so[i] = *elem(sold, fs[i]);
    if (!DEF && i < SPT) {
      pv[DEF ? 0 : i] = *elem(pd, fs[i]);
      xv[DEF ? 0 : i] = *elem(xd, fs[i]);
    }
  }
  end_of_loads();
  const double2 gd = head_finish2<WH>(pin, a.nb, pv2, lds);
  const double gam = gd.x, del = gd.y;
  CGMState* out = sout + rhs;
  const CGMState* sp = sin + rhs;
  auto put = [&](int active, int iters, int conv, int early, double rho, double al, double be, int pend,
                 const double* hal, const double* hbe) {
    if constexpr (DEF)
      put_state_d(out, active, iters, conv, early, rho, al, be, pend, hal, hbe, s);
    else
      put_state1(out, active, iters, conv, early, rho, al, be, s);
  };
  if (!s.active) {
    if (lead) put(0, s.iters, s.converged, s.early, s.rho, s.alpha, s.beta, s.pend, sp->hal, sp->hbe);
    return;
  }
  const int it = s.iters;
  double al, be;
  if (it == 0) {
    // ||b|| < tol returns before the loop only in the imposed start
    // (conjugategradient.c:123-128); the null start tested its prologue
    // residual instead (:22)
    if (s.start == GNX_START_IMPOSED && sqrt(gam) < s.tol) {
      if (lead) {
        put(0, 0, 1, 1, gam, 0.0, 0.0, 0, sp->hal, sp->hbe);
        CGDevResult res = {1, 1, 0, 0, sqrt(gam), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    al = gam / del;
    be = 0.0;
  } else {
    const bool conv = sqrt(gam) < s.tol;
    if (conv || it >= s.max_iter) {
      if (DEF) {
        // the pending updates, then the exit
#pragma unroll
        for (int i = 0; i < SPT; i++) cgd_flush_site(a, off, fs[i], it - s.pend, s.pend, sp->hal, sp->hbe);
      }
      if (lead) {
        put(0, it, conv ? 1 : 0, 0, gam, s.alpha, s.beta, 0, sp->hal, sp->hbe);
        CGDevResult res = {conv ? 1 : 0, 0, it, 0, sqrt(gam), {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    be = gam / s.rho;
    al = gam / (del - be * gam / s.alpha);
  }
  // the first iteration has no previous direction: p = r, s = A r
  const bool first = it == 0;
  double2 rt[SPT];
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int e = i * CGM_BS + (int)threadIdx.x;
    if (i >= SPT && e >= TL::NEXT) continue;
    const SiteCode st = kc[i];
    const double2 m0 = rmul(smv[i], tv[i][0]);
    const double2 w = LINK ? link_hops<1>(st, m0, &tv[i][1], &tv[i][4], Uo[LINK ? i : 0], Um[LINK ? i : 0])
                           : apply_M_vals(st, smv[i], tv[i][0], &tv[i][1], &tv[i][4]);
    const double2 sn = first ? w : cadd(w, rmul(be, so[i]));
    const double2 rr = cadd(rv[i], rmul(-al, sn));
    box[bx[i]] = rr;
    if (i < SPT) {
      rt[i < SPT ? i : 0] = rr;
      st_nt(elem(snew, fs[i]), sn);
      st_nt(elem(rn, fs[i]), rr);
      if (!DEF) {
        const double2 pn = first ? rv[i] : cadd(rv[i], rmul(be, pv[DEF ? 0 : i]));
        st_nt(elem(pd, fs[i]), pn);
        st_nt(elem(xd, fs[i]), cadd(xv[DEF ? 0 : i], rmul(al, pn)));
      }
    }
  }
  __syncthreads();
  double ag = 0.0, ad = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const SiteCode st = kc[i];
    const int b = bx[i];
    const double2 vpp[3] = {box[b + TL::S0], box[b + TL::S1], box[b + 1]};
    const double2 vmm[3] = {box[b - TL::S0], box[b - TL::S1], box[b - 1]};
    const double2 qq = LINK ? link_hops<-1>(st, rmul(smv[i], rt[i]), vpp, vmm, Uo[LINK ? i : 0], Um[LINK ? i : 0])
                            : apply_Mdag_fast_vals(st, smv[i], rt[i], vpp, vmm);
    st_nt(elem(qn, fs[i]), qq);
    ag += cnorm2(rt[i]);
    ad += cnorm2(qq);
  }
  const double2 t = block_sum2(ag, ad, lds);
  if (threadIdx.x == 0) fpart(a, kp ^ 1, rhs)[partial_slot(blockIdx.x, a.nb)] = t;
  if (DEF) {
    // record this iteration's update; apply the pending ones every CGD_M
    double hal[CGD_M], hbe[CGD_M];
#pragma unroll
    for (int j = 0; j < CGD_M; j++) {
      hal[j] = j == s.pend ? al : sp->hal[j];
      hbe[j] = j == s.pend ? be : sp->hbe[j];
    }
    int pend = s.pend + 1;
    if (pend == CGD_M) {
#pragma unroll
      for (int i = 0; i < SPT; i++) cgd_flush_site(a, off, fs[i], it + 1 - CGD_M, CGD_M, hal, hbe);
      pend = 0;
    }
    if (lead) put(1, it + 1, 0, 0, gam, al, be, pend, hal, hbe);
  } else if (lead) {
    put(1, it + 1, 0, 0, gam, al, be, 0, nullptr, nullptr);
  }
}

// ---- shapes the tuner chooses from -----------------------------------------------
struct CGMShape {
  int spt;
  bool xcd;
  bool link;
  bool one;        // one-reduction form: A = k_cg1_X, B = k_cg1_Y (or k_cgd_*)
  bool deferred;   // the one-reduction form with deferred x and p updates (k_cgd_X, k_cgd_Y)
  const void* A;   // block-wide head reduction
  const void* B;
  const void* Aw;  // per-wave head reduction (up to 512 partials)
  const void* Bw;
  const void* start_imposed;
  const void* null_p1;
  const void* null_p2;
  const void* null_p3;
  bool fused;      // one kernel per iteration (k_cgf_Z; A is the prologue X(0), k_cgd_X)
  int T0, T1, T2;  // its tile
};

#define GNX_SHAPE(S, X, K)                                                                                  \
  {S, X, K, false, false, (const void*)k_cgm_A<S, X, K, false>, (const void*)k_cgm_B<S, X, K, false>,               \
   (const void*)k_cgm_A<S, X, K, true>, (const void*)k_cgm_B<S, X, K, true>, (const void*)k_cgm_start_imposed<S, X>, \
   (const void*)k_cgm_null_p1<S, X, K>, (const void*)k_cgm_null_p2<S, X, K>, (const void*)k_cgm_null_p3<S, X>}
#define GNX_SHAPE1(S, X, K)                                                                                 \
  {S, X, K, true, false, (const void*)k_cg1_X<S, X, K, false>, (const void*)k_cg1_Y<S, X, K, false>,                \
   (const void*)k_cg1_X<S, X, K, true>, (const void*)k_cg1_Y<S, X, K, true>, (const void*)k_cgm_start_imposed<S, X>, \
   (const void*)k_cgm_null_p1<S, X, K>, (const void*)k_cgm_null_p2<S, X, K>, (const void*)k_cgm_null_p3<S, X>}
#define GNX_SHAPED(S, X, K)                                                                                 \
  {S, X, K, true, true, (const void*)k_cgd_X<S, X, K, false>, (const void*)k_cgd_Y<S, X, K, false>,          \
   (const void*)k_cgd_X<S, X, K, true>, (const void*)k_cgd_Y<S, X, K, true>, (const void*)k_cgm_start_imposed<S, X>, \
   (const void*)k_cgm_null_p1<S, X, K>, (const void*)k_cgm_null_p2<S, X, K>, (const void*)k_cgm_null_p3<S, X>}
#define GNX_SHAPEF(T0, T1, T2, K, D)                                                                      \
  {(T0) * (T1) * (T2) / CGM_BS,                                                                             \
   true,                                                                                                    \
   K,                                                                                                       \
   true,                                                                                                    \
   D,                                                                                                       \
   (const void*)k_cgd_X<(T0) * (T1) * (T2) / CGM_BS, true, K, false>,                                       \
   (const void*)k_cgf_Z<T0, T1, T2, K, false, D>,                                                           \
   (const void*)k_cgd_X<(T0) * (T1) * (T2) / CGM_BS, true, K, true>,                                        \
   (const void*)k_cgf_Z<T0, T1, T2, K, true, D>,                                                            \
   (const void*)k_cgm_start_imposed<(T0) * (T1) * (T2) / CGM_BS, true>,                                     \
   (const void*)k_cgm_null_p1<(T0) * (T1) * (T2) / CGM_BS, true, K>,                                        \
   (const void*)k_cgm_null_p2<(T0) * (T1) * (T2) / CGM_BS, true, K>,                                        \
   (const void*)k_cgm_null_p3<(T0) * (T1) * (T2) / CGM_BS, true>,                                           \
   true,                                                                                                    \
   T0,                                                                                                      \
   T1,                                                                                                      \
   T2}
// per operator (the live operator first, then the link-coupled one): 0..5
// the two-reduction kernels, 6..11 the one-reduction kernels, 12..17 the
// one-reduction kernels with deferred x and p updates, each with 1, 1, 2, 2,
// 4, 4 sites per thread, odd indices with the XCD slab mapping; 18..21 the
// fused kernels (one per iteration) on 4 x 8 x 16 and 8 x 8 x 8 tiles,
// without and with deferral
static const CGMShape kShapes[] = {
    GNX_SHAPE(1, false, false),  GNX_SHAPE(1, true, false),  GNX_SHAPE(2, false, false),
    GNX_SHAPE(2, true, false),   GNX_SHAPE(4, false, false), GNX_SHAPE(4, true, false),
    GNX_SHAPE1(1, false, false), GNX_SHAPE1(1, true, false), GNX_SHAPE1(2, false, false),
    GNX_SHAPE1(2, true, false),  GNX_SHAPE1(4, false, false), GNX_SHAPE1(4, true, false),
    GNX_SHAPED(1, false, false), GNX_SHAPED(1, true, false), GNX_SHAPED(2, false, false),
    GNX_SHAPED(2, true, false),  GNX_SHAPED(4, false, false), GNX_SHAPED(4, true, false),
    GNX_SHAPEF(4, 8, 16, false, false), GNX_SHAPEF(4, 8, 16, false, true),
    GNX_SHAPEF(8, 8, 8, false, false),  GNX_SHAPEF(8, 8, 8, false, true),
    GNX_SHAPE(1, false, true),   GNX_SHAPE(1, true, true),   GNX_SHAPE(2, false, true),
    GNX_SHAPE(2, true, true),    GNX_SHAPE(4, false, true),  GNX_SHAPE(4, true, true),
    GNX_SHAPE1(1, false, true),  GNX_SHAPE1(1, true, true),  GNX_SHAPE1(2, false, true),
    GNX_SHAPE1(2, true, true),   GNX_SHAPE1(4, false, true), GNX_SHAPE1(4, true, true),
    GNX_SHAPED(1, false, true),  GNX_SHAPED(1, true, true),  GNX_SHAPED(2, false, true),
    GNX_SHAPED(2, true, true),   GNX_SHAPED(4, false, true), GNX_SHAPED(4, true, true),
    GNX_SHAPEF(4, 8, 16, true, false), GNX_SHAPEF(4, 8, 16, true, true),
    GNX_SHAPEF(8, 8, 8, true, false),  GNX_SHAPEF(8, 8, 8, true, true)};
#undef GNX_SHAPE
#undef GNX_SHAPE1
#undef GNX_SHAPED
#undef GNX_SHAPEF
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
constexpr int kShapesPerOp = kNumShapes / 2;

static int shape_blocks(const Geom& g, int shape) {
  int per = CGM_BS * kShapes[shape].spt;
  return (g.V + per - 1) / per;
}

// whether a shape runs on lattice g: a fused shape's tiles must divide it
static bool shape_fits(const Geom& g, int shape) {
  const CGMShape& sh = kShapes[shape];
  return !sh.fused || (g.L % sh.T0 == 0 && g.L % sh.T1 == 0 && g.LT % sh.T2 == 0);
}

// ---- host side ---------------------------------------------------------------
struct CGMachine {
  GraphEngine eng;
  double* sm = nullptr;
  double2* U = nullptr;  // [3V] link copy (link-coupled solves)
  double2* x = nullptr;
  double2* r = nullptr;  // CGD_R residual buffers of cap elements (the deferred form's ring), r the first
  double2* p0 = nullptr;
  double2* p1 = nullptr;
  double2* t = nullptr;
  double2* t1 = nullptr;  // the fused form's second q and s buffers
  double2* s1 = nullptr;
  CGDevResult* res = nullptr;
  long cap = 0;          // complex elements per vector
  int forced_shape = -1; // GNX_CG_SHAPE (index within an operator's shapes)
  int autotune = 0;      // 1: choose shapes by timing (GNX_AUTOTUNE=1, gnx_set_autotune)
  void free_vectors() {
    void* bufs[] = {sm, U, x, r, p0, p1, t, t1, s1, res};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    sm = nullptr;
    U = nullptr;
    x = r = p0 = p1 = t = t1 = s1 = nullptr;
    res = nullptr;
    cap = 0;
  }
  ~CGMachine() { free_vectors(); }
};

void cgm_release(void* m) { delete (CGMachine*)m; }

static CGMachine* cgm_get(void** slotp) {
  if (!*slotp) {
    CGMachine* m = new CGMachine();
    const char* env = getenv("GNX_CG_SHAPE");
    if (env) {
      int v = atoi(env);
      if (v >= 0 && v < kShapesPerOp) m->forced_shape = v;
    }
    const char* at = getenv("GNX_AUTOTUNE");
    if (at && atoi(at) != 0) m->autotune = 1;
    *slotp = m;
  }
  return (CGMachine*)*slotp;
}

static hipError_t cgm_ensure(CGMachine* m, const Geom& g, int nrhs) {
  const int nb_max = (g.V + CGM_BS - 1) / CGM_BS;
  const int cap_rhs = nrhs < 2 ? 2 : nrhs;
  // partials: two slots of nrhs x nb pairs (the fused form's double buffer)
  hipError_t e = m->eng.ensure(cap_rhs, 1, sizeof(CGMState), 4 * (size_t)cap_rhs * nb_max);
  if (e != hipSuccess) return e;
  const long need = (long)g.V * cap_rhs;
  if (m->cap >= need && m->res) return hipSuccess;
  m->free_vectors();
  m->eng.invalidate_graphs();  // captured arguments point at the old vectors
  e = hipMalloc(&m->sm, sizeof(double) * g.V);
  if (e == hipSuccess) e = hipMalloc(&m->U, sizeof(double2) * 3 * (size_t)g.V);
  if (e == hipSuccess) e = hipMalloc(&m->x, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->r, sizeof(double2) * need * CGD_R);
  if (e == hipSuccess) e = hipMalloc(&m->p0, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->p1, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->t, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->t1, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->s1, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->res, sizeof(CGDevResult) * cap_rhs);
  if (e != hipSuccess) {
    m->free_vectors();
    return e;
  }
  m->cap = need;
  return hipSuccess;
}

static CGMFix cgm_fix(CGMachine* m, const Geom& g, int nrhs, int shape) {
  CGMFix a;
  a.g = g;
  a.sm = m->sm;
  a.U = m->U;
  a.x = m->x;
  a.r = m->r;
  a.p0 = m->p0;
  a.p1 = m->p1;
  a.t = m->t;
  a.part = m->eng.dev_part();
  a.state = (CGMState*)m->eng.dev_state();
  a.res = m->res;
  a.rring = m->r;
  a.rstride = (size_t)m->cap;
  a.t1 = m->t1;
  a.s1 = m->s1;
  a.nrhs = nrhs;
  a.nb = shape_blocks(g, shape);
  return a;
}

// `iters` iterations (every A and B kernel) on stream cs, ENG_CHUNK or
// ENG_TAIL: captured into graphs by cgm_graph, launched directly by the
// static solve
static void cgm_chunk(const CGMFix& a, int nrhs, int shape, hipStream_t cs, int iters) {
  dim3 grid(a.nb, nrhs), blk(CGM_BS);
  const CGMState* s0 = a.state;
  const CGMState* s1 = a.state + nrhs;
  CGMState* w0 = a.state;
  CGMState* w1 = a.state + nrhs;
  const double* q0 = slot(a, 0, 0);
  const double* q1 = slot(a, 1, 0);
  double* o0 = a.part;
  double* o1 = a.part + (size_t)nrhs * a.nb;
  const CGMShape& sh = kShapes[shape];
  const bool wh = head_wave(a.nb);
  CGMFix ac = a;
  static_assert(ENG_CHUNK % CGD_R == 0 && ENG_TAIL % CGD_R == 0, "the residual ring turns over within a replay");
  if (sh.fused) {
    // Z(k) for k = 1 + i (the prologue ran X(0) and Z(0)): node i reads
    // state[i & 1] and writes the other half, so the chunk ends in state[0]
    for (int i = 0; i < iters; i++) {
      int k8 = (i + 1) & (CGD_R - 1);
      const CGMState* sin = (i & 1) ? s1 : s0;
      CGMState* sout = (i & 1) ? w0 : w1;
      void* args[] = {(void*)&ac, (void*)&k8, (void*)&sin, (void*)&sout};
      (void)hipLaunchKernel(wh ? sh.Bw : sh.B, grid, blk, args, 0, cs);
    }
    return;
  }
  for (int i = 0; i < iters; i++) {
    // iteration parity (p buffers of the two-reduction form), or the
    // residual ring slot of the deferred form
    int odd = sh.deferred ? (i & (CGD_R - 1)) : (i & 1);
    // A: state[0] -> state[1], reduces slot 1, writes slot 0; B the reverse.
    // The one-reduction kernels take (a, odd, state in, state out) only.
    void* argsA[] = {(void*)&ac, (void*)&odd, (void*)&s0, (void*)&w1, (void*)&q1, (void*)&o0};
    void* argsB[] = {(void*)&ac, (void*)&odd, (void*)&s1, (void*)&w0, (void*)&q0, (void*)&o1};
    (void)hipLaunchKernel(wh ? sh.Aw : sh.A, grid, blk, argsA, 0, cs);
    (void)hipLaunchKernel(wh ? sh.Bw : sh.B, grid, blk, argsB, 0, cs);
  }
}

static hipError_t cgm_graph(CGMachine* m, const Geom& g, int nrhs, int shape, EngGraph* exec) {
  const CGMFix a = cgm_fix(m, g, nrhs, shape);
  auto body = [&](hipStream_t cs, int iters) { cgm_chunk(a, nrhs, shape, cs, iters); };
  return m->eng.graph(shape, nrhs, body, exec);
}

// the start kernels of one solve into the buffers of a
static hipError_t cgm_start(const CGMFix& a, const double* sm, const double2* b, const double2* xin, double tol,
                            long max_iter, int start, int nrhs, int shape, hipStream_t st) {
  CGMFix ac = a;
  dim3 grid(a.nb, nrhs), blk(CGM_BS);
  hipError_t e;
  if (start == GNX_START_NULL) {
    void* a1[] = {(void*)&ac, (void*)&sm, (void*)&xin};
    e = hipLaunchKernel(kShapes[shape].null_p1, grid, blk, a1, 0, st);
    void* a2[] = {(void*)&ac, (void*)&b};
    if (e == hipSuccess) e = hipLaunchKernel(kShapes[shape].null_p2, grid, blk, a2, 0, st);
    void* a3[] = {(void*)&ac, (void*)&b, (void*)&tol, (void*)&max_iter};
    if (e == hipSuccess) e = hipLaunchKernel(kShapes[shape].null_p3, grid, blk, a3, 0, st);
  } else {
    void* a0[] = {(void*)&ac, (void*)&sm, (void*)&b, (void*)&tol, (void*)&max_iter};
    e = hipLaunchKernel(kShapes[shape].start_imposed, grid, blk, a0, 0, st);
  }
  const CGMShape& sh = kShapes[shape];
  if (e == hipSuccess && sh.fused) {
    // the fused form's prologue: X(0) (state[0] -> state[1], q_0 into t,
    // partials into slot 0), then Z(0) (state[1] -> state[0])
    const bool wh = head_wave(a.nb);
    int k0 = 0;
    const CGMState* s0 = a.state;
    const CGMState* s1 = a.state + nrhs;
    CGMState* w0 = a.state;
    CGMState* w1 = a.state + nrhs;
    void* ax[] = {(void*)&ac, (void*)&k0, (void*)&s0, (void*)&w1};
    e = hipLaunchKernel(wh ? sh.Aw : sh.A, grid, blk, ax, 0, st);
    void* az[] = {(void*)&ac, (void*)&k0, (void*)&s1, (void*)&w0};
    if (e == hipSuccess) e = hipLaunchKernel(wh ? sh.Bw : sh.B, grid, blk, az, 0, st);
  }
  return e;
}

// run the start kernels of one solve
static hipError_t cgm_begin(CGMachine* m, const Geom& g, const double* sm, const double2* U, const double2* b,
                            const double2* xin, double tol, long max_iter, int start, int nrhs, int shape,
                            hipStream_t st) {
  if (kShapes[shape].link) {
    hipError_t e = hipMemcpyAsync(m->U, U, sizeof(double2) * 3 * (size_t)g.V, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return cgm_start(cgm_fix(m, g, nrhs, shape), sm, b, xin, tol, max_iter, start, nrhs, shape, st);
}

// Launch shape of a solve.  By default a fixed rule, so that a solve's
// iterates (which depend on the form and on the partial-sum partition) are
// the same in every process: on the live operator, the fused form (one
// kernel per iteration, 4 x 8 x 16 tiles) where the batch's tiles occupy
// 160 to 256 CUs one each; otherwise the one-reduction kernels (with
// deferred x and p updates once the batch holds 400 000 sites) in the XCD
// slab mapping with the most sites per thread (4, 2, 1) that still gives the
// batch 512 blocks of at least 8 per system (the sites per thread the timing
// tuner picks at 64^3 single and batched).  GNX_CG_SHAPE pins a shape; autotune
// (GNX_AUTOTUNE=1 or gnx_set_autotune) times every shape on 2*CHUNK
// iterations of the caller's own system (imposed start into the solver's
// buffers; the caller's x is not touched) and keeps the fastest per key.
static int cgm_rule(const Geom& g, int nrhs, bool link = false) {
  // the fused form saves a kernel boundary per iteration but recomputes each
  // tile's halo: with one tile per CU the halo hides behind the boundary
  // (48^3 single 7.1 against 8.7 us per iteration; 32^3 x 3 and x 4 7.7 and
  // 7.9 against 8.5 and 9.1), with fewer tiles it ties (32^3 x 1 and x 2),
  // with two or more per CU the halo work loses (64^3 single 14.4 against
  // 11.0, 48^3 x 2 12.4 against 11.3)
  constexpr int kFused = 18;
  if (!link && shape_fits(g, kFused)) {
    const long tiles = (long)shape_blocks(g, kFused) * nrhs;
    if (tiles >= 160 && tiles <= 256) return kFused;
  }
  // the deferred form pays where the iteration is bound by its bytes (64^3
  // x 2 and larger: -12 % per iteration at x 2, -16 % at x 16, -4 % at 80^3
  // single), not where it is bound by latency (+9 % at 64^3 single, +2 % at
  // 48^3)
  const int base = (long)g.V * nrhs >= 400000 ? 13 : 7;
  for (int k = 2; k >= 1; k--) {
    const long nb = (g.V + (long)CGM_BS * (1 << k) - 1) / ((long)CGM_BS * (1 << k));
    if (nb >= NUM_XCD && nb * nrhs >= 512) return base + 2 * k;
  }
  return base;
}

static int cgm_key(int nrhs, bool link) { return nrhs + (link ? 1000 : 0); }

static hipError_t cgm_choose(CGMachine* m, const Geom& g, const double* sm, const double2* U, const double2* b,
                             int nrhs, hipStream_t st, int* best) {
  GraphEngine& E = m->eng;
  const bool link = U != nullptr;
  const int key = cgm_key(nrhs, link);
  const int first = link ? kShapesPerOp : 0;
  const int cached = E.shapes.find(key);
  if (cached >= 0) {
    *best = cached;
    return hipSuccess;
  }
  int best_shape = first + cgm_rule(g, nrhs, link);
  hipError_t e = hipSuccess;
  if (m->forced_shape >= 0 && shape_fits(g, first + m->forced_shape)) {
    best_shape = first + m->forced_shape;
  } else if (m->autotune) {
    E.tune_runs++;
    float best_ms = 1e30f;
    for (int sh = first; sh < first + kShapesPerOp && e == hipSuccess; sh++) {
      if (!shape_fits(g, sh)) continue;
      EngGraph exec;
      e = cgm_graph(m, g, nrhs, sh, &exec);
      float ms_min = 1e30f;
      for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
        e = cgm_begin(m, g, sm, U, b, nullptr, 0.0, 2 * ENG_CHUNK, GNX_START_IMPOSED, nrhs, sh, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        float ms = 0.f;
        if (e == hipSuccess) e = E.time_replays(exec, 2, st, &ms);
        if (rep > 0 && ms < ms_min) ms_min = ms;
      }
      if (e == hipSuccess && ms_min < best_ms) {
        best_ms = ms_min;
        best_shape = sh;
      }
    }
    if (e != hipSuccess) return e;
  }
  E.shapes.put(key, best_shape);
  *best = best_shape;
  return hipSuccess;
}

hipError_t cgm_f64_run(void** machine_slot, const Geom& g, const double* sm, const double2* U, const double2* b,
                       double2* x, double tol, long max_iter, int start, int nrhs, hipStream_t st,
                       CGDevResult* host_res) {
  CGMachine* m = cgm_get(machine_slot);
  hipError_t e = cgm_ensure(m, g, nrhs);
  if (e != hipSuccess) return e;
  int shape = 0;
  e = cgm_choose(m, g, sm, U, b, nrhs, st, &shape);
  if (e != hipSuccess) return e;
  EngGraph exec;
  e = cgm_graph(m, g, nrhs, shape, &exec);
  if (e != hipSuccess) return e;
  e = cgm_begin(m, g, sm, U, b, x, tol, max_iter, start, nrhs, shape, st);
  if (e != hipSuccess) return e;
  const FinishFn out = [&](hipStream_t s) {
    hipError_t f = hipMemcpyAsync(x, m->x, sizeof(double2) * (size_t)g.V * nrhs, hipMemcpyDeviceToDevice, s);
    if (f == hipSuccess && host_res)
      f = hipMemcpyAsync(host_res, m->res, sizeof(CGDevResult) * nrhs, hipMemcpyDeviceToHost, s);
    return f;
  };
  return m->eng.run(exec, nrhs, max_iter, st, nullptr, eng_key(shape, nrhs, tol, start), nullptr, &out);
}

// ---- solve without allocation ----------------------------------------------------
// gnx_cg_noalloc runs the kernels of gnx_cg in the same launch shape (the
// fixed rule) with every buffer of the solver -- sigma mass copy, x, r, the
// two direction buffers, q, the partial sums, the state and the result -- in
// one static device array, so the call itself never allocates.  Each device
// that loads this library holds its own copy of the array until the process
// ends; the host uses the address hipGetSymbolAddress reports for the
// current device, never the array's name (a host-side stand-in).  Every call
// on a device shares that device's array, and it keeps the last call's
// contents: the start kernels write every word a solve reads before it reads
// it, so nothing needs clearing, and calls that overlap on one device race.
constexpr size_t kStaticScratchBytes = (size_t)80 << 20;
constexpr int kStaticMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static unsigned char* g_scratch_dev[kStaticMaxDevices];
static std::mutex g_scratch_mu;

static size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

// The shape of the static solve: the rule's for one system, the deferred
// form replaced by its equivalent without the residual ring (the same
// iterates); the fused form (not deferred) keeps two residual slots.
static int static_shape(const Geom& g) {
  const int rule = cgm_rule(g, 1);
  if (kShapes[rule].fused) return rule;
  return rule >= 12 ? rule - 6 : rule;
}

// byte offsets of the solver's buffers in the array for one system of g
// (the fused form's second residual, q and s buffers only for it)
struct StaticLayout {
  size_t sm, x, r, p0, p1, t, t1, s1, part, state, res, total;
  size_t vec;
};
static StaticLayout static_layout(const Geom& g, int nb, bool fused) {
  StaticLayout l;
  const size_t vec = up256(sizeof(double2) * (size_t)g.V);
  l.vec = vec;
  l.sm = 0;
  l.x = l.sm + up256(sizeof(double) * (size_t)g.V);
  l.r = l.x + vec;  // and, fused, the second residual slot right after it
  l.p0 = l.r + (fused ? 2 : 1) * vec;
  l.p1 = l.p0 + vec;
  l.t = l.p1 + vec;
  l.t1 = l.t + vec;
  l.s1 = l.t1 + (fused ? vec : 0);
  l.part = l.s1 + (fused ? vec : 0);
  l.state = l.part + up256(sizeof(double) * (fused ? 4 : 2) * (size_t)nb);
  l.res = l.state + up256(sizeof(CGMState) * 2);
  l.total = l.res + up256(sizeof(CGDevResult));
  return l;
}

size_t cgm_static_bytes(const Geom& g) {
  const int sh = static_shape(g);
  return static_layout(g, shape_blocks(g, sh), kShapes[sh].fused).total;
}

size_t cgm_static_capacity() { return kStaticScratchBytes; }

// device address of this device's array (looked up once per device, with
// that device current)
static hipError_t static_scratch(int device, unsigned char** out) {
  if (device < 0 || device >= kStaticMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = (unsigned char*)ptr;
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

hipError_t cgm_f64_run_static(const Geom& g, int device, const double* sm, const double2* b, double2* x, double tol,
                              long max_iter, int start, hipStream_t st, void* pinned, CGDevResult* host_res) {
  unsigned char* base = nullptr;
  hipError_t e = static_scratch(device, &base);
  if (e != hipSuccess) return e;
  const int shape = static_shape(g);
  const bool fused = kShapes[shape].fused;
  CGMFix a;
  a.g = g;
  a.nrhs = 1;
  a.nb = shape_blocks(g, shape);
  const StaticLayout l = static_layout(g, a.nb, fused);
  if (l.total > kStaticScratchBytes) return hipErrorInvalidValue;  // the caller checks the size first
  a.sm = (double*)(base + l.sm);
  a.U = nullptr;
  a.x = (double2*)(base + l.x);
  a.r = (double2*)(base + l.r);
  a.p0 = (double2*)(base + l.p0);
  a.p1 = (double2*)(base + l.p1);
  a.t = (double2*)(base + l.t);
  a.part = (double*)(base + l.part);
  a.state = (CGMState*)(base + l.state);
  a.res = (CGDevResult*)(base + l.res);
  a.rring = a.r;
  a.rstride = fused ? l.vec / sizeof(double2) : 0;
  a.t1 = fused ? (double2*)(base + l.t1) : nullptr;
  a.s1 = fused ? (double2*)(base + l.s1) : nullptr;
  e = cgm_start(a, sm, b, x, tol, max_iter, start, 1, shape, st);
  if (e != hipSuccess) return e;
  // chunks of ENG_CHUNK iterations launched directly, the state's `active`
  // and `iters` polled through the caller's pinned buffer after each
  int* flags = (int*)pinned;
  for (long done = 0;; done += ENG_CHUNK) {
    cgm_chunk(a, 1, shape, st, ENG_CHUNK);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(flags, a.state, 2 * sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (!flags[0]) break;
    if (done > max_iter + 2 * ENG_CHUNK) return hipErrorUnknown;  // the state machine failed to stop
  }
  e = hipMemcpyAsync(x, a.x, sizeof(double2) * (size_t)g.V, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(pinned, a.res, sizeof(CGDevResult), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) memcpy(host_res, pinned, sizeof(CGDevResult));
  return e;
}

int cgm_shape_of(void* machine, int nrhs, int* spt, int* xcd) {
  CGMachine* m = (CGMachine*)machine;
  if (!m) return -1;
  const int sh = m->eng.shapes.peek(cgm_key(nrhs, false));
  if (sh < 0) return -1;
  if (spt) *spt = kShapes[sh].spt;
  if (xcd)
    *xcd = (kShapes[sh].xcd ? 1 : 0) | (kShapes[sh].one ? 2 : 0) | (kShapes[sh].deferred ? 4 : 0) |
           (kShapes[sh].fused ? 8 : 0);
  return sh;
}

void cgm_set_autotune(void** machine_slot, int on) { cgm_get(machine_slot)->autotune = on; }

long cgm_tune_runs(void* machine) { return machine ? ((CGMachine*)machine)->eng.tune_runs : 0; }
void cgm_replay_stats(void* machine, long* slots, long* polls) {
  *slots = machine ? ((CGMachine*)machine)->eng.launched_slots : 0;
  *polls = machine ? ((CGMachine*)machine)->eng.polls : 0;
}

}  // namespace gnx
