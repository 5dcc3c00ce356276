// gnx_cgmix.hip -- mixed-precision CG on A = M M^dag with FP64 reliable
// updates (Sleijpen & van der Vorst's reliable-update scheme as used for
// lattice Dirac solvers).
//
// Inner iterations keep r, p, t in reduced-precision storage (FP32, or BF16
// or FP16 with FP32 arithmetic) and the solution correction xs in FP32; every global
// sum is accumulated in FP64.  When the iterated residual has dropped by a
// factor `delta` since the last update, claims convergence or reaches the
// iteration cap, a reliable update folds xs into the FP64 solution and
// recomputes the true residual r = b - M M^dag x in FP64; convergence is
// declared only on that true residual (||r|| < tol, absolute, as
// conjugategradient.c:87-89).  The search direction is kept across updates
// (no restart), so the Krylov space is not lost.  FP64 is used only for x,
// the true residual, the sums and the scalars.
//
// Kernel slots, per system (same graph engine as the FP64 solver):
//   X: reduce rho; after an update, test the true residual; otherwise test
//      ||r_it|| < delta * rmax -> start an update (u = M^dag (x + xs) in
//      FP64, p copied forward so the slot parity keeps its meaning); else
//      p = r + beta p, t = M^dag p.
//   Y: finish an update (x += xs, xs = 0, r = (S)(b - M u), partial |r|^2 in
//      FP64) or the normal step (xs += alpha p, r -= alpha M t, partial |r|^2).
// The loads-first kernels (k_mix_Xe / k_mix_Ye) move the xs update out of Y:
// it is folded into the X kernel of every even slot and into the update.
// As in the FP64 solver, each kernel issues its loads (state and partials,
// then the operands) before its first wait, and kernel arguments are passed
// by value.
#include <stdlib.h>
#include <string.h>

#include "gnx_cgstate.h"
#include "gnx_engine.h"
#include "gnx_internal.h"
#include "gnx_group.h"
#include "gnx_solver.h"

namespace gnx {

// ---- storage formats --------------------------------------------------------
// Store policy of the inner iterations, per solve (nt, a kernel argument):
// while a batch's working set (r, t, two p, xs) fits the 256 MB Infinity
// Cache, ordinary stores keep it there for the next kernel (FP16 x32 at 64^3,
// 201 MB: 824-832 against 738 RHS-solves/s with non-temporal stores; BF16
// 714 against 652; FP32 x16 497 against 444); past it (FP32 x32 at 64^3, 335
// MB; 96^3 batches) non-temporal stores win by 2-5 %, as for the FP64 kernels.
// The branch is uniform and follows every load of a kernel.
GNX_DEVONLY void st_mix(float2* p, float2 v, int nt) {
  if (nt) {
    gnx_f2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, (gnx_f2v*)p);
  } else {
    *p = v;
  }
}
GNX_DEVONLY void st_mix(unsigned int* p, unsigned int v, int nt) {
  if (nt) __builtin_nontemporal_store(v, p);
  else *p = v;
}

struct StoreF32 {
  using T = float2;
  static constexpr bool kScaled = false;
  GNX_DEVONLY static float2 ld(const T* p, int i) { return *elem(p, i); }
  GNX_DEVONLY static void st(T* p, int i, float2 v, int nt) { st_mix(elem(p, i), v, nt); }
};

// bfloat16 pair (re in the low half), round-to-nearest-even on store: one
// v_cvt_pk_bf16_f32 (gfx950) for the pair
typedef __bf16 gnx_bf16x2 __attribute__((ext_vector_type(2)));
typedef float gnx_f32x2 __attribute__((ext_vector_type(2)));
struct StoreBF16 {
  using T = unsigned int;
  static constexpr bool kScaled = false;
  GNX_DEVONLY static float up(unsigned int h) { return __uint_as_float(h << 16); }
  GNX_DEVONLY static float2 cvt(T v) { return make_float2(up(v & 0xFFFFu), up(v >> 16)); }
  GNX_DEVONLY static T pack(float2 v) {
    const gnx_f32x2 w = {v.x, v.y};
    return __builtin_bit_cast(unsigned int, __builtin_convertvector(w, gnx_bf16x2));
  }
  GNX_DEVONLY static float2 ld(const T* p, int i) { return cvt(*elem(p, i)); }
  GNX_DEVONLY static void st(T* p, int i, float2 v, int nt) { st_mix(elem(p, i), pack(v), nt); }
};

// IEEE half pair (re in the low half), round-to-nearest-even on store (one
// v_cvt_pk_f16_f32).  Three more mantissa bits than BF16 in the same four
// bytes, but a range of 6e-5 .. 65504, so r, p, t are stored times a
// per-system scale s = sqrt(V) / ||r|| set at every reliable update (the
// first slot of a solve is one): elements of r are then O(1).  Only the
// loads-first kernels implement the scale.
typedef _Float16 gnx_f16x2 __attribute__((ext_vector_type(2)));
struct StoreF16 {
  using T = unsigned int;
  static constexpr bool kScaled = true;
  GNX_DEVONLY static float2 cvt(T v) {
    const gnx_f32x2 w = __builtin_convertvector(__builtin_bit_cast(gnx_f16x2, v), gnx_f32x2);
    return make_float2(w.x, w.y);
  }
  GNX_DEVONLY static T pack(float2 v) {
    const gnx_f32x2 w = {v.x, v.y};
    return __builtin_bit_cast(unsigned int, __builtin_convertvector(w, gnx_f16x2));
  }
  GNX_DEVONLY static float2 ld(const T* p, int i) { return cvt(*elem(p, i)); }
  GNX_DEVONLY static void st(T* p, int i, float2 v, int nt) { st_mix(elem(p, i), pack(v), nt); }
};

constexpr int MIX_BS = ENG_BS;

enum MixPhase { PH_NORMAL = 0, PH_UPDATE = 1, PH_UPDATED = 2 };

struct MixState {
  int active;
  int iters;      // CG iterations (normal slot pairs)
  int phase;
  int nrel;       // reliable updates done
  int converged;
  int early;
  int start;
  int pad;
  double rho;     // rho of the current direction (alpha = rho / gamma)
  double rmax;    // ||r|| at the last reliable update
  double tol;
  double delta;
  long max_iter;
  // loads-first form: step lengths whose alpha p terms are not yet in xs
  // (alpha_a for the newest direction, alpha_b for the one before)
  double alpha_a;
  double alpha_b;
  // scaled storage (FP16): r, p, t hold s times their values; an update's
  // X kernel sets scale_next for its Y kernel, which rescales to it
  double scale;
  double scale_next;
};

GNX_DEVONLY void put_mstate(MixState* o, int active, int iters, int phase, int nrel, int conv, int early,
                            double rho, double rmax, const MixState& cfg) {
  o->active = active;
  o->iters = iters;
  o->phase = phase;
  o->nrel = nrel;
  o->converged = conv;
  o->early = early;
  o->start = cfg.start;
  o->pad = cfg.pad;  // every loaded word stays live: a dead load destination is waited on early
  o->rho = rho;
  o->rmax = rmax;
  o->tol = cfg.tol;
  o->delta = cfg.delta;
  o->max_iter = cfg.max_iter;
  o->alpha_a = cfg.alpha_a;
  o->alpha_b = cfg.alpha_b;
  o->scale = cfg.scale;
  o->scale_next = cfg.scale_next;
}
GNX_DEVONLY void put_alphas(MixState* o, double alpha_a, double alpha_b) {
  o->alpha_a = alpha_a;
  o->alpha_b = alpha_b;
}

// kernel arguments, fixed per solver instance (buffers the solver owns)
struct MixFix {
  Geom g;
  double* sm;       // FP64 copy of the caller's sigma mass
  float* sm32;
  double2* b;       // copy of the right-hand sides
  double2* x;       // FP64 solutions
  double2* u;       // FP64 M^dag (x + xs) of a reliable update
  float2* xs;       // FP32 solution correction
  void* r;          // storage-format residual
  void* p0;  // direction buffers, slot parity: X reads p[sl], writes p[sl+1]
  void* p1;
  void* t;
  double* part;     // [2 slots][ptotal]
  MixState* state;  // [2][nrhs]
  CGDevResult* res;
  int nrhs;
  int nb;
  int nt;           // non-temporal stores: the working set exceeds the Infinity Cache
  long ptotal;      // partials per slot (nrhs * nb, or the uneven systems' sum)
  const CGUSys* sys;  // uneven batches: system k's lattice and offsets
  const long* caps;   // uneven batches: system k's iteration cap
};

// Where a block's system lives, a kernel template argument.  MixOne: every
// system on the lattice of MixFix (a batch sharing one sigma mass; system
// blockIdx.y at offset blockIdx.y V).  MixUneven: system blockIdx.y on its
// own lattice, vectors, sigma mass and partials at the offsets of its
// descriptor (an uneven batch); a block past the system's sites leaves.
struct MixOne {
  static constexpr bool kUneven = false;
  Geom g;
  size_t off, soff, poff;
  int nb;
  GNX_DEVONLY bool init(const struct MixFix& a);
};
struct MixUneven {
  static constexpr bool kUneven = true;
  Geom g;
  size_t off, soff, poff;
  int nb;
  GNX_DEVONLY bool init(const struct MixFix& a);
};
GNX_DEVONLY bool MixOne::init(const MixFix& a) {
  g = a.g;
  nb = a.nb;
  off = (size_t)blockIdx.y * a.g.V;
  soff = 0;
  poff = (size_t)blockIdx.y * a.nb;
  return true;
}
GNX_DEVONLY bool MixUneven::init(const MixFix& a) {
  const CGUSys d = a.sys[blockIdx.y];
  g = d.g;
  nb = d.nb;
  off = (size_t)d.voff;
  soff = (size_t)d.soff;
  poff = (size_t)d.poff;
  return (int)blockIdx.x < d.nb;
}

GNX_DEV double* mslot(const MixFix& a, int s, int rhs) { return a.part + ((size_t)s * a.nrhs + rhs) * a.nb; }
// slot s of the block's system
template <typename SV>
GNX_DEVONLY double* mslot_v(const MixFix& a, const SV& v, int s) {
  return a.part + (size_t)s * a.ptotal + v.poff;
}

GNX_DEVONLY void* mix_p(const MixFix& a, int k) { return (k & 1) ? a.p1 : a.p0; }

// Slots a finished system used (its polled state): one pair per normal step
// (iters), one per reliable update (nrel, the first update of a scaled
// format's start included), and the pair of the X kernel that stops it.
static int mix_slots_of(const char* state) {
  const MixState* s = (const MixState*)state;
  return s->iters + s->nrel + 2;
}

GNX_DEVONLY double dnorm2(float2 v) { return (double)v.x * v.x + (double)v.y * v.y; }

// ---- start kernels ------------------------------------------------------------
// Every kernel of a solve uses the same SPT (sites per thread), so the
// per-block partial slots have one layout: nb = ceil(V / (SPT * MIX_BS)).
// Thread sites: f = blockIdx.x * SPT * MIX_BS + i * MIX_BS + threadIdx.x.
template <int SPT, bool XCD, typename SV>
GNX_DEVONLY int mix_site(const SV& v, int i) {
  return SiteMap<SPT, XCD>(v.nb).f(v.g, i);
}

// imposed start: x = 0, r = p0 = (S)b, xs = 0, sm and b copied; partial |b|^2
// (an uneven batch passes the solver's own gathered sm and b: nothing to copy)
template <typename S, int SPT, bool XCD, typename SV = MixOne>
__global__ __launch_bounds__(MIX_BS) void k_mix_start_imposed(MixFix a, const double* __restrict__ sm_in,
                                                              const double2* __restrict__ b_in, double tol,
                                                              double delta, long max_iter) {
  __shared__ double lds[MIX_BS / 64];
  SV v;
  if (!v.init(a)) return;
  const int rhs = blockIdx.y;
  const size_t off = v.off;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = mix_site<SPT, XCD>(v, i);
    if (f < 0) break;
    double2 bb = b_in[off + f];
    if (!SV::kUneven) a.b[off + f] = bb;
    a.x[off + f] = make_double2(0.0, 0.0);
    // a scaled format stores r = s b in the first slot (a reliable update),
    // once ||b|| is known; zeros here keep every loaded value finite
    float2 bs = S::kScaled ? make_float2(0.f, 0.f) : make_float2((float)bb.x, (float)bb.y);
    S::st((typename S::T*)a.r + off, f, bs, a.nt);
    S::st((typename S::T*)a.p0 + off, f, bs, a.nt);
    S::st((typename S::T*)a.p1 + off, f, make_float2(0.f, 0.f), a.nt);
    a.xs[off + f] = make_float2(0.f, 0.f);
    if (SV::kUneven || rhs == 0) {
      if (!SV::kUneven) a.sm[f] = sm_in[f];
      a.sm32[v.soff + f] = (float)sm_in[v.soff + f];
    }
    acc += cnorm2(bb);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) mslot_v(a, v, 1)[partial_slot(blockIdx.x, v.nb)] = acc;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MixState cfg = {};
    cfg.scale = cfg.scale_next = 1.0;
    cfg.start = GNX_START_IMPOSED;
    cfg.tol = tol;
    cfg.delta = delta;
    cfg.max_iter = SV::kUneven ? a.caps[rhs] : max_iter;
    put_mstate(a.state + rhs, 1, 0, PH_NORMAL, 0, 0, 0, 0.0, 0.0, cfg);
  }
}

// null start (conjugategradient.c:18-33) pass 1: u = M^dag x_in, x = x_in,
// b and sm copied
template <int SPT, bool XCD, typename SV = MixOne>
__global__ __launch_bounds__(MIX_BS) void k_mix_null_p1(MixFix a, const double* __restrict__ sm_in,
                                                         const double2* __restrict__ b_in,
                                                         const double2* __restrict__ xin) {
  SV v;
  if (!v.init(a)) return;
  const size_t off = v.off;
  const double2* xv = xin + off;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = mix_site<SPT, XCD>(v, i);
    if (f < 0) break;
    a.u[off + f] = apply_Mdag(v.g, decode(v.g, f), sm_in[v.soff + f], [&](int q) { return xv[q]; });
    if (!SV::kUneven) {
      a.x[off + f] = xv[f];
      a.b[off + f] = b_in[off + f];
    }
    if (SV::kUneven || blockIdx.y == 0) {
      if (!SV::kUneven) a.sm[f] = sm_in[f];
      a.sm32[v.soff + f] = (float)sm_in[v.soff + f];
    }
  }
}

// pass 2: partials |b - M u|^2 (slot 0) and |b|^2 (slot 1)
template <int SPT, bool XCD, typename SV = MixOne>
__global__ __launch_bounds__(MIX_BS) void k_mix_null_p2(MixFix a) {
  __shared__ double lds[MIX_BS / 64];
  SV v;
  if (!v.init(a)) return;
  const size_t off = v.off;
  const double2* u = a.u + off;
  double ar = 0.0, ab = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = mix_site<SPT, XCD>(v, i);
    if (f < 0) break;
    double2 bb = a.b[off + f];
    double2 rr = apply_sub_M(v.g, decode(v.g, f), a.sm[v.soff + f], bb, [&](int q) { return u[q]; });
    ar += cnorm2(rr);
    ab += cnorm2(bb);
  }
  ar = block_sum(ar, lds);
  ab = block_sum(ab, lds);
  if (threadIdx.x == 0) {
    mslot_v(a, v, 0)[partial_slot(blockIdx.x, v.nb)] = ar;
    mslot_v(a, v, 1)[partial_slot(blockIdx.x, v.nb)] = ab;
  }
}

// pass 3: an early exit keeps x_in, otherwise as the imposed start (slot 1 kept)
template <typename S, int SPT, bool XCD, typename SV = MixOne>
__global__ __launch_bounds__(MIX_BS) void k_mix_null_p3(MixFix a, double tol, double delta, long max_iter) {
  __shared__ double lds[MIX_BS / 64];
  SV v;
  if (!v.init(a)) return;
  const int rhs = blockIdx.y;
  const size_t off = v.off;
  const double nr = block_sum(partial_share(mslot_v(a, v, 0), v.nb), lds);
  const bool early = sqrt(nr) < tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MixState cfg = {};
    cfg.scale = cfg.scale_next = 1.0;
    cfg.start = GNX_START_NULL;
    cfg.tol = tol;
    cfg.delta = delta;
    cfg.max_iter = SV::kUneven ? a.caps[rhs] : max_iter;
    put_mstate(a.state + rhs, early ? 0 : 1, 0, PH_NORMAL, 0, early ? 1 : 0, early ? 1 : 0, 0.0, 0.0, cfg);
    if (early) {
      CGDevResult r = {1, 1, 0, 0, sqrt(nr), {0, 0, 0}};
      a.res[rhs] = r;
    }
  }
  if (early) return;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = mix_site<SPT, XCD>(v, i);
    if (f < 0) break;
    double2 bb = a.b[off + f];
    a.x[off + f] = make_double2(0.0, 0.0);
    float2 bs = S::kScaled ? make_float2(0.f, 0.f) : make_float2((float)bb.x, (float)bb.y);
    S::st((typename S::T*)a.r + off, f, bs, a.nt);
    S::st((typename S::T*)a.p0 + off, f, bs, a.nt);
    S::st((typename S::T*)a.p1 + off, f, make_float2(0.f, 0.f), a.nt);
    a.xs[off + f] = make_float2(0.f, 0.f);
  }
}

// ---- iteration kernels -----------------------------------------------------------
// odd: slot parity (p_old = odd ? p1 : p0); sin/sout state halves, pin/pout slots
template <typename S, int SPT, bool XCD>
__global__ __launch_bounds__(MIX_BS) void k_mix_X(MixFix a, int odd, const MixState* __restrict__ sin,
                                                  MixState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[MIX_BS / 64];
  using T = typename S::T;
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const T* r = (const T*)a.r + off;
  const T* pold = (const T*)(odd ? a.p1 : a.p0) + off;
  T* pnew = (T*)(odd ? a.p0 : a.p1) + off;
  // loads of the normal step, issued before the reduction
  const MixState s = sin[rhs];
  const double share = partial_share(pin + (size_t)rhs * a.nb, a.nb);
  int fs[SPT];
  SiteCode kc[SPT];
  float2 rv[SPT][7], pv[SPT][7];
  float smv[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = mix_site<SPT, XCD>(a, i);
    if (fs[i] < 0) continue;
    const SiteGeo sgeo = site_geo(g, fs[i]);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fs[i], n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = a.sm32[fs[i]];
#pragma unroll
    for (int k = 0; k < 7; k++) {
      rv[i][k] = S::ld(r, q[k]);
      pv[i][k] = S::ld(pold, q[k]);
    }
  }
  const double rho = block_sum(share, lds);
  const double nr = sqrt(rho);
  MixState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return;
  }
  double beta = 0.0;
  double rmax = s.rmax;
  int nrel = s.nrel;
  bool update = false;
  if (s.phase == PH_UPDATED) {
    // rho is the true FP64 residual norm^2 after a reliable update
    nrel = s.nrel + 1;
    if (nr < s.tol || s.iters >= s.max_iter) {
      const int conv = nr < s.tol ? 1 : 0;
      if (lead) {
        put_mstate(out, 0, s.iters, PH_NORMAL, nrel, conv, 0, rho, nr, s);
        CGDevResult res = {conv, 0, s.iters, nrel, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    rmax = nr;
    beta = rho / s.rho;
  } else if (s.iters == 0) {
    if (s.start == GNX_START_IMPOSED && nr < s.tol) {  // conjugategradient.c:123-128
      if (lead) {
        put_mstate(out, 0, 0, PH_NORMAL, 0, 1, 1, rho, nr, s);
        CGDevResult res = {1, 1, 0, 0, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    rmax = nr;
  } else {
    // fold the correction and test the true residual when the iterated one
    // has dropped by delta, claims convergence, or the iteration cap is hit
    update = (nr < s.delta * s.rmax) || (nr < s.tol) || (s.iters >= s.max_iter);
    beta = rho / s.rho;
  }
  if (update) {
    // reliable update, part 1: u = M^dag (x + xs) in FP64 (x and xs are not
    // written in this kernel, so neighbours read consistent values); the
    // direction is carried into the other parity buffer unchanged
    const double2* x = a.x + off;
    const float2* xs = a.xs + off;
    auto ldx = [&](int q) {
      float2 c = xs[q];
      return cadd(x[q], make_double2((double)c.x, (double)c.y));
    };
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      if (fs[i] < 0) continue;
      a.u[off + fs[i]] = apply_Mdag(g, decode(g, fs[i]), a.sm[fs[i]], ldx);
      S::st(pnew, fs[i], pv[i][0], a.nt);
    }
    if (lead) put_mstate(out, 1, s.iters, PH_UPDATE, nrel, 0, 0, s.rho, rmax, s);
    return;
  }
  // normal step: p = r + beta p, t = M^dag p
  const float fb = (float)beta;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const SiteCode st = kc[i];
    float2 p0v = cfma(fb, pv[i][0], rv[i][0]);
    float2 ppv[3], pmv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      ppv[d] = cfma(fb, pv[i][1 + d], rv[i][1 + d]);
      pmv[d] = cfma(fb, pv[i][4 + d], rv[i][4 + d]);
    }
    S::st(pnew, fs[i], p0v, a.nt);
    float2 tt = apply_Mdag_fast_vals(st, smv[i], p0v, ppv, pmv);
    S::st((T*)a.t + off, fs[i], tt, a.nt);
    acc += dnorm2(tt);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) pout[(size_t)rhs * a.nb + partial_slot(blockIdx.x, a.nb)] = acc;
  if (lead) put_mstate(out, 1, s.iters, PH_NORMAL, nrel, 0, 0, rho, rmax, s);
}

template <typename S, int SPT, bool XCD>
__global__ __launch_bounds__(MIX_BS) void k_mix_Y(MixFix a, int odd, const MixState* __restrict__ sin,
                                                  MixState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[MIX_BS / 64];
  using T = typename S::T;
  const Geom g = a.g;
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * g.V;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const MixState s = sin[rhs];
  const double share = partial_share(pin + (size_t)rhs * a.nb, a.nb);
  MixState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return;
  }
  double acc = 0.0;
  if (s.phase == PH_UPDATE) {
    // reliable update, part 2: x += xs, xs = 0, r = (S)(b - M u)
    const double2* u = a.u + off;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      const int f = mix_site<SPT, XCD>(a, i);
      if (f < 0) break;
      float2 c = a.xs[off + f];
      a.x[off + f] = cadd(a.x[off + f], make_double2((double)c.x, (double)c.y));
      a.xs[off + f] = make_float2(0.f, 0.f);
      double2 rr = apply_sub_M(g, decode(g, f), a.sm[f], a.b[off + f], [&](int q) { return u[q]; });
      S::st((T*)a.r + off, f, make_float2((float)rr.x, (float)rr.y), a.nt);
      acc += cnorm2(rr);
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) pout[(size_t)rhs * a.nb + partial_slot(blockIdx.x, a.nb)] = acc;
    if (lead) put_mstate(out, 1, s.iters, PH_UPDATED, s.nrel, 0, 0, s.rho, s.rmax, s);
    return;
  }
  // normal step: loads first, then gamma
  const T* pnew = (const T*)(odd ? a.p0 : a.p1) + off;
  const T* t = (const T*)a.t + off;
  T* r = (T*)a.r + off;
  int fs[SPT];
  float2 w[SPT], pv[SPT], xv[SPT], rv[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = mix_site<SPT, XCD>(a, i);
    if (fs[i] < 0) continue;
    w[i] = apply_M(g, decode(g, fs[i]), a.sm32[fs[i]], [&](int q) { return S::ld(t, q); });
    pv[i] = S::ld(pnew, fs[i]);
    xv[i] = a.xs[off + fs[i]];
    rv[i] = S::ld(r, fs[i]);
  }
  const double gamma = block_sum(share, lds);
  const float al = (float)(s.rho / gamma);
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    st_mix(&a.xs[off + fs[i]], cfma(al, pv[i], xv[i]), a.nt);
    float2 rn = cfma(-al, w[i], rv[i]);
    S::st(r, fs[i], rn, a.nt);
    acc += dnorm2(rn);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) pout[(size_t)rhs * a.nb + partial_slot(blockIdx.x, a.nb)] = acc;
  if (lead) put_mstate(out, 1, s.iters + 1, PH_NORMAL, s.nrel, 0, 0, s.rho, s.rmax, s);
}

// ---- iteration kernels, loads-first form ------------------------------------------
// The steps of k_mix_X / k_mix_Y with two changes.
// (1) Every load of a kernel (state, partials, then all operands of the
// normal step, whatever the phase) is issued before its first wait, and the
// partial reduction runs while the operands are in flight.  This holds every
// operand in registers at once (lower occupancy), so which form is faster
// depends on the storage format and batch size: the tuner times both.
// (2) The FP32 correction xs is not touched by the Y kernel: Y leaves its
// step length alpha pending in the state, and the X kernel of every even
// slot adds both pending terms, xs += alpha_b p_prev + alpha_a p_old, in
// that order, while it reads p_old anyway.  Y reads only t and r (the
// 16-byte xs read-modify-write and the p read leave the every-slot path),
// X gains xs and p_prev on every other slot.  A reliable update folds the
// pending terms into x: its u = M^dag (x + c) and the following x += c use
// the same c = xs + alpha_b p_prev + alpha_a p_old, computed alike.
// Two direction buffers, by slot parity: p_prev is p[sl+1], the buffer X
// writes, so it is read only at the thread's own sites, before that thread's
// store (a third and fourth buffer would push a batch of 32 at 64^3 past the
// 256 MB Infinity Cache); an update's X gathers p_prev at the neighbours and
// so writes no direction, its Y carries p_old into p[sl+1] after the fold.
template <typename S, int SPT, bool XCD, bool FLUSH, typename SV = MixOne>
__global__ __launch_bounds__(MIX_BS) void k_mix_Xe(MixFix a, int sl, const MixState* __restrict__ sin,
                                                  MixState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[MIX_BS / 64], lds2[MIX_BS / 64];  // head and tail reductions
  using T = typename S::T;
  SV v;
  if (!v.init(a)) return;
  const Geom g = v.g;
  const int rhs = blockIdx.y;
  const size_t off = v.off;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const T* r = (const T*)a.r + off;
  const T* pold = (const T*)mix_p(a, sl) + off;
  const T* pprev = (const T*)mix_p(a, sl + 1) + off;  // = pnew: own sites only, read before the store
  T* pnew = (T*)mix_p(a, sl + 1) + off;
  // loads in the order they are needed (state and partials, then the
  // normal step's operands), all issued before the first wait; the
  // reduction runs while the operands are in flight
  const MixState s = load_state(sin + rhs);
  double pl[PART_K];
  partial_issue(pin + v.poff, v.nb, pl);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  float2 rv[SPT][7], pv[SPT][7], xsv[SPT], ppr[SPT];
  float smv[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = mix_site<SPT, XCD>(v, i);
    const int fc = fs[i] < 0 ? 0 : fs[i];  // past the last site: load site 0, store nothing
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm32 + v.soff, fc);
#pragma unroll
    for (int k = 0; k < 7; k++) {
      rv[i][k] = S::ld(r, q[k]);
      pv[i][k] = S::ld(pold, q[k]);
    }
    if (FLUSH) {
      xsv[i] = *elem(a.xs + off, fc);
      ppr[i] = S::ld(pprev, fc);
    }
  }
  end_of_loads();
  const double rho = block_sum_once(partial_finish(pin + v.poff, v.nb, pl), lds);
  const double nr = S::kScaled ? sqrt(rho) / s.scale : sqrt(rho);  // the unscaled norm
  MixState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return;
  }
  double beta = 0.0;
  double rmax = s.rmax;
  int nrel = s.nrel;
  bool update = false;
  if (s.phase == PH_UPDATED) {
    // rho is the true FP64 residual norm^2 after a reliable update
    nrel = s.nrel + 1;
    if (nr < s.tol || s.iters >= s.max_iter) {
      const int conv = nr < s.tol ? 1 : 0;
      if (lead) {
        put_mstate(out, 0, s.iters, PH_NORMAL, nrel, conv, 0, rho, nr, s);
        CGDevResult res = {conv, 0, s.iters, nrel, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    rmax = nr;
    // a scaled format's first update is the start (the direction is zero)
    beta = (S::kScaled && s.iters == 0 && s.nrel == 0) ? 0.0 : rho / s.rho;
  } else if (s.iters == 0) {
    if (s.start == GNX_START_IMPOSED && nr < s.tol) {  // conjugategradient.c:123-128
      if (lead) {
        put_mstate(out, 0, 0, PH_NORMAL, 0, 1, 1, rho, nr, s);
        CGDevResult res = {1, 1, 0, 0, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    rmax = nr;
    update = S::kScaled;  // scaled storage: the first slot stores r = s b as an update
  } else {
    // fold the correction and test the true residual when the iterated one
    // has dropped by delta, claims convergence, or the iteration cap is hit
    update = (nr < s.delta * s.rmax) || (nr < s.tol) || (s.iters >= s.max_iter);
    beta = rho / s.rho;
  }
  const float fa = (float)s.alpha_a, fp = (float)s.alpha_b;
  if (update) {
    // reliable update, part 1: u = M^dag (x + c) in FP64 with the pending
    // terms in c (x, xs and both directions are not written in this kernel,
    // so neighbours read consistent values); the pending alphas stay for Y
    const double2* x = a.x + off;
    const float2* xs = a.xs + off;
    auto ldx = [&](int q) {
      float2 c = cfma(fa, S::ld(pold, q), cfma(fp, S::ld(pprev, q), xs[q]));
      return cadd(x[q], make_double2((double)c.x, (double)c.y));
    };
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      if (fs[i] < 0) continue;
      a.u[off + fs[i]] = apply_Mdag(g, decode(g, fs[i]), a.sm[v.soff + fs[i]], ldx);
    }
    if (lead) {
      put_mstate(out, 1, s.iters, PH_UPDATE, nrel, 0, 0, s.rho, rmax, s);
      // scale of the coming r: its elements O(1) if ||r|| stays near ||r_it||
      if (S::kScaled) out->scale_next = nr > 0.0 ? sqrt((double)g.V) / nr : s.scale;
    }
    return;
  }
  // normal step: (even slots) xs += alpha_b p_prev + alpha_a p_old;
  // p = r + beta p, t = M^dag p
  const float fb = (float)beta;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    if (FLUSH) st_mix(elem(a.xs + off, fs[i]), cfma(fa, pv[i][0], cfma(fp, ppr[i], xsv[i])), a.nt);
    const SiteCode st = kc[i];
    float2 p0v = cfma(fb, pv[i][0], rv[i][0]);
    float2 ppv[3], pmv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      ppv[d] = cfma(fb, pv[i][1 + d], rv[i][1 + d]);
      pmv[d] = cfma(fb, pv[i][4 + d], rv[i][4 + d]);
    }
    S::st(pnew, fs[i], p0v, a.nt);
    float2 tt = apply_Mdag_fast_vals(st, smv[i], p0v, ppv, pmv);
    S::st((T*)a.t + off, fs[i], tt, a.nt);
    acc += dnorm2(tt);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[v.poff + partial_slot(blockIdx.x, v.nb)] = acc;
  if (lead) {
    put_mstate(out, 1, s.iters, PH_NORMAL, nrel, 0, 0, rho, rmax, s);
    if (FLUSH) put_alphas(out, 0.0, 0.0);
  }
}

template <typename S, int SPT, bool XCD, typename SV = MixOne>
__global__ __launch_bounds__(MIX_BS) void k_mix_Ye(MixFix a, int sl, const MixState* __restrict__ sin,
                                                  MixState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[MIX_BS / 64], lds2[MIX_BS / 64];  // head and tail reductions
  using T = typename S::T;
  SV v;
  if (!v.init(a)) return;
  const Geom g = v.g;
  const int rhs = blockIdx.y;
  const size_t off = v.off;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const T* t = (const T*)a.t + off;
  T* r = (T*)a.r + off;
  // loads as in k_mix_Xe: state and partials, then the normal step's
  // operands, issued whatever the phase (an update or a finished system
  // ignores them), so that none waits for the state
  const MixState s = load_state(sin + rhs);
  double pl[PART_K];
  partial_issue(pin + v.poff, v.nb, pl);
  end_of_loads();
  int fs[SPT];
  SiteCode kc[SPT];
  float smv[SPT];
  float2 tv[SPT][7], rv[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = mix_site<SPT, XCD>(v, i);
    const int fc = fs[i] < 0 ? 0 : fs[i];
    const SiteGeo sgeo = site_geo(g, fc);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fc, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(a.sm32 + v.soff, fc);
#pragma unroll
    for (int k = 0; k < 7; k++) tv[i][k] = S::ld(t, q[k]);
    rv[i] = S::ld(r, fc);
  }
  end_of_loads();
  const double gamma = block_sum_once(partial_finish(pin + v.poff, v.nb, pl), lds);
  MixState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return;
  }
  double acc = 0.0;
  if (s.phase == PH_UPDATE) {
    // reliable update, part 2: x += c (the c of part 1), xs = 0,
    // r = (S)(b - M u); the direction is carried into p[sl+1] (p_prev's
    // buffer, read first) unchanged
    const double2* u = a.u + off;
    const T* pold = (const T*)mix_p(a, sl) + off;
    T* pnext = (T*)mix_p(a, sl + 1) + off;
    const float fa = (float)s.alpha_a, fp = (float)s.alpha_b;
    // scaled storage: r is stored at the new scale, the direction and the
    // previous rho are converted to it
    const double sn = S::kScaled ? s.scale_next : 1.0;
    const double ratio = S::kScaled ? s.scale_next / s.scale : 1.0;
    const float fr = (float)ratio;
#pragma unroll
    for (int i = 0; i < SPT; i++) {
      const int f = fs[i];
      if (f < 0) break;
      const float2 po = S::ld(pold, f);
      const float2 c = cfma(fa, po, cfma(fp, S::ld(pnext, f), a.xs[off + f]));
      S::st(pnext, f, S::kScaled ? make_float2(fr * po.x, fr * po.y) : po, a.nt);
      a.x[off + f] = cadd(a.x[off + f], make_double2((double)c.x, (double)c.y));
      a.xs[off + f] = make_float2(0.f, 0.f);
      double2 rr = apply_sub_M(g, decode(g, f), a.sm[v.soff + f], a.b[off + f], [&](int q) { return u[q]; });
      if (S::kScaled) rr = make_double2(sn * rr.x, sn * rr.y);
      S::st(r, f, make_float2((float)rr.x, (float)rr.y), a.nt);
      acc += cnorm2(rr);
    }
    acc = block_sum_once(acc, lds2);
    if (threadIdx.x == 0) pout[v.poff + partial_slot(blockIdx.x, v.nb)] = acc;
    if (lead) {
      put_mstate(out, 1, s.iters, PH_UPDATED, s.nrel, 0, 0, s.rho * ratio * ratio, s.rmax, s);
      put_alphas(out, 0.0, 0.0);
      if (S::kScaled) out->scale = sn;
    }
    return;
  }
  // normal step: r -= alpha M t; alpha p is left pending for the X kernel
  const float al = (float)(s.rho / gamma);
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    if (fs[i] < 0) continue;
    const float2 w = apply_M_vals(kc[i], smv[i], tv[i][0], &tv[i][1], &tv[i][4]);
    float2 rn = cfma(-al, w, rv[i]);
    S::st(r, fs[i], rn, a.nt);
    acc += dnorm2(rn);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[v.poff + partial_slot(blockIdx.x, v.nb)] = acc;
  if (lead) {
    put_mstate(out, 1, s.iters + 1, PH_NORMAL, s.nrel, 0, 0, s.rho, s.rmax, s);
    // pending step length of the unscaled direction: alpha / s
    put_alphas(out, S::kScaled ? (double)al / s.scale : (double)al, s.alpha_a);
  }
}

// ---- loads-first kernels on groups of four consecutive sites ------------------------
// For the 4-byte storage formats (BF16, FP16) on lattices whose time extent
// is a multiple of 8: a thread owns one group of four consecutive f-sites
// (gnx_group.h), so its operands of one vector are seven aligned 16-byte
// loads for four sites, against twenty-eight 4-byte loads in k_mix_Xe /
// k_mix_Ye.  Each site's arithmetic is theirs (the same helpers on the same
// operands, the same state logic and reliable updates); the partition of
// the global sums differs (a thread's four sites are consecutive), so the
// iterates are this shape's own.  The blocks and partial layout are those of
// SiteMap<4, true>, so the start kernels of that shape are shared.
// Measured against the loads-first shapes at 64^3 and 96^3 (BF16, FP16; 1,
// 16 and 32 systems) it is never more than 2 % faster than the rule's
// shape and up to 6 % slower on batches: the gathers stay bound by their
// cache lines, not by the number of load instructions.  Reached by pin or
// by the tuner, not by the rule.
GNX_DEVONLY gnx_u4 ld_u4(const unsigned int* p, int i) { return *(const gnx_u4*)elem(p, i); }
GNX_DEVONLY void st_u4(unsigned int* p, int i, gnx_u4 v, int nt) {
  if (nt)
    __builtin_nontemporal_store(v, (gnx_u4*)elem(p, i));
  else
    *(gnx_u4*)elem(p, i) = v;
}
typedef float gnx_f4v __attribute__((ext_vector_type(4)));

template <typename S, bool FLUSH>
__global__ __launch_bounds__(MIX_BS) void k_mix_Xq(MixFix a, int sl, const MixState* __restrict__ sin,
                                                  MixState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[MIX_BS / 64], lds2[MIX_BS / 64];  // head and tail reductions
  using T = typename S::T;
  static_assert(sizeof(T) == 4, "four sites per 16-byte load");
  MixOne v;
  v.init(a);
  const Geom g = v.g;
  const int rhs = blockIdx.y;
  const size_t off = v.off;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const T* r = (const T*)a.r + off;
  const T* pold = (const T*)mix_p(a, sl) + off;
  const T* pprev = (const T*)mix_p(a, sl + 1) + off;  // = pnew: own sites only, read before the store
  T* pnew = (T*)mix_p(a, sl + 1) + off;
  const MixState s = load_state(sin + rhs);
  double pl[PART_K];
  partial_issue(pin + v.poff, v.nb, pl);
  end_of_loads();
  const int f0 = group_first(g, v.nb, blockIdx.x, threadIdx.x);
  const int fc = f0 < 0 ? 0 : f0;  // past the last group: load group 0, store nothing
  const GroupGeo q = group_geo(g, fc);
  const int gi[7] = {q.own, q.pp0, q.mm0, q.pp1, q.mm1, q.g2, q.s2};
  const gnx_f4v sm4 = *(const gnx_f4v*)elem(a.sm32 + v.soff, fc);
  gnx_u4 rw[7], pw[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    rw[k] = ld_u4(r, gi[k]);
    pw[k] = ld_u4(pold, gi[k]);
  }
  gnx_f4v xs01 = {0.f, 0.f, 0.f, 0.f}, xs23 = {0.f, 0.f, 0.f, 0.f};
  gnx_u4 ppw = {0u, 0u, 0u, 0u};
  if (FLUSH) {
    xs01 = *(const gnx_f4v*)elem(a.xs + off, fc);
    xs23 = *(const gnx_f4v*)elem(a.xs + off, fc + 2);
    ppw = ld_u4(pprev, fc);
  }
  end_of_loads();
  const double rho = block_sum_once(partial_finish(pin + v.poff, v.nb, pl), lds);
  const double nr = S::kScaled ? sqrt(rho) / s.scale : sqrt(rho);  // the unscaled norm
  MixState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return;
  }
  double beta = 0.0;
  double rmax = s.rmax;
  int nrel = s.nrel;
  bool update = false;
  if (s.phase == PH_UPDATED) {
    nrel = s.nrel + 1;
    if (nr < s.tol || s.iters >= s.max_iter) {
      const int conv = nr < s.tol ? 1 : 0;
      if (lead) {
        put_mstate(out, 0, s.iters, PH_NORMAL, nrel, conv, 0, rho, nr, s);
        CGDevResult res = {conv, 0, s.iters, nrel, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    rmax = nr;
    beta = (S::kScaled && s.iters == 0 && s.nrel == 0) ? 0.0 : rho / s.rho;
  } else if (s.iters == 0) {
    if (s.start == GNX_START_IMPOSED && nr < s.tol) {  // conjugategradient.c:123-128
      if (lead) {
        put_mstate(out, 0, 0, PH_NORMAL, 0, 1, 1, rho, nr, s);
        CGDevResult res = {1, 1, 0, 0, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return;
    }
    rmax = nr;
    update = S::kScaled;
  } else {
    update = (nr < s.delta * s.rmax) || (nr < s.tol) || (s.iters >= s.max_iter);
    beta = rho / s.rho;
  }
  const float fa = (float)s.alpha_a, fp = (float)s.alpha_b;
  if (update) {
    // reliable update, part 1, per site as in k_mix_Xe
    const double2* x = a.x + off;
    const float2* xs = a.xs + off;
    auto ldx = [&](int qq) {
      float2 c = cfma(fa, S::ld(pold, qq), cfma(fp, S::ld(pprev, qq), xs[qq]));
      return cadd(x[qq], make_double2((double)c.x, (double)c.y));
    };
    if (f0 >= 0) {
#pragma unroll
      for (int j = 0; j < 4; j++)
        a.u[off + f0 + j] = apply_Mdag(g, decode(g, f0 + j), a.sm[v.soff + f0 + j], ldx);
    }
    if (lead) {
      put_mstate(out, 1, s.iters, PH_UPDATE, nrel, 0, 0, s.rho, rmax, s);
      if (S::kScaled) out->scale_next = nr > 0.0 ? sqrt((double)g.V) / nr : s.scale;
    }
    return;
  }
  // normal step: (even slots) xs += alpha_b p_prev + alpha_a p_old;
  // p = r + beta p, t = M^dag p
  const float fb = (float)beta;
  double acc = 0.0;
  if (f0 >= 0) {
    T ro[4][7], po[4][7];
    group_operands<T>(rw, q.even, ro);
    group_operands<T>(pw, q.even, po);
    const T ppr[4] = {ppw.x, ppw.y, ppw.z, ppw.w};
    const float2 xsv[4] = {make_float2(xs01.x, xs01.y), make_float2(xs01.z, xs01.w), make_float2(xs23.x, xs23.y),
                           make_float2(xs23.z, xs23.w)};
    const float smv[4] = {sm4.x, sm4.y, sm4.z, sm4.w};
    float2 xn[4];
    T pn[4], tn[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (FLUSH) xn[j] = cfma(fa, S::cvt(po[j][0]), cfma(fp, S::cvt(ppr[j]), xsv[j]));
      float2 pv[7];
#pragma unroll
      for (int k = 0; k < 7; k++) pv[k] = cfma(fb, S::cvt(po[j][k]), S::cvt(ro[j][k]));
      pn[j] = S::pack(pv[0]);
      const float2 tt = apply_Mdag_fast_vals(q.k[j], smv[j], pv[0], &pv[1], &pv[4]);
      tn[j] = S::pack(tt);
      acc += dnorm2(tt);
    }
    if (FLUSH) {
      const gnx_f4v w01 = {xn[0].x, xn[0].y, xn[1].x, xn[1].y}, w23 = {xn[2].x, xn[2].y, xn[3].x, xn[3].y};
      if (a.nt) {
        __builtin_nontemporal_store(w01, (gnx_f4v*)elem(a.xs + off, f0));
        __builtin_nontemporal_store(w23, (gnx_f4v*)elem(a.xs + off, f0 + 2));
      } else {
        *(gnx_f4v*)elem(a.xs + off, f0) = w01;
        *(gnx_f4v*)elem(a.xs + off, f0 + 2) = w23;
      }
    }
    const gnx_u4 pn4 = {pn[0], pn[1], pn[2], pn[3]}, tn4 = {tn[0], tn[1], tn[2], tn[3]};
    st_u4(pnew, f0, pn4, a.nt);
    st_u4((T*)a.t + off, f0, tn4, a.nt);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[v.poff + partial_slot(blockIdx.x, v.nb)] = acc;
  if (lead) {
    put_mstate(out, 1, s.iters, PH_NORMAL, nrel, 0, 0, rho, rmax, s);
    if (FLUSH) put_alphas(out, 0.0, 0.0);
  }
}

template <typename S>
__global__ __launch_bounds__(MIX_BS) void k_mix_Yq(MixFix a, int sl, const MixState* __restrict__ sin,
                                                  MixState* __restrict__ sout, const double* __restrict__ pin,
                                                  double* __restrict__ pout) {
  __shared__ double lds[MIX_BS / 64], lds2[MIX_BS / 64];  // head and tail reductions
  using T = typename S::T;
  static_assert(sizeof(T) == 4, "four sites per 16-byte load");
  MixOne v;
  v.init(a);
  const Geom g = v.g;
  const int rhs = blockIdx.y;
  const size_t off = v.off;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const T* t = (const T*)a.t + off;
  T* r = (T*)a.r + off;
  const MixState s = load_state(sin + rhs);
  double pl[PART_K];
  partial_issue(pin + v.poff, v.nb, pl);
  end_of_loads();
  const int f0 = group_first(g, v.nb, blockIdx.x, threadIdx.x);
  const int fc = f0 < 0 ? 0 : f0;
  const GroupGeo q = group_geo(g, fc);
  const int gi[7] = {q.own, q.pp0, q.mm0, q.pp1, q.mm1, q.g2, q.s2};
  const gnx_f4v sm4 = *(const gnx_f4v*)elem(a.sm32 + v.soff, fc);
  gnx_u4 tw[7];
#pragma unroll
  for (int k = 0; k < 7; k++) tw[k] = ld_u4(t, gi[k]);
  const gnx_u4 rw = ld_u4(r, fc);
  end_of_loads();
  const double gamma = block_sum_once(partial_finish(pin + v.poff, v.nb, pl), lds);
  MixState* out = sout + rhs;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return;
  }
  double acc = 0.0;
  if (s.phase == PH_UPDATE) {
    // reliable update, part 2, per site as in k_mix_Ye
    const double2* u = a.u + off;
    const T* pold = (const T*)mix_p(a, sl) + off;
    T* pnext = (T*)mix_p(a, sl + 1) + off;
    const float fa = (float)s.alpha_a, fp = (float)s.alpha_b;
    const double sn = S::kScaled ? s.scale_next : 1.0;
    const double ratio = S::kScaled ? s.scale_next / s.scale : 1.0;
    const float fr = (float)ratio;
    if (f0 >= 0) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int f = f0 + j;
        const float2 po = S::ld(pold, f);
        const float2 c = cfma(fa, po, cfma(fp, S::ld(pnext, f), a.xs[off + f]));
        S::st(pnext, f, S::kScaled ? make_float2(fr * po.x, fr * po.y) : po, a.nt);
        a.x[off + f] = cadd(a.x[off + f], make_double2((double)c.x, (double)c.y));
        a.xs[off + f] = make_float2(0.f, 0.f);
        double2 rr = apply_sub_M(g, decode(g, f), a.sm[v.soff + f], a.b[off + f], [&](int qq) { return u[qq]; });
        if (S::kScaled) rr = make_double2(sn * rr.x, sn * rr.y);
        S::st(r, f, make_float2((float)rr.x, (float)rr.y), a.nt);
        acc += cnorm2(rr);
      }
    }
    acc = block_sum_once(acc, lds2);
    if (threadIdx.x == 0) pout[v.poff + partial_slot(blockIdx.x, v.nb)] = acc;
    if (lead) {
      put_mstate(out, 1, s.iters, PH_UPDATED, s.nrel, 0, 0, s.rho * ratio * ratio, s.rmax, s);
      put_alphas(out, 0.0, 0.0);
      if (S::kScaled) out->scale = sn;
    }
    return;
  }
  // normal step: r -= alpha M t; alpha p is left pending for the X kernel
  const float al = (float)(s.rho / gamma);
  if (f0 >= 0) {
    T to[4][7];
    group_operands<T>(tw, q.even, to);
    const T rv[4] = {rw.x, rw.y, rw.z, rw.w};
    const float smv[4] = {sm4.x, sm4.y, sm4.z, sm4.w};
    T rn[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float2 tv[7];
#pragma unroll
      for (int k = 0; k < 7; k++) tv[k] = S::cvt(to[j][k]);
      const float2 w = apply_M_vals(q.k[j], smv[j], tv[0], &tv[1], &tv[4]);
      const float2 rr = cfma(-al, w, S::cvt(rv[j]));
      rn[j] = S::pack(rr);
      acc += dnorm2(rr);
    }
    const gnx_u4 rn4 = {rn[0], rn[1], rn[2], rn[3]};
    st_u4(r, f0, rn4, a.nt);
  }
  acc = block_sum_once(acc, lds2);
  if (threadIdx.x == 0) pout[v.poff + partial_slot(blockIdx.x, v.nb)] = acc;
  if (lead) {
    put_mstate(out, 1, s.iters + 1, PH_NORMAL, s.nrel, 0, 0, s.rho, s.rmax, s);
    put_alphas(out, S::kScaled ? (double)al / s.scale : (double)al, s.alpha_a);
  }
}

// ---- paired iteration kernels -------------------------------------------------------
// Two systems per block (blockIdx.y = pair r0 = 2 y, r0 + 1): the site decode,
// neighbours, phases and addresses are shared by the pair; each system keeps
// its own state, scalars and phase, decided uniformly per block, so the
// pair runs exactly the steps of k_mix_X / k_mix_Y on each system (same
// partial sums, same order: a paired solve is bit-identical to an unpaired one
// of the same sites-per-thread and mapping).
enum MixMode { MM_OFF = 0, MM_NORMAL = 1, MM_UPDATE = 2 };

// X-kernel decision of one system (the branches of k_mix_X); finished
// systems get their final state and result written here
GNX_DEVONLY int mix_x_decide(const MixFix& a, const MixState& s, int rhs, double rho, MixState* out, bool lead,
                             double* beta, double* rmax, int* nrel) {
  const double nr = sqrt(rho);
  *beta = 0.0;
  *rmax = s.rmax;
  *nrel = s.nrel;
  if (!s.active) {
    if (lead) put_mstate(out, 0, s.iters, s.phase, s.nrel, s.converged, s.early, s.rho, s.rmax, s);
    return MM_OFF;
  }
  if (s.phase == PH_UPDATED) {
    *nrel = s.nrel + 1;
    if (nr < s.tol || s.iters >= s.max_iter) {
      const int conv = nr < s.tol ? 1 : 0;
      if (lead) {
        put_mstate(out, 0, s.iters, PH_NORMAL, *nrel, conv, 0, rho, nr, s);
        CGDevResult res = {conv, 0, s.iters, *nrel, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return MM_OFF;
    }
    *rmax = nr;
    *beta = rho / s.rho;
    return MM_NORMAL;
  }
  if (s.iters == 0) {
    if (s.start == GNX_START_IMPOSED && nr < s.tol) {
      if (lead) {
        put_mstate(out, 0, 0, PH_NORMAL, 0, 1, 1, rho, nr, s);
        CGDevResult res = {1, 1, 0, 0, nr, {0, 0, 0}};
        a.res[rhs] = res;
      }
      return MM_OFF;
    }
    *rmax = nr;
    return MM_NORMAL;
  }
  *beta = rho / s.rho;
  return ((nr < s.delta * s.rmax) || (nr < s.tol) || (s.iters >= s.max_iter)) ? MM_UPDATE : MM_NORMAL;
}

template <typename S, int SPT, bool XCD>
__global__ __launch_bounds__(MIX_BS) void k_mix_X2(MixFix a, int odd, const MixState* __restrict__ sin,
                                                   MixState* __restrict__ sout, const double* __restrict__ pin,
                                                   double* __restrict__ pout) {
  __shared__ double lds[2 * (MIX_BS / 64)];
  using T = typename S::T;
  const Geom g = a.g;
  const int r0 = 2 * blockIdx.y;
  const bool two = r0 + 1 < a.nrhs;
  const int rr[2] = {r0, two ? r0 + 1 : r0};
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const MixState s0 = sin[rr[0]];
  const MixState s1 = sin[rr[1]];
  const double sh0 = partial_share(pin + (size_t)rr[0] * a.nb, a.nb);
  const double sh1 = two ? partial_share(pin + (size_t)rr[1] * a.nb, a.nb) : 0.0;
  const T* r[2];
  const T* pold[2];
  T* pnew[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const size_t off = (size_t)rr[k] * g.V;
    r[k] = (const T*)a.r + off;
    pold[k] = (const T*)(odd ? a.p1 : a.p0) + off;
    pnew[k] = (T*)(odd ? a.p0 : a.p1) + off;
  }
  int fs[SPT];
  SiteCode kc[SPT];
  float smv[SPT];
  float2 rv[2][SPT][7], pv[2][SPT][7];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = mix_site<SPT, XCD>(a, i);
    if (fs[i] < 0) continue;
    const SiteGeo sgeo = site_geo(g, fs[i]);
    const NbrF n = sgeo.n;
    kc[i] = sgeo.k;
    int q[7] = {fs[i], n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = a.sm32[fs[i]];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int j = 0; j < 7; j++) {
        rv[k][i][j] = S::ld(r[k], q[j]);
        pv[k][i][j] = S::ld(pold[k], q[j]);
      }
  }
  const double2 rho = block_sum2(sh0, sh1, lds);
  const double rhok[2] = {rho.x, rho.y};
  const MixState* sk[2] = {&s0, &s1};
  int mode[2];
  double beta[2], rmax[2];
  int nrel[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (k == 1 && !two) {
      mode[1] = MM_OFF;
      continue;
    }
    mode[k] = mix_x_decide(a, *sk[k], rr[k], rhok[k], sout + rr[k], lead, &beta[k], &rmax[k], &nrel[k]);
  }
  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const size_t off = (size_t)rr[k] * g.V;
    if (mode[k] == MM_UPDATE) {
      // reliable update, part 1 (as k_mix_X)
      const double2* x = a.x + off;
      const float2* xs = a.xs + off;
      auto ldx = [&](int q) {
        float2 c = xs[q];
        return cadd(x[q], make_double2((double)c.x, (double)c.y));
      };
#pragma unroll
      for (int i = 0; i < SPT; i++) {
        if (fs[i] < 0) continue;
        a.u[off + fs[i]] = apply_Mdag(g, decode(g, fs[i]), a.sm[fs[i]], ldx);
        S::st(pnew[k], fs[i], pv[k][i][0], a.nt);
      }
      if (lead) put_mstate(sout + rr[k], 1, sk[k]->iters, PH_UPDATE, nrel[k], 0, 0, sk[k]->rho, rmax[k], *sk[k]);
    } else if (mode[k] == MM_NORMAL) {
      const float fb = (float)beta[k];
#pragma unroll
      for (int i = 0; i < SPT; i++) {
        if (fs[i] < 0) continue;
        const SiteCode st = kc[i];
        float2 p0v = cfma(fb, pv[k][i][0], rv[k][i][0]);
        float2 ppv[3], pmv[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          ppv[d] = cfma(fb, pv[k][i][1 + d], rv[k][i][1 + d]);
          pmv[d] = cfma(fb, pv[k][i][4 + d], rv[k][i][4 + d]);
        }
        S::st(pnew[k], fs[i], p0v, a.nt);
        float2 tt = apply_Mdag_fast_vals(st, smv[i], p0v, ppv, pmv);
        S::st((T*)a.t + off, fs[i], tt, a.nt);
        acc[k] += dnorm2(tt);
      }
    }
  }
  if (mode[0] != MM_NORMAL && mode[1] != MM_NORMAL) return;  // uniform
  const double2 tot = block_sum2(acc[0], acc[1], lds);
  const double totk[2] = {tot.x, tot.y};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (mode[k] != MM_NORMAL) continue;
    if (threadIdx.x == 0) pout[(size_t)rr[k] * a.nb + partial_slot(blockIdx.x, a.nb)] = totk[k];
    if (lead) put_mstate(sout + rr[k], 1, sk[k]->iters, PH_NORMAL, nrel[k], 0, 0, rhok[k], rmax[k], *sk[k]);
  }
}

template <typename S, int SPT, bool XCD>
__global__ __launch_bounds__(MIX_BS) void k_mix_Y2(MixFix a, int odd, const MixState* __restrict__ sin,
                                                   MixState* __restrict__ sout, const double* __restrict__ pin,
                                                   double* __restrict__ pout) {
  __shared__ double lds[2 * (MIX_BS / 64)];
  using T = typename S::T;
  const Geom g = a.g;
  const int r0 = 2 * blockIdx.y;
  const bool two = r0 + 1 < a.nrhs;
  const int rr[2] = {r0, two ? r0 + 1 : r0};
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const MixState s0 = sin[rr[0]];
  const MixState s1 = sin[rr[1]];
  const double sh0 = partial_share(pin + (size_t)rr[0] * a.nb, a.nb);
  const double sh1 = two ? partial_share(pin + (size_t)rr[1] * a.nb, a.nb) : 0.0;
  const MixState* sk[2] = {&s0, &s1};
  int mode[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (k == 1 && !two) {
      mode[1] = MM_OFF;
      continue;
    }
    if (!sk[k]->active) {
      if (lead)
        put_mstate(sout + rr[k], 0, sk[k]->iters, sk[k]->phase, sk[k]->nrel, sk[k]->converged, sk[k]->early,
                   sk[k]->rho, sk[k]->rmax, *sk[k]);
      mode[k] = MM_OFF;
    } else {
      mode[k] = sk[k]->phase == PH_UPDATE ? MM_UPDATE : MM_NORMAL;
    }
  }
  if (mode[0] == MM_OFF && mode[1] == MM_OFF) return;  // uniform
  // normal-step loads (both systems), issued before the reduction
  int fs[SPT];
  Site st[SPT];
  float2 w[2][SPT], pv[2][SPT], xv[2][SPT], rv[2][SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    fs[i] = mix_site<SPT, XCD>(a, i);
    if (fs[i] < 0) continue;
    st[i] = decode(g, fs[i]);
    NbrF n = neighbours_f(g, st[i]);
    const float smv = a.sm32[fs[i]];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (mode[k] != MM_NORMAL) continue;
      const size_t off = (size_t)rr[k] * g.V;
      const T* t = (const T*)a.t + off;
      float2 vpp[3], vmm[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        vpp[d] = S::ld(t, n.pp[d]);
        vmm[d] = S::ld(t, n.mm[d]);
      }
      w[k][i] = apply_M_vals(g, st[i], smv, S::ld(t, fs[i]), vpp, vmm);
      pv[k][i] = S::ld((const T*)(odd ? a.p0 : a.p1) + off, fs[i]);
      xv[k][i] = a.xs[off + fs[i]];
      rv[k][i] = S::ld((const T*)a.r + off, fs[i]);
    }
  }
  const double2 gam = block_sum2(sh0, sh1, lds);
  const double gamk[2] = {gam.x, gam.y};
  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const size_t off = (size_t)rr[k] * g.V;
    if (mode[k] == MM_UPDATE) {
      // reliable update, part 2 (as k_mix_Y)
      const double2* u = a.u + off;
#pragma unroll
      for (int i = 0; i < SPT; i++) {
        if (fs[i] < 0) continue;
        const int f = fs[i];
        float2 c = a.xs[off + f];
        a.x[off + f] = cadd(a.x[off + f], make_double2((double)c.x, (double)c.y));
        a.xs[off + f] = make_float2(0.f, 0.f);
        double2 rn = apply_sub_M(g, st[i], a.sm[f], a.b[off + f], [&](int q) { return u[q]; });
        S::st((T*)a.r + off, f, make_float2((float)rn.x, (float)rn.y), a.nt);
        acc[k] += cnorm2(rn);
      }
    } else if (mode[k] == MM_NORMAL) {
      const float al = (float)(sk[k]->rho / gamk[k]);
#pragma unroll
      for (int i = 0; i < SPT; i++) {
        if (fs[i] < 0) continue;
        st_mix(&a.xs[off + fs[i]], cfma(al, pv[k][i], xv[k][i]), a.nt);
        float2 rn = cfma(-al, w[k][i], rv[k][i]);
        S::st((T*)a.r + off, fs[i], rn, a.nt);
        acc[k] += dnorm2(rn);
      }
    }
  }
  const double2 tot = block_sum2(acc[0], acc[1], lds);
  const double totk[2] = {tot.x, tot.y};
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (mode[k] == MM_OFF) continue;
    if (threadIdx.x == 0) pout[(size_t)rr[k] * a.nb + partial_slot(blockIdx.x, a.nb)] = totk[k];
    if (lead) {
      if (mode[k] == MM_UPDATE)
        put_mstate(sout + rr[k], 1, sk[k]->iters, PH_UPDATED, sk[k]->nrel, 0, 0, sk[k]->rho, sk[k]->rmax, *sk[k]);
      else
        put_mstate(sout + rr[k], 1, sk[k]->iters + 1, PH_NORMAL, sk[k]->nrel, 0, 0, sk[k]->rho, sk[k]->rmax, *sk[k]);
    }
  }
}

// ---- host side ---------------------------------------------------------------
struct MixMachine {
  GraphEngine eng;
  double* sm = nullptr;
  float* sm32 = nullptr;
  double2* b = nullptr;
  double2* x = nullptr;
  double2* u = nullptr;
  float2* xs = nullptr;
  void* r = nullptr;
  void* p0 = nullptr;
  void* p1 = nullptr;
  void* t = nullptr;
  CGDevResult* res = nullptr;
  long cap = 0;  // complex elements per vector
  int autotune = 0;  // 1: choose shapes by timing (GNX_AUTOTUNE=1, gnx_set_autotune)
  MixMachine() {
    const char* at = getenv("GNX_AUTOTUNE");
    if (at && atoi(at) != 0) autotune = 1;
  }
  void free_vectors() {
    void* bufs[] = {sm, sm32, b, x, u, xs, r, p0, p1, t, res};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
    sm = nullptr;
    sm32 = nullptr;
    b = x = u = nullptr;
    xs = nullptr;
    r = p0 = p1 = t = nullptr;
    res = nullptr;
    cap = 0;
  }
  ~MixMachine() { free_vectors(); }
};

void cgmix_release(void* m) { delete (MixMachine*)m; }

static hipError_t mix_ensure(MixMachine* m, const Geom& g, int nrhs) {
  const int nb = (g.V + MIX_BS - 1) / MIX_BS;
  const int cap_rhs = nrhs < 2 ? 2 : nrhs;
  hipError_t e = m->eng.ensure(cap_rhs, 1, sizeof(MixState), 2 * (size_t)cap_rhs * nb);
  if (e != hipSuccess) return e;
  const long need = (long)g.V * cap_rhs;
  if (m->cap >= need && m->res) return hipSuccess;
  m->free_vectors();
  m->eng.invalidate_graphs();
  e = hipMalloc(&m->sm, sizeof(double) * g.V);
  if (e == hipSuccess) e = hipMalloc(&m->sm32, sizeof(float) * g.V);
  if (e == hipSuccess) e = hipMalloc(&m->b, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->x, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->u, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->xs, sizeof(float2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->r, sizeof(float2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->p0, sizeof(float2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->p1, sizeof(float2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->t, sizeof(float2) * need);
  if (e == hipSuccess) e = hipMalloc(&m->res, sizeof(CGDevResult) * cap_rhs);
  if (e != hipSuccess) {
    m->free_vectors();
    return e;
  }
  m->cap = need;
  return hipSuccess;
}

// launch shapes: sites per thread (the kernels of one solve share it)
struct MixShape {
  int spt;
  bool xcd;
  bool pair;   // two systems per block (k_mix_X2 / k_mix_Y2)
  bool early;  // loads-first kernels (k_mix_Xe / k_mix_Ye)
  const void* X;   // X kernel of odd slots
  const void* Xf;  // X kernel of even slots (differs only in the loads-first form: the xs flush)
  const void* Y;
  const void* start_imposed;
  const void* null_p1;
  const void* null_p2;
  const void* null_p3;
  bool quad = false;  // four consecutive sites per thread (k_mix_Xq / k_mix_Yq)
};

#define GNX_MIX_SHAPE(S, N, X)                                                                     \
  {N, X, false, false, (const void*)k_mix_X<S, N, X>, (const void*)k_mix_X<S, N, X>,             \
   (const void*)k_mix_Y<S, N, X>,                                                                 \
   (const void*)k_mix_start_imposed<S, N, X>, (const void*)k_mix_null_p1<N, X>,                   \
   (const void*)k_mix_null_p2<N, X>, (const void*)k_mix_null_p3<S, N, X>}
#define GNX_MIX_PAIR(S, N, X)                                                                      \
  {N, X, true, false, (const void*)k_mix_X2<S, N, X>, (const void*)k_mix_X2<S, N, X>,            \
   (const void*)k_mix_Y2<S, N, X>,                                                                \
   (const void*)k_mix_start_imposed<S, N, X>, (const void*)k_mix_null_p1<N, X>,                   \
   (const void*)k_mix_null_p2<N, X>, (const void*)k_mix_null_p3<S, N, X>}
#define GNX_MIX_EARLY(S, N, X)                                                                     \
  {N, X, false, true, (const void*)k_mix_Xe<S, N, X, false>, (const void*)k_mix_Xe<S, N, X, true>, \
   (const void*)k_mix_Ye<S, N, X>,                                                                \
   (const void*)k_mix_start_imposed<S, N, X>, (const void*)k_mix_null_p1<N, X>,                   \
   (const void*)k_mix_null_p2<N, X>, (const void*)k_mix_null_p3<S, N, X>}
// the group kernels exist for the 4-byte storage formats only
template <typename S, bool FLUSH>
static const void* mix_xq() {
  if constexpr (sizeof(typename S::T) == 4) return (const void*)k_mix_Xq<S, FLUSH>;
  return nullptr;
}
template <typename S>
static const void* mix_yq() {
  if constexpr (sizeof(typename S::T) == 4) return (const void*)k_mix_Yq<S>;
  return nullptr;
}
#define GNX_MIX_QUAD(S)                                                                            \
  {4, true, false, true, mix_xq<S, false>(), mix_xq<S, true>(), mix_yq<S>(),                      \
   (const void*)k_mix_start_imposed<S, 4, true>, (const void*)k_mix_null_p1<4, true>,             \
   (const void*)k_mix_null_p2<4, true>, (const void*)k_mix_null_p3<S, 4, true>, true}
// shapes 0..5: one system per block; 6..9: two (batches of two or more);
// 10..15: one system per block, loads-first kernels; 16: loads-first on
// groups of four consecutive sites (4-byte storage, LT a multiple of 8; the
// start kernels and partial layout of shape 15)
template <typename S>
static const MixShape* mix_shapes() {
  static const MixShape t[] = {GNX_MIX_SHAPE(S, 1, false), GNX_MIX_SHAPE(S, 1, true), GNX_MIX_SHAPE(S, 2, false),
                               GNX_MIX_SHAPE(S, 2, true),  GNX_MIX_SHAPE(S, 4, false), GNX_MIX_SHAPE(S, 4, true),
                               GNX_MIX_PAIR(S, 1, false),  GNX_MIX_PAIR(S, 1, true),   GNX_MIX_PAIR(S, 2, false),
                               GNX_MIX_PAIR(S, 2, true),   GNX_MIX_EARLY(S, 1, false), GNX_MIX_EARLY(S, 1, true),
                               GNX_MIX_EARLY(S, 2, false), GNX_MIX_EARLY(S, 2, true),  GNX_MIX_EARLY(S, 4, false),
                               GNX_MIX_EARLY(S, 4, true),  GNX_MIX_QUAD(S)};
  return t;
}
#undef GNX_MIX_SHAPE
#undef GNX_MIX_PAIR
#undef GNX_MIX_EARLY
#undef GNX_MIX_QUAD
constexpr int kMixShapes = 17;

// whether a shape runs this storage format, lattice and batch
template <typename S>
static bool mix_fits(const Geom& g, int nrhs, int shape) {
  const MixShape& sh = mix_shapes<S>()[shape];
  if (sh.pair && nrhs < 2) return false;
  if (S::kScaled && !sh.early) return false;  // the scale lives in the loads-first form
  if (sh.quad && (sh.X == nullptr || g.LT % 8 != 0)) return false;
  return true;
}

static MixFix mix_fix(MixMachine* m, const Geom& g, int nrhs, int spt, int store_bytes) {
  MixFix a;
  a.g = g;
  a.sm = m->sm;
  a.sm32 = m->sm32;
  a.b = m->b;
  a.x = m->x;
  a.u = m->u;
  a.xs = m->xs;
  a.r = m->r;
  a.p0 = m->p0;
  a.p1 = m->p1;
  a.t = m->t;
  a.part = m->eng.dev_part();
  a.state = (MixState*)m->eng.dev_state();
  a.res = m->res;
  a.nrhs = nrhs;
  a.nb = (g.V + spt * MIX_BS - 1) / (spt * MIX_BS);
  a.ptotal = (long)nrhs * a.nb;
  a.sys = nullptr;
  a.caps = nullptr;
  // r, t, two p in the storage format and xs in FP32, per site and system
  const double ws = (double)g.V * nrhs * (4.0 * store_bytes + 8.0);
  a.nt = ws > 240.0 * (1 << 20) ? 1 : 0;
  return a;
}

template <typename S>
static hipError_t mix_graph(MixMachine* m, const Geom& g, int nrhs, int shape, int tag, EngGraph* exec) {
  const MixShape& sh = mix_shapes<S>()[shape];
  MixFix a = mix_fix(m, g, nrhs, sh.spt, (int)sizeof(typename S::T));
  dim3 grid(a.nb, sh.pair ? (nrhs + 1) / 2 : nrhs), blk(MIX_BS);
  const MixState* s0 = a.state;
  const MixState* s1 = a.state + nrhs;
  MixState* w0 = a.state;
  MixState* w1 = a.state + nrhs;
  const double* q0 = mslot(a, 0, 0);
  const double* q1 = mslot(a, 1, 0);
  double* o0 = mslot(a, 0, 0);
  double* o1 = mslot(a, 1, 0);
  auto body = [&](hipStream_t cs, int iters) {
    static_assert(ENG_CHUNK % 2 == 0 && ENG_TAIL % 2 == 0, "the direction buffers alternate by slot parity");
    for (int i = 0; i < iters; i++) {
      int sl = i & 1;  // slot parity
      // X: state[0] -> state[1], reduces slot 1, writes slot 0; Y the reverse
      void* argsX[] = {(void*)&a, (void*)&sl, (void*)&s0, (void*)&w1, (void*)&q1, (void*)&o0};
      void* argsY[] = {(void*)&a, (void*)&sl, (void*)&s1, (void*)&w0, (void*)&q0, (void*)&o1};
      (void)hipLaunchKernel((i & 1) ? sh.X : sh.Xf, grid, blk, argsX, 0, cs);
      (void)hipLaunchKernel(sh.Y, grid, blk, argsY, 0, cs);
    }
  };
  return m->eng.graph(tag + 10 * shape, nrhs, body, exec);
}

template <typename S>
static hipError_t mix_begin(MixMachine* m, const Geom& g, const double* sm, const double2* b, const double2* xin,
                            double tol, double delta, long max_iter, int start, int nrhs, int shape,
                            hipStream_t st) {
  const MixShape& sh = mix_shapes<S>()[shape];
  MixFix a = mix_fix(m, g, nrhs, sh.spt, (int)sizeof(typename S::T));
  dim3 grid(a.nb, nrhs), blk(MIX_BS);
  hipError_t e;
  if (start == GNX_START_NULL) {
    void* a1[] = {(void*)&a, (void*)&sm, (void*)&b, (void*)&xin};
    e = hipLaunchKernel(sh.null_p1, grid, blk, a1, 0, st);
    void* a2[] = {(void*)&a};
    if (e == hipSuccess) e = hipLaunchKernel(sh.null_p2, grid, blk, a2, 0, st);
    void* a3[] = {(void*)&a, (void*)&tol, (void*)&delta, (void*)&max_iter};
    if (e == hipSuccess) e = hipLaunchKernel(sh.null_p3, grid, blk, a3, 0, st);
  } else {
    void* a0[] = {(void*)&a, (void*)&sm, (void*)&b, (void*)&tol, (void*)&delta, (void*)&max_iter};
    e = hipLaunchKernel(sh.start_imposed, grid, blk, a0, 0, st);
  }
  return e;
}

// Launch shape of a mixed solve.  By default a fixed rule (the same
// iterates in every process): the XCD slab mapping with the most sites per
// thread (4, 2, 1) that still gives the batch 512 blocks of at least 8 per
// system; loads-first kernels for 2-byte storage (FP16 requires them: the
// per-system scale lives there) and for single systems, the interleaved
// form for FP32 batches (bandwidth-bound, where the loads-first form loses
// occupancy) -- the tuner's picks at 64^3.  GNX_MIX_SHAPE=<0..16> pins one
// (ignored where it does not fit); autotune times two chunks of every shape
// that fits on the caller's own system and keeps the fastest per (batch
// size, storage format).
template <typename S>
static int mix_rule(const Geom& g, int nrhs) {
  int k = 0;
  for (int kk = 2; kk >= 1 && k == 0; kk--) {
    const long nb = (g.V + (long)MIX_BS * (1 << kk) - 1) / ((long)MIX_BS * (1 << kk));
    if (nb >= NUM_XCD && nb * nrhs >= 512) k = kk;
  }
  const bool lf = sizeof(typename S::T) == 4 || S::kScaled || nrhs == 1;
  return (lf ? 10 : 0) + 2 * k + 1;
}

template <typename S>
static hipError_t mix_choose(MixMachine* m, const Geom& g, const double* sm, const double2* b, double delta,
                             int nrhs, int tag, hipStream_t st, int* best) {
  GraphEngine& E = m->eng;
  const int key = nrhs + 1000 * tag;
  const int cached = E.shapes.find(key);
  if (cached >= 0) {
    *best = cached;
    return hipSuccess;
  }
  int best_shape = mix_rule<S>(g, nrhs);
  hipError_t e = hipSuccess;
  const char* env = getenv("GNX_MIX_SHAPE");
  if (env && atoi(env) >= 0 && atoi(env) < kMixShapes && mix_fits<S>(g, nrhs, atoi(env))) {
    best_shape = atoi(env);
  } else if (m->autotune) {
    E.tune_runs++;
    float best_ms = 1e30f;
    for (int sh = 0; sh < kMixShapes && e == hipSuccess; sh++) {
      if (!mix_fits<S>(g, nrhs, sh)) continue;
      EngGraph exec;
      e = mix_graph<S>(m, g, nrhs, sh, tag, &exec);
      float ms_min = 1e30f;
      for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
        e = mix_begin<S>(m, g, sm, b, nullptr, 0.0, delta, 2 * ENG_CHUNK, GNX_START_IMPOSED, nrhs, sh, st);
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

template <typename S>
static hipError_t mix_run_t(MixMachine* m, const Geom& g, const double* sm, const double2* b, double2* x,
                            double tol, double delta, long max_iter, int start, int nrhs, hipStream_t st,
                            CGDevResult* host_res, int tag) {
  hipError_t e = mix_ensure(m, g, nrhs);
  if (e != hipSuccess) return e;
  int shape = 0;
  e = mix_choose<S>(m, g, sm, b, delta, nrhs, tag, st, &shape);
  if (e != hipSuccess) return e;
  EngGraph exec;
  e = mix_graph<S>(m, g, nrhs, shape, tag, &exec);
  if (e != hipSuccess) return e;
  e = mix_begin<S>(m, g, sm, b, x, tol, delta, max_iter, start, nrhs, shape, st);
  if (e != hipSuccess) return e;
  // at most max_iter normal slot pairs plus one update pair after each
  const FinishFn out = [&](hipStream_t s) {
    hipError_t f = hipMemcpyAsync(x, m->x, sizeof(double2) * (size_t)g.V * nrhs, hipMemcpyDeviceToDevice, s);
    if (f == hipSuccess && host_res)
      f = hipMemcpyAsync(host_res, m->res, sizeof(CGDevResult) * nrhs, hipMemcpyDeviceToHost, s);
    return f;
  };
  return m->eng.run(exec, nrhs, 2 * max_iter + 4, st, nullptr, eng_key(tag + 10 * shape, nrhs, tol, start),
                    mix_slots_of, &out);
}

hipError_t cgmix_run(void** machine_slot, const Geom& g, const double* sm, const double2* b, double2* x, double tol,
                     double delta, long max_iter, int start, int prec, int nrhs, hipStream_t st,
                     CGDevResult* host_res) {
  if (!*machine_slot) *machine_slot = new MixMachine();
  MixMachine* m = (MixMachine*)*machine_slot;
  if (prec == GNX_PREC_F32)
    return mix_run_t<StoreF32>(m, g, sm, b, x, tol, delta, max_iter, start, nrhs, st, host_res, 100);
  if (prec == GNX_PREC_F16)
    return mix_run_t<StoreF16>(m, g, sm, b, x, tol, delta, max_iter, start, nrhs, st, host_res, 102);
  return mix_run_t<StoreBF16>(m, g, sm, b, x, tol, delta, max_iter, start, nrhs, st, host_res, 101);
}

int cgmix_shape_of(void* machine, int nrhs, int prec, int* spt, int* xcd, int* pair, int* early) {
  MixMachine* m = (MixMachine*)machine;
  if (!m) return -1;
  const int key = nrhs + 1000 * (prec == GNX_PREC_F32 ? 100 : prec == GNX_PREC_F16 ? 102 : 101);
  const int t = m->eng.shapes.peek(key);
  if (t < 0) return -1;
  const MixShape& sh = mix_shapes<StoreF32>()[t];
  *spt = sh.spt;
  *xcd = (sh.xcd ? 1 : 0) | (sh.quad ? 2 : 0);
  *pair = sh.pair ? 1 : 0;
  *early = sh.early ? 1 : 0;
  return t;
}

void cgmix_set_autotune(void** machine_slot, int on) {
  if (!*machine_slot) *machine_slot = new MixMachine();
  ((MixMachine*)*machine_slot)->autotune = on;
}

long cgmix_tune_runs(void* machine) { return machine ? ((MixMachine*)machine)->eng.tune_runs : 0; }
void cgmix_replay_stats(void* machine, long* slots, long* polls) {
  *slots = machine ? ((MixMachine*)machine)->eng.launched_slots : 0;
  *polls = machine ? ((MixMachine*)machine)->eng.polls : 0;
}

// ---- uneven batches ------------------------------------------------------------------
// The loads-first kernels over the systems of an uneven solver (gnx_cgu.hip),
// each on its own lattice: one site per thread in the slab mapping, the
// partition and order of sums of the single-lattice shape 11
// (GNX_MIX_SHAPE=11), so a system's solution is bit-identical to its own
// gnx_cg solve in that shape with the same reliable-update factor.  The
// systems' sigma masses, right-hand sides and (null start) x_in are gathered
// into the machine's concatenated buffers by one kernel, the solutions
// scattered back by another; the descriptors and partial offsets are the
// FP64 uneven solver's.
struct MixMulti {
  GraphEngine eng;
  double* sm = nullptr;
  float* sm32 = nullptr;
  double2 *b = nullptr, *x = nullptr, *u = nullptr;
  float2* xs = nullptr;
  void *r = nullptr, *p0 = nullptr, *p1 = nullptr, *t = nullptr;
  CGDevResult* res = nullptr;
  long* caps = nullptr;      // [n] iteration cap per system
  long* hcaps = nullptr;     // pinned
  CGUPtrs* dptr = nullptr;   // [n] the caller's vectors of the current solve
  CGUPtrs* hptr = nullptr;   // pinned
  ~MixMulti() {
    void* dev[] = {sm, sm32, b, x, u, xs, r, p0, p1, t, res, caps, dptr};
    for (void* p : dev)
      if (p) (void)hipFree(p);
    if (hcaps) (void)hipHostFree(hcaps);
    if (hptr) (void)hipHostFree(hptr);
  }
};

void cgmix_multi_release(void* m) { delete (MixMulti*)m; }

__global__ __launch_bounds__(MIX_BS) void k_mixu_gather(MixFix a, const CGUPtrs* __restrict__ ptr, int with_x) {
  const CGUSys d = a.sys[blockIdx.y];
  const int f = (int)blockIdx.x * MIX_BS + (int)threadIdx.x;
  if (f >= d.g.V) return;
  const CGUPtrs p = ptr[blockIdx.y];
  a.sm[d.soff + f] = p.sm[f];
  a.b[d.voff + f] = p.b[f];
  if (with_x) a.x[d.voff + f] = p.x[f];
}

__global__ __launch_bounds__(MIX_BS) void k_mixu_scatter(MixFix a, const CGUPtrs* __restrict__ ptr) {
  const CGUSys d = a.sys[blockIdx.y];
  const int f = (int)blockIdx.x * MIX_BS + (int)threadIdx.x;
  if (f >= d.g.V) return;
  ptr[blockIdx.y].x[f] = a.x[d.voff + f];
}

// buffers for the solver's systems (sized once: an uneven solver's systems never change)
static hipError_t mixu_ensure(MixMulti* mm, const MultiView& mv) {
  if (mm->res) return hipSuccess;
  const size_t vt = (size_t)mv.vtotal, n = (size_t)mv.n;
  hipError_t e = mm->eng.ensure(mv.n, 1, sizeof(MixState), 2 * (size_t)mv.ptotal);
  if (e == hipSuccess) e = hipMalloc(&mm->sm, sizeof(double) * (size_t)mv.stotal);
  if (e == hipSuccess) e = hipMalloc(&mm->sm32, sizeof(float) * (size_t)mv.stotal);
  if (e == hipSuccess) e = hipMalloc(&mm->b, sizeof(double2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->x, sizeof(double2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->u, sizeof(double2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->xs, sizeof(float2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->r, sizeof(float2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->p0, sizeof(float2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->p1, sizeof(float2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->t, sizeof(float2) * vt);
  if (e == hipSuccess) e = hipMalloc(&mm->caps, sizeof(long) * n);
  if (e == hipSuccess) e = hipHostMalloc(&mm->hcaps, sizeof(long) * n, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&mm->dptr, sizeof(CGUPtrs) * n);
  if (e == hipSuccess) e = hipHostMalloc(&mm->hptr, sizeof(CGUPtrs) * n, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&mm->res, sizeof(CGDevResult) * n);
  return e;
}

static MixFix mixu_fix(MixMulti* mm, const MultiView& mv, int store_bytes) {
  MixFix a;
  a.g = mv.hsys[0].g;  // unused: every kernel of the uneven form reads its system's descriptor
  a.sm = mm->sm;
  a.sm32 = mm->sm32;
  a.b = mm->b;
  a.x = mm->x;
  a.u = mm->u;
  a.xs = mm->xs;
  a.r = mm->r;
  a.p0 = mm->p0;
  a.p1 = mm->p1;
  a.t = mm->t;
  a.part = mm->eng.dev_part();
  a.state = (MixState*)mm->eng.dev_state();
  a.res = mm->res;
  a.nrhs = mv.n;
  a.nb = mv.nbmax;
  a.ptotal = mv.ptotal;
  a.sys = mv.dsys;
  a.caps = mm->caps;
  const double ws = (double)mv.vtotal * (4.0 * store_bytes + 8.0);
  a.nt = ws > 240.0 * (1 << 20) ? 1 : 0;
  return a;
}

template <typename S>
static hipError_t mixu_run_t(MixMulti* mm, const MultiView& mv, const double* const* sm, const double* const* b,
                             double* const* x, double tol, double delta, long max_iter, int start,
                             CGDevResult* host_res, int tag) {
  hipError_t e = mixu_ensure(mm, mv);
  if (e != hipSuccess) return e;
  const hipStream_t st = mv.stream;
  const int n = mv.n;
  // the pinned tables are rewritten here: no copy of an earlier call may still read them
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  long cap_max = 0;
  for (int k = 0; k < n; k++) {
    // the reference's cap V+1 (macro.h:134) of each system when max_iter <= 0
    mm->hcaps[k] = max_iter > 0 ? max_iter : (long)mv.hsys[k].g.V + 1;
    cap_max = mm->hcaps[k] > cap_max ? mm->hcaps[k] : cap_max;
    mm->hptr[k] = CGUPtrs{sm[k], (const double2*)b[k], (double2*)x[k], nullptr};
  }
  e = hipMemcpyAsync(mm->caps, mm->hcaps, sizeof(long) * n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(mm->dptr, mm->hptr, sizeof(CGUPtrs) * n, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  MixFix a = mixu_fix(mm, mv, (int)sizeof(typename S::T));
  const dim3 grid((mv.nbmax + NUM_XCD - 1) / NUM_XCD * NUM_XCD, n), blk(MIX_BS);
  const CGUPtrs* dp = mm->dptr;
  int with_x = start == GNX_START_NULL ? 1 : 0;
  void* ag[] = {(void*)&a, (void*)&dp, (void*)&with_x};
  e = hipLaunchKernel((const void*)k_mixu_gather, grid, blk, ag, 0, st);
  if (e != hipSuccess) return e;
  // start kernels on the gathered buffers (the cap comes from caps)
  const double* smv = mm->sm;
  const double2* bv = mm->b;
  const double2* xv = mm->x;
  long unused_cap = 0;
  if (start == GNX_START_NULL) {
    void* a1[] = {(void*)&a, (void*)&smv, (void*)&bv, (void*)&xv};
    e = hipLaunchKernel((const void*)k_mix_null_p1<1, true, MixUneven>, grid, blk, a1, 0, st);
    void* a2[] = {(void*)&a};
    if (e == hipSuccess) e = hipLaunchKernel((const void*)k_mix_null_p2<1, true, MixUneven>, grid, blk, a2, 0, st);
    void* a3[] = {(void*)&a, (void*)&tol, (void*)&delta, (void*)&unused_cap};
    if (e == hipSuccess)
      e = hipLaunchKernel((const void*)k_mix_null_p3<S, 1, true, MixUneven>, grid, blk, a3, 0, st);
  } else {
    void* a0[] = {(void*)&a, (void*)&smv, (void*)&bv, (void*)&tol, (void*)&delta, (void*)&unused_cap};
    e = hipLaunchKernel((const void*)k_mix_start_imposed<S, 1, true, MixUneven>, grid, blk, a0, 0, st);
  }
  if (e != hipSuccess) return e;
  // the chunk and tail graphs: slot pairs over all systems, as mix_graph
  const MixState* s0 = a.state;
  const MixState* s1 = a.state + n;
  MixState* w0 = a.state;
  MixState* w1 = a.state + n;
  const double* q0 = a.part;
  const double* q1 = a.part + a.ptotal;
  double* o0 = a.part;
  double* o1 = a.part + a.ptotal;
  auto body = [&](hipStream_t cs, int iters) {
    for (int i = 0; i < iters; i++) {
      int sl = i & 1;
      void* argsX[] = {(void*)&a, (void*)&sl, (void*)&s0, (void*)&w1, (void*)&q1, (void*)&o0};
      void* argsY[] = {(void*)&a, (void*)&sl, (void*)&s1, (void*)&w0, (void*)&q0, (void*)&o1};
      (void)hipLaunchKernel((i & 1) ? (const void*)k_mix_Xe<S, 1, true, false, MixUneven>
                                    : (const void*)k_mix_Xe<S, 1, true, true, MixUneven>,
                            grid, blk, argsX, 0, cs);
      (void)hipLaunchKernel((const void*)k_mix_Ye<S, 1, true, MixUneven>, grid, blk, argsY, 0, cs);
    }
  };
  EngGraph exec;
  e = mm->eng.graph(tag, n, body, &exec);
  if (e != hipSuccess) return e;
  // at most max_iter normal slot pairs plus one update pair after each
  // the solutions scattered to the callers' vectors and the results copied
  // after each burst (the gathered x_in were read by the start kernels)
  const FinishFn out = [&](hipStream_t s) {
    void* as[] = {(void*)&a, (void*)&dp};
    hipError_t f = hipLaunchKernel((const void*)k_mixu_scatter, grid, blk, as, 0, s);
    if (f == hipSuccess) f = hipMemcpyAsync(host_res, mm->res, sizeof(CGDevResult) * n, hipMemcpyDeviceToHost, s);
    return f;
  };
  return mm->eng.run(exec, n, 2 * cap_max + 4, st, nullptr, eng_key(tag, n, tol, start), mix_slots_of, &out);
}

hipError_t cgmix_multi_run(const MultiView& mv, const double* const* sm, const double* const* b, double* const* x,
                           double tol, double delta, long max_iter, int start, int prec, CGDevResult* host_res) {
  if (!*mv.mixed) *mv.mixed = new MixMulti();
  MixMulti* mm = (MixMulti*)*mv.mixed;
  if (prec == GNX_PREC_F32)
    return mixu_run_t<StoreF32>(mm, mv, sm, b, x, tol, delta, max_iter, start, host_res, 100);
  if (prec == GNX_PREC_F16)
    return mixu_run_t<StoreF16>(mm, mv, sm, b, x, tol, delta, max_iter, start, host_res, 102);
  return mixu_run_t<StoreBF16>(mm, mv, sm, b, x, tol, delta, max_iter, start, host_res, 101);
}

}  // namespace gnx
