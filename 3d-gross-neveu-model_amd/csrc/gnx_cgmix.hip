// gnx_cgmix.hip -- mixed-precision CG on A = M M^dag with FP64 reliable
// updates (Sleijpen & van der Vorst's reliable-update scheme as used for
// lattice Dirac solvers).
//
// Inner iterations keep r, p, t in reduced-precision storage (FP32, or BF16
// with FP32 arithmetic) and the solution correction xs in FP32; every global
// sum is accumulated in FP64.  When the iterated residual has dropped by a
// factor `delta` since the last update (or claims convergence), a reliable
// update folds xs into the FP64 solution and recomputes the true residual
// r = b - M M^dag x in FP64; convergence is declared only on that true
// residual (||r|| < tol, absolute, as conjugategradient.c:87-89).  The search
// direction is kept across updates (no restart), so the Krylov space is not
// lost.  FP64 is used only for x, the true residual, the sums and the
// scalars.
//
// Kernel slots, per system (same graph engine as the FP64 solver):
//   X: reduce rho; if a reliable update just finished, test the true
//      residual; otherwise test ||r_it|| < delta * rmax -> start an update
//      (X does u = M^dag (x + xs) in FP64); else p = r + beta p, t = M^dag p.
//   Y: finish an update (x += xs, xs = 0, r64 = b - M u, r = (S)r64,
//      partial |r64|^2) or do the normal step (xs += alpha p, r -= alpha M t,
//      partial |r|^2).
#include <stdlib.h>
#include <string.h>

#include "gnx_engine.h"
#include "gnx_internal.h"

namespace gnx {

// ---- storage formats --------------------------------------------------------
struct StoreF32 {
  using T = float2;
  static constexpr int bytes = 8;
  GNX_DEVONLY static float2 ld(const T* p, int i) { return p[i]; }
  GNX_DEVONLY static void st(T* p, int i, float2 v) { p[i] = v; }
};

// bfloat16 pair (re in the low half), round-to-nearest-even on store
struct StoreBF16 {
  using T = unsigned int;
  static constexpr int bytes = 4;
  GNX_DEVONLY static float up(unsigned int h) { return __uint_as_float(h << 16); }
  GNX_DEVONLY static unsigned int down(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
  }
  GNX_DEVONLY static float2 ld(const T* p, int i) {
    unsigned int v = p[i];
    return make_float2(up(v & 0xFFFFu), up(v >> 16));
  }
  GNX_DEVONLY static void st(T* p, int i, float2 v) { p[i] = down(v.x) | (down(v.y) << 16); }
};

constexpr int MIX_BS = ENG_BS;

enum MixPhase { PH_NORMAL = 0, PH_UPDATE = 1, PH_UPDATED = 2 };

struct MixState {
  int active;
  int iters;      // CG iterations (normal X/Y pairs)
  int phase;
  int nrel;       // reliable updates done
  int converged;
  int early;
  double rho;     // rho of the current direction (alpha = rho / gamma)
  double rmax;    // ||r|| at the last reliable update
};

struct MixArgs {
  Geom g;
  const double* sm;
  float* sm32;
  const double2* b;
  double2* x;       // FP64 solution (caller)
  double2* r64;     // FP64 true residual
  double2* u64;     // FP64 M^dag (x + xs)
  float2* xs;       // FP32 solution correction
  void* r;          // storage-format residual
  void* p0;
  void* p1;
  void* t;
  double* part;     // [2 slots][nrhs][nb]
  MixState* state;  // [2][nrhs]
  CGDevResult* res;
  double tol;
  double delta;
  long max_iter;
  int nrhs;
  int nb;
  int start;
};

GNX_DEVONLY double* mslot(const MixArgs& a, int s, int rhs) { return a.part + ((size_t)s * a.nrhs + rhs) * a.nb; }

GNX_DEVONLY double mreduce(const MixArgs& a, int s, int rhs, double* lds) {
  const double* src = mslot(a, s, rhs);
  double acc = 0.0;
  for (int i = threadIdx.x; i < a.nb; i += MIX_BS) acc += src[i];
  return block_sum(acc, lds);
}

GNX_DEVONLY double dnorm2(float2 v) { return (double)v.x * v.x + (double)v.y * v.y; }

// ---- start kernels ------------------------------------------------------------
// imposed start: x = 0, r64 = b, r = p0 = (S)b, xs = 0, sm32; partial |b|^2
template <typename S>
__global__ __launch_bounds__(MIX_BS) void k_mix_start_imposed(const MixArgs* __restrict__ ap) {
  const MixArgs& a = *ap;
  __shared__ double lds[MIX_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const int f = blockIdx.x * MIX_BS + threadIdx.x;
  double acc = 0.0;
  if (f < a.g.V) {
    double2 bb = a.b[off + f];
    a.x[off + f] = make_double2(0.0, 0.0);
    a.r64[off + f] = bb;
    float2 bs = make_float2((float)bb.x, (float)bb.y);
    S::st((typename S::T*)a.r + off, f, bs);
    S::st((typename S::T*)a.p0 + off, f, bs);
    a.xs[off + f] = make_float2(0.f, 0.f);
    if (rhs == 0) a.sm32[f] = (float)a.sm[f];
    acc = cnorm2(bb);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) mslot(a, 1, rhs)[blockIdx.x] = acc;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MixState s = {1, 0, PH_NORMAL, 0, 0, 0, 0.0, 0.0};
    a.state[rhs] = s;
  }
}

// null start (conjugategradient.c:18-33) pass 1: u64 = M^dag x_in
__global__ __launch_bounds__(MIX_BS) void k_mix_null_p1(const MixArgs* __restrict__ ap) {
  const MixArgs& a = *ap;
  const size_t off = (size_t)blockIdx.y * a.g.V;
  const int f = blockIdx.x * MIX_BS + threadIdx.x;
  if (f >= a.g.V) return;
  const double2* xin = a.x + off;
  a.u64[off + f] = apply_Mdag(a.g, decode(a.g, f), a.sm[f], [&](int q) { return xin[q]; });
}

// pass 2: r64 = b - M u64, partials |r64|^2 (slot 0), |b|^2 (slot 1)
__global__ __launch_bounds__(MIX_BS) void k_mix_null_p2(const MixArgs* __restrict__ ap) {
  const MixArgs& a = *ap;
  __shared__ double lds[MIX_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const int f = blockIdx.x * MIX_BS + threadIdx.x;
  double ar = 0.0, ab = 0.0;
  if (f < a.g.V) {
    const double2* u = a.u64 + off;
    double2 bb = a.b[off + f];
    double2 rr = apply_sub_M(a.g, decode(a.g, f), a.sm[f], bb, [&](int q) { return u[q]; });
    ar = cnorm2(rr);
    ab = cnorm2(bb);
  }
  ar = block_sum(ar, lds);
  ab = block_sum(ab, lds);
  if (threadIdx.x == 0) {
    mslot(a, 0, rhs)[blockIdx.x] = ar;
    mslot(a, 1, rhs)[blockIdx.x] = ab;
  }
}

// pass 3: early exit keeps x_in, otherwise as the imposed start (slot 1 kept)
template <typename S>
__global__ __launch_bounds__(MIX_BS) void k_mix_null_p3(const MixArgs* __restrict__ ap) {
  const MixArgs& a = *ap;
  __shared__ double lds[MIX_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const double nr = mreduce(a, 0, rhs, lds);
  const bool early = sqrt(nr) < a.tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MixState s = {early ? 0 : 1, 0, PH_NORMAL, 0, early ? 1 : 0, early ? 1 : 0, 0.0, 0.0};
    a.state[rhs] = s;
    if (early) {
      CGDevResult r = {1, 1, 0, 0, sqrt(nr), {0, 0, 0}};
      a.res[rhs] = r;
    }
  }
  const int f = blockIdx.x * MIX_BS + threadIdx.x;
  if (f < a.g.V && rhs == 0) a.sm32[f] = (float)a.sm[f];
  if (early || f >= a.g.V) return;
  double2 bb = a.b[off + f];
  a.x[off + f] = make_double2(0.0, 0.0);
  a.r64[off + f] = bb;
  float2 bs = make_float2((float)bb.x, (float)bb.y);
  S::st((typename S::T*)a.r + off, f, bs);
  S::st((typename S::T*)a.p0 + off, f, bs);
  a.xs[off + f] = make_float2(0.f, 0.f);
}

// ---- iteration kernels -----------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(MIX_BS) void k_mix_X(const MixArgs* __restrict__ ap) {
  const MixArgs& a = *ap;
  __shared__ double lds[MIX_BS / 64];
  const int rhs = blockIdx.y;
  const MixState s = a.state[rhs];
  MixState* out = a.state + a.nrhs + rhs;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!s.active) {
    if (lead) *out = s;
    return;
  }
  const Geom g = a.g;
  const size_t off = (size_t)rhs * g.V;
  const double rho = mreduce(a, 1, rhs, lds);
  const double nr = sqrt(rho);
  MixState n = s;
  double beta = 0.0;
  bool update = false;
  if (s.phase == PH_UPDATED) {
    // rho is the true FP64 residual norm^2 after a reliable update
    n.nrel = s.nrel + 1;
    if (nr < a.tol) {
      if (lead) {
        n.active = 0;
        n.converged = 1;
        n.phase = PH_NORMAL;
        *out = n;
        CGDevResult r = {1, 0, s.iters, n.nrel, nr, {0, 0, 0}};
        a.res[rhs] = r;
      }
      return;
    }
    if (s.iters >= a.max_iter) {  // cap reached: x already holds every correction
      if (lead) {
        n.active = 0;
        n.phase = PH_NORMAL;
        *out = n;
        CGDevResult r = {0, 0, s.iters, n.nrel, nr, {0, 0, 0}};
        a.res[rhs] = r;
      }
      return;
    }
    n.rmax = nr;
    beta = rho / s.rho;
  } else if (s.iters == 0) {
    if (a.start == GNX_START_IMPOSED && nr < a.tol) {  // conjugategradient.c:123-128
      if (lead) {
        n.active = 0;
        n.converged = 1;
        n.early = 1;
        *out = n;
        CGDevResult r = {1, 1, 0, 0, nr, {0, 0, 0}};
        a.res[rhs] = r;
      }
      return;
    }
    n.rmax = nr;
  } else {
    // fold the correction and test the true residual when the iterated one
    // has dropped by delta, claims convergence, or the iteration cap is hit
    update = (nr < a.delta * s.rmax) || (nr < a.tol) || (s.iters >= a.max_iter);
    beta = rho / s.rho;
  }
  const int f = blockIdx.x * MIX_BS + threadIdx.x;
  if (update) {
    // reliable update, part 1: u = M^dag (x + xs) in FP64 (x and xs are not
    // written in this kernel, so neighbours read consistent values)
    if (f < g.V) {
      const double2* x = a.x + off;
      const float2* xs = a.xs + off;
      auto ldx = [&](int q) {
        float2 c = xs[q];
        return cadd(x[q], make_double2((double)c.x, (double)c.y));
      };
      a.u64[off + f] = apply_Mdag(g, decode(g, f), a.sm[f], ldx);
    }
    if (lead) {
      n.phase = PH_UPDATE;  // keep n.rho: beta after the update is rho_true / rho_k
      *out = n;
    }
    return;
  }
  // normal step: p = r + beta p, t = M^dag p
  using T = typename S::T;
  const T* r = (const T*)a.r + off;
  const T* pold = (const T*)((s.iters & 1) ? a.p1 : a.p0) + off;
  T* pnew = (T*)((s.iters & 1) ? a.p0 : a.p1) + off;
  const float fb = (float)beta;
  double acc = 0.0;
  if (f < g.V) {
    auto ldp = [&](int q) { return cadd(S::ld(r, q), rmul(fb, S::ld(pold, q))); };
    S::st(pnew, f, ldp(f));
    float2 tt = apply_Mdag(g, decode(g, f), a.sm32[f], ldp);
    S::st((T*)a.t + off, f, tt);
    acc = dnorm2(tt);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) mslot(a, 0, rhs)[blockIdx.x] = acc;
  if (lead) {
    n.phase = PH_NORMAL;
    n.rho = rho;
    *out = n;
  }
}

template <typename S>
__global__ __launch_bounds__(MIX_BS) void k_mix_Y(const MixArgs* __restrict__ ap) {
  const MixArgs& a = *ap;
  __shared__ double lds[MIX_BS / 64];
  const int rhs = blockIdx.y;
  const MixState s = a.state[a.nrhs + rhs];
  MixState* out = a.state + rhs;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!s.active) {
    if (lead) *out = s;
    return;
  }
  const Geom g = a.g;
  const size_t off = (size_t)rhs * g.V;
  const int f = blockIdx.x * MIX_BS + threadIdx.x;
  using T = typename S::T;
  MixState n = s;
  double acc = 0.0;
  if (s.phase == PH_UPDATE) {
    // reliable update, part 2: x += xs, xs = 0, r64 = b - M u, r = (S) r64
    if (f < g.V) {
      float2 c = a.xs[off + f];
      a.x[off + f] = cadd(a.x[off + f], make_double2((double)c.x, (double)c.y));
      a.xs[off + f] = make_float2(0.f, 0.f);
      const double2* u = a.u64 + off;
      double2 rr = apply_sub_M(g, decode(g, f), a.sm[f], a.b[off + f], [&](int q) { return u[q]; });
      a.r64[off + f] = rr;
      S::st((T*)a.r + off, f, make_float2((float)rr.x, (float)rr.y));
      acc = cnorm2(rr);
    }
    n.phase = PH_UPDATED;
  } else {
    const double gamma = mreduce(a, 0, rhs, lds);
    const float al = (float)(s.rho / gamma);
    if (f < g.V) {
      const T* pnew = (const T*)((s.iters & 1) ? a.p0 : a.p1) + off;
      const T* t = (const T*)a.t + off;
      T* r = (T*)a.r + off;
      float2 w = apply_M(g, decode(g, f), a.sm32[f], [&](int q) { return S::ld(t, q); });
      float2 pv = S::ld(pnew, f);
      float2 xv = a.xs[off + f];
      a.xs[off + f] = make_float2(xv.x + al * pv.x, xv.y + al * pv.y);
      float2 rv = S::ld(r, f);
      float2 rn = make_float2(rv.x - al * w.x, rv.y - al * w.y);
      S::st(r, f, rn);
      acc = dnorm2(rn);
    }
    n.iters = s.iters + 1;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) mslot(a, 1, rhs)[blockIdx.x] = acc;
  if (lead) *out = n;
}

// ---- host side ---------------------------------------------------------------
struct MixWork {
  double2* r64 = nullptr;
  double2* u64 = nullptr;
  float2* xs = nullptr;
  float* sm32 = nullptr;
  void* r = nullptr;
  void* p0 = nullptr;
  void* p1 = nullptr;
  void* t = nullptr;
  CGDevResult* res = nullptr;
  long cap = 0;  // complex elements per vector
  int store_bytes = 0;
};

struct MixMachine {
  GraphEngine eng;
  MixWork w;
  ~MixMachine() {
    void* bufs[] = {w.r64, w.u64, w.xs, w.sm32, w.r, w.p0, w.p1, w.t, w.res};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
  }
};

void cgmix_release(void* m) { delete (MixMachine*)m; }

static hipError_t mix_work(MixMachine* m, long V, int nrhs, int sbytes) {
  MixWork& w = m->w;
  long need = V * nrhs;
  if (w.cap >= need && w.store_bytes >= sbytes) return hipSuccess;
  void* bufs[] = {w.r64, w.u64, w.xs, w.sm32, w.r, w.p0, w.p1, w.t, w.res};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  w = MixWork();
  hipError_t e = hipMalloc(&w.r64, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&w.u64, sizeof(double2) * need);
  if (e == hipSuccess) e = hipMalloc(&w.xs, sizeof(float2) * need);
  if (e == hipSuccess) e = hipMalloc(&w.sm32, sizeof(float) * V);
  if (e == hipSuccess) e = hipMalloc(&w.r, (size_t)8 * need);
  if (e == hipSuccess) e = hipMalloc(&w.p0, (size_t)8 * need);
  if (e == hipSuccess) e = hipMalloc(&w.p1, (size_t)8 * need);
  if (e == hipSuccess) e = hipMalloc(&w.t, (size_t)8 * need);
  if (e == hipSuccess) e = hipMalloc(&w.res, sizeof(CGDevResult) * nrhs);
  if (e != hipSuccess) return e;
  w.cap = need;
  w.store_bytes = 8;
  return hipSuccess;
}

template <typename S>
static hipError_t mix_run_t(MixMachine* m, const Geom& g, const double* sm, const double2* b, double2* x,
                            double tol, double delta, long max_iter, int start, int nrhs, hipStream_t st,
                            CGDevResult* host_res, int tag) {
  const int nb = (g.V + MIX_BS - 1) / MIX_BS;
  hipError_t e = m->eng.ensure(nrhs, sizeof(MixArgs), sizeof(MixState), 2 * (size_t)(nrhs < 2 ? 2 : nrhs) * nb);
  if (e != hipSuccess) return e;
  e = mix_work(m, g.V, nrhs, S::bytes);
  if (e != hipSuccess) return e;
  EngineKernels k = {(const void*)k_mix_X<S>, (const void*)k_mix_Y<S>, nb, tag};
  hipGraphExec_t exec;
  e = m->eng.graph(k, nrhs, &exec);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  MixArgs& a = *(MixArgs*)m->eng.host_args();
  MixWork& w = m->w;
  a.g = g;
  a.sm = sm;
  a.sm32 = w.sm32;
  a.b = b;
  a.x = x;
  a.r64 = w.r64;
  a.u64 = w.u64;
  a.xs = w.xs;
  a.r = w.r;
  a.p0 = w.p0;
  a.p1 = w.p1;
  a.t = w.t;
  a.part = m->eng.dev_part();
  a.state = (MixState*)m->eng.dev_state();
  a.res = w.res;
  a.tol = tol;
  a.delta = delta;
  a.max_iter = max_iter;
  a.nrhs = nrhs;
  a.nb = nb;
  a.start = start;
  e = m->eng.stage_args(st);
  if (e != hipSuccess) return e;
  const MixArgs* dp = (const MixArgs*)m->eng.dev_args();
  dim3 grid(nb, nrhs), blk(MIX_BS);
  if (start == GNX_START_NULL) {
    hipLaunchKernelGGL(k_mix_null_p1, grid, blk, 0, st, dp);
    hipLaunchKernelGGL(k_mix_null_p2, grid, blk, 0, st, dp);
    hipLaunchKernelGGL(k_mix_null_p3<S>, grid, blk, 0, st, dp);
  } else {
    hipLaunchKernelGGL(k_mix_start_imposed<S>, grid, blk, 0, st, dp);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // a system needs at most max_iter normal slots plus one per reliable update;
  // updates happen at most once per normal step
  e = m->eng.run(exec, nrhs, 2 * max_iter + 4, st, nullptr);
  if (e != hipSuccess) return e;
  if (host_res) {
    e = hipMemcpyAsync(host_res, w.res, sizeof(CGDevResult) * nrhs, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
  }
  return e;
}

hipError_t cgmix_run(void** machine_slot, const Geom& g, const double* sm, const double2* b, double2* x, double tol,
                     double delta, long max_iter, int start, int prec, int nrhs, hipStream_t st,
                     CGDevResult* host_res) {
  if (!*machine_slot) *machine_slot = new MixMachine();
  MixMachine* m = (MixMachine*)*machine_slot;
  if (prec == GNX_PREC_F32)
    return mix_run_t<StoreF32>(m, g, sm, b, x, tol, delta, max_iter, start, nrhs, st, host_res, 100);
  return mix_run_t<StoreBF16>(m, g, sm, b, x, tol, delta, max_iter, start, nrhs, st, host_res, 101);
}

}  // namespace gnx
