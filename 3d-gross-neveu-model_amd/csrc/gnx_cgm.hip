// gnx_cgm.hip -- FP64 conjugate gradient on A = M M^dag (conjugategradient.c:13-182)
// as a device-side state machine of two kernels per iteration, replayed from
// a captured hipGraph.
//
// Why kernels and not a persistent kernel: on gfx950 a grid-wide barrier
// inside a kernel costs 10-80 us (cooperative grid sync or an atomic counter,
// growing with the block count), a kernel node inside a hipGraph about
// 1.7 us.  Standard CG needs two global reductions per iteration, so each
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
// Every block reduces the previous kernel's per-block partials in the same
// fixed order, so all blocks hold bit-identical scalars.  Per-system state
// (rho, active, iteration count) is double-buffered: kernel j reads
// state[j & 1] and its block 0 writes state[(j+1) & 1]; no block reads what
// another block of the same kernel writes.  Systems of a batch (blockIdx.y)
// are independent: each keeps its alpha, beta and exit iteration, and a
// finished system's blocks return at once.  The host replays a fixed graph of
// CHUNK iterations until every system is inactive.
#include <stdlib.h>
#include <string.h>

#include "gnx_engine.h"
#include "gnx_internal.h"

namespace gnx {

constexpr int CGM_BS = ENG_BS;

// per-system state; `active` and `iters` lead, as the replay engine polls them
struct CGMState {
  int active;
  int iters;
  int converged;
  int early;
  double rho;
};

struct CGMArgs {
  Geom g;
  const double* sm;
  const double2* b;
  double2* x;
  double2* r;
  double2* p0;
  double2* p1;
  double2* t;
  double* part;        // [2 slots][nrhs][nb]
  CGMState* state;     // [2][nrhs]
  CGDevResult* res;    // [nrhs]
  double tol;
  long max_iter;
  int nrhs;
  int nb;              // blocks of the iteration kernels (per system)
  int nb0;             // blocks of the start kernels (one site per thread)
  int nbs;             // stride of a partial slot (>= nb, nb0)
  int start;           // GNX_START_NULL / GNX_START_IMPOSED
};

// Launch shapes: SPT sites per thread; XCD = remap blocks so that each XCD
// (workgroups are dispatched round-robin over the 8 XCDs) sweeps one
// contiguous lex slab of both parities, keeping a slab's halo in its own L2.
constexpr int NUM_XCD = 8;

GNX_DEVONLY int xcd_block(int b, int nb) {
  int q = nb / NUM_XCD, rem = nb % NUM_XCD, x = b % NUM_XCD;
  return x * q + (x < rem ? x : rem) + b / NUM_XCD;
}

template <int SPT, bool XCD>
struct SiteMap {
  int base;
  GNX_DEVONLY SiteMap(int nb) { base = (XCD ? xcd_block(blockIdx.x, nb) : (int)blockIdx.x) * SPT * CGM_BS; }
  // j-th logical site of this thread, or -1
  GNX_DEVONLY int f(const Geom& g, int i) const {
    int j = base + i * CGM_BS + (int)threadIdx.x;
    if (j >= g.V) return -1;
    return XCD ? (j >> 1) + (j & 1) * g.half : j;
  }
};

GNX_DEVONLY double* slot(const CGMArgs& a, int s, int rhs) {
  return a.part + ((size_t)s * a.nrhs + rhs) * a.nbs;
}

// sum of the first n partials of a slot, fixed order, identical in every block
GNX_DEVONLY double reduce_slot(const CGMArgs& a, int s, int rhs, int n, double* lds) {
  const double* src = slot(a, s, rhs);
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += CGM_BS) acc += src[i];
  return block_sum(acc, lds);
}

// ---- start kernels ---------------------------------------------------------
// imposed start (conjugategradient.c:121-128): x = 0, r = p = b, partial |b|^2
__global__ __launch_bounds__(CGM_BS) void k_cgm_start_imposed(const CGMArgs* __restrict__ ap) {
  const CGMArgs a = *ap;
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const int f = blockIdx.x * CGM_BS + threadIdx.x;
  double acc = 0.0;
  if (f < a.g.V) {
    double2 bb = a.b[off + f];
    a.x[off + f] = make_double2(0.0, 0.0);
    a.r[off + f] = bb;
    a.p0[off + f] = bb;
    acc = cnorm2(bb);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) slot(a, 1, rhs)[blockIdx.x] = acc;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGMState s = {1, 0, 0, 0, 0.0};
    a.state[rhs] = s;
  }
}

// null start, pass 1 (conjugategradient.c:20): t = M^dag x_in
__global__ __launch_bounds__(CGM_BS) void k_cgm_null_p1(const CGMArgs* __restrict__ ap) {
  const CGMArgs a = *ap;
  const size_t off = (size_t)blockIdx.y * a.g.V;
  const int f = blockIdx.x * CGM_BS + threadIdx.x;
  if (f >= a.g.V) return;
  const double2* xin = a.x + off;
  a.t[off + f] = apply_Mdag(a.g, decode(a.g, f), a.sm[f], [&](int q) { return xin[q]; });
}

// null start, pass 2 (conjugategradient.c:19-21): r = b - M t; partials |r|^2
// (slot 0) and |b|^2 (slot 1)
__global__ __launch_bounds__(CGM_BS) void k_cgm_null_p2(const CGMArgs* __restrict__ ap) {
  const CGMArgs a = *ap;
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const int f = blockIdx.x * CGM_BS + threadIdx.x;
  double ar = 0.0, ab = 0.0;
  if (f < a.g.V) {
    const double2* t = a.t + off;
    double2 bb = a.b[off + f];
    double2 rr = apply_sub_M(a.g, decode(a.g, f), a.sm[f], bb, [&](int q) { return t[q]; });
    a.r[off + f] = rr;
    ar = cnorm2(rr);
    ab = cnorm2(bb);
  }
  ar = block_sum(ar, lds);
  ab = block_sum(ab, lds);
  if (threadIdx.x == 0) {
    slot(a, 0, rhs)[blockIdx.x] = ar;
    slot(a, 1, rhs)[blockIdx.x] = ab;
  }
}

// null start, pass 3 (conjugategradient.c:22-33): early exit keeps x_in;
// otherwise x = 0, r = p = b.  The |b|^2 partials stay in slot 1 for A(0).
__global__ __launch_bounds__(CGM_BS) void k_cgm_null_p3(const CGMArgs* __restrict__ ap) {
  const CGMArgs a = *ap;
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const size_t off = (size_t)rhs * a.g.V;
  const double nr = reduce_slot(a, 0, rhs, a.nb0, lds);
  const bool early = sqrt(nr) < a.tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CGMState s = {early ? 0 : 1, 0, early ? 1 : 0, early ? 1 : 0, 0.0};
    a.state[rhs] = s;
    if (early) {
      CGDevResult r = {1, 1, 0, 0, sqrt(nr), {0, 0, 0}};
      a.res[rhs] = r;
    }
  }
  if (early) return;
  const int f = blockIdx.x * CGM_BS + threadIdx.x;
  if (f < a.g.V) {
    double2 bb = a.b[off + f];
    a.x[off + f] = make_double2(0.0, 0.0);
    a.r[off + f] = bb;
    a.p0[off + f] = bb;
  }
}

// ---- iteration kernels -----------------------------------------------------
template <int SPT, bool XCD>
__global__ __launch_bounds__(CGM_BS) void k_cgm_A(const CGMArgs* __restrict__ ap) {
  const CGMArgs& a = *ap;
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const CGMState s = a.state[rhs];          // state[0]: written by the previous B / start
  CGMState* out = a.state + a.nrhs + rhs;   // state[1]
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!s.active) {
    if (lead) *out = s;
    return;
  }
  const Geom g = a.g;
  // slot 1 holds the start kernel's partials before the first iteration
  const double rho = reduce_slot(a, 1, rhs, s.iters == 0 ? a.nb0 : a.nb, lds);
  double beta = 0.0;
  if (s.iters == 0) {
    // ||b|| < tol returns before the loop only in the imposed start
    // (conjugategradient.c:123-128); the null start already tested its
    // prologue residual instead (:22)
    if (a.start == GNX_START_IMPOSED && sqrt(rho) < a.tol) {
      if (lead) {
        CGMState n = {0, 0, 1, 1, rho};
        *out = n;
        CGDevResult r = {1, 1, 0, 0, sqrt(rho), {0, 0, 0}};
        a.res[rhs] = r;
      }
      return;
    }
  } else {
    const bool conv = sqrt(rho) < a.tol;
    if (conv || s.iters >= a.max_iter) {
      if (lead) {
        CGMState n = {0, s.iters, conv ? 1 : 0, 0, rho};
        *out = n;
        CGDevResult r = {conv ? 1 : 0, 0, s.iters, 0, sqrt(rho), {0, 0, 0}};
        a.res[rhs] = r;
      }
      return;
    }
    beta = rho / s.rho;
  }
  const size_t off = (size_t)rhs * g.V;
  const double2* r = a.r + off;
  const double2* pold = ((s.iters & 1) ? a.p1 : a.p0) + off;
  double2* pnew = ((s.iters & 1) ? a.p0 : a.p1) + off;
  double2* t = a.t + off;
  const SiteMap<SPT, XCD> map(a.nb);
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(g, i);
    if (f < 0) break;
    auto ldp = [&](int q) { return cadd(r[q], rmul(beta, pold[q])); };
    pnew[f] = ldp(f);
    double2 tt = apply_Mdag(g, decode(g, f), a.sm[f], ldp);
    t[f] = tt;
    acc += cnorm2(tt);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) slot(a, 0, rhs)[blockIdx.x] = acc;
  if (lead) {
    CGMState n = {1, s.iters, 0, 0, rho};
    *out = n;
  }
}

template <int SPT, bool XCD>
__global__ __launch_bounds__(CGM_BS) void k_cgm_B(const CGMArgs* __restrict__ ap) {
  const CGMArgs& a = *ap;
  __shared__ double lds[CGM_BS / 64];
  const int rhs = blockIdx.y;
  const CGMState s = a.state[a.nrhs + rhs];  // state[1]
  CGMState* out = a.state + rhs;             // state[0]
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!s.active) {
    if (lead) *out = s;
    return;
  }
  const Geom g = a.g;
  const size_t off = (size_t)rhs * g.V;
  const double2* pnew = ((s.iters & 1) ? a.p0 : a.p1) + off;
  const double2* t = a.t + off;
  double2* r = a.r + off;
  double2* x = a.x + off;
  const SiteMap<SPT, XCD> map(a.nb);
  // the stencil does not depend on alpha: issue it before the reduction
  double2 w[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(g, i);
    if (f >= 0) w[i] = apply_M(g, decode(g, f), a.sm[f], [&](int q) { return t[q]; });
  }
  const double gamma = reduce_slot(a, 0, rhs, a.nb, lds);
  const double al = s.rho / gamma;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int f = map.f(g, i);
    if (f < 0) break;
    x[f] = cadd(x[f], rmul(al, pnew[f]));
    double2 rr = cadd(r[f], rmul(-al, w[i]));
    r[f] = rr;
    acc += cnorm2(rr);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) slot(a, 1, rhs)[blockIdx.x] = acc;
  if (lead) {
    CGMState n = {1, s.iters + 1, 0, 0, s.rho};
    *out = n;
  }
}

// launch shapes the tuner may pick from
struct CGMShape {
  int spt;
  bool xcd;
  const void* A;
  const void* B;
};

static const CGMShape kShapes[] = {
    {1, false, (const void*)k_cgm_A<1, false>, (const void*)k_cgm_B<1, false>},
    {1, true, (const void*)k_cgm_A<1, true>, (const void*)k_cgm_B<1, true>},
    {2, false, (const void*)k_cgm_A<2, false>, (const void*)k_cgm_B<2, false>},
    {2, true, (const void*)k_cgm_A<2, true>, (const void*)k_cgm_B<2, true>},
    {4, false, (const void*)k_cgm_A<4, false>, (const void*)k_cgm_B<4, false>},
    {4, true, (const void*)k_cgm_A<4, true>, (const void*)k_cgm_B<4, true>},
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// ---- host side ---------------------------------------------------------------
static int shape_blocks(const Geom& g, int shape) {
  int per = CGM_BS * kShapes[shape].spt;
  return (g.V + per - 1) / per;
}

static EngineKernels shape_kernels(const Geom& g, int shape) {
  EngineKernels k;
  k.X = kShapes[shape].A;
  k.Y = kShapes[shape].B;
  k.blocks = shape_blocks(g, shape);
  k.tag = shape;
  return k;
}

struct CGMachine {
  GraphEngine eng;
  int forced_shape = -1;  // GNX_CG_SHAPE
};

void cgm_release(void* m) { delete (CGMachine*)m; }

static CGMachine* cgm_get(void** slotp) {
  if (!*slotp) {
    CGMachine* m = new CGMachine();
    const char* env = getenv("GNX_CG_SHAPE");
    if (env) {
      int v = atoi(env);
      if (v >= 0 && v < kNumShapes) m->forced_shape = v;
    }
    *slotp = m;
  }
  return (CGMachine*)*slotp;
}

// stage the arguments and run the start kernels
static hipError_t cgm_begin(CGMachine* m, const Geom& g, const double* sm, const double2* b, double2* x, double tol,
                            long max_iter, int start, CGWork& w, int nrhs, int shape, hipStream_t st) {
  const int nb0 = (g.V + CGM_BS - 1) / CGM_BS;
  CGMArgs& a = *(CGMArgs*)m->eng.host_args();
  hipError_t e = hipStreamSynchronize(st);  // the staging buffer may still be in flight
  if (e != hipSuccess) return e;
  a.g = g;
  a.sm = sm;
  a.b = b;
  a.x = x;
  a.r = w.r;
  a.p0 = w.p0;
  a.p1 = w.p1;
  a.t = w.t;
  a.part = m->eng.dev_part();
  a.state = (CGMState*)m->eng.dev_state();
  a.res = w.res;
  a.tol = tol;
  a.max_iter = max_iter;
  a.nrhs = nrhs;
  a.nb = shape_blocks(g, shape);
  a.nb0 = nb0;
  a.nbs = nb0;
  a.start = start;
  e = m->eng.stage_args(st);
  if (e != hipSuccess) return e;
  const CGMArgs* dp = (const CGMArgs*)m->eng.dev_args();
  dim3 grid(nb0, nrhs), blk(CGM_BS);
  if (start == GNX_START_NULL) {
    hipLaunchKernelGGL(k_cgm_null_p1, grid, blk, 0, st, dp);
    hipLaunchKernelGGL(k_cgm_null_p2, grid, blk, 0, st, dp);
    hipLaunchKernelGGL(k_cgm_null_p3, grid, blk, 0, st, dp);
  } else {
    hipLaunchKernelGGL(k_cgm_start_imposed, grid, blk, 0, st, dp);
  }
  return hipGetLastError();
}

// Autotuner: time every launch shape on 2*CHUNK iterations of the caller's
// own system (imposed start into a scratch solution, so the caller's x is not
// touched) and keep the fastest per batch size.
static hipError_t cgm_tune(CGMachine* m, const Geom& g, const double* sm, const double2* b, CGWork& w, int nrhs,
                           hipStream_t st, int* best) {
  GraphEngine& E = m->eng;
  for (int i = 0; i < 8; i++)
    if (E.tuned_tag[i] >= 0 && E.tuned_nrhs[i] == nrhs) {
      *best = E.tuned_tag[i];
      return hipSuccess;
    }
  int best_shape = 0;
  hipError_t e = hipSuccess;
  if (m->forced_shape >= 0) {
    best_shape = m->forced_shape;
  } else {
    double2* xs = nullptr;
    e = hipMalloc(&xs, sizeof(double2) * (size_t)g.V * nrhs);
    if (e != hipSuccess) return e;
    float best_ms = 1e30f;
    for (int sh = 0; sh < kNumShapes && e == hipSuccess; sh++) {
      hipGraphExec_t exec;
      e = E.graph(shape_kernels(g, sh), nrhs, &exec);
      float ms_min = 1e30f;
      for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
        e = cgm_begin(m, g, sm, b, xs, 0.0, 2 * ENG_CHUNK, GNX_START_IMPOSED, w, nrhs, sh, st);
        float ms = 0.f;
        if (e == hipSuccess) e = E.time_replays(exec, 2, st, &ms);
        if (rep > 0 && ms < ms_min) ms_min = ms;
      }
      if (e == hipSuccess && ms_min < best_ms) {
        best_ms = ms_min;
        best_shape = sh;
      }
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(xs);
    if (e != hipSuccess) return e;
  }
  for (int i = 0; i < 8; i++)
    if (E.tuned_tag[i] < 0) {
      E.tuned_tag[i] = best_shape;
      E.tuned_nrhs[i] = nrhs;
      break;
    }
  *best = best_shape;
  return hipSuccess;
}

hipError_t cgm_f64_run(void** machine_slot, const Geom& g, const double* sm, const double2* b, double2* x,
                       double tol, long max_iter, int start, CGWork& w, int nrhs, hipStream_t st,
                       CGDevResult* host_res) {
  const int nbs = (g.V + CGM_BS - 1) / CGM_BS;
  CGMachine* m = cgm_get(machine_slot);
  hipError_t e = m->eng.ensure(nrhs, sizeof(CGMArgs), sizeof(CGMState), 2 * (size_t)(nrhs < 2 ? 2 : nrhs) * nbs);
  if (e != hipSuccess) return e;
  int shape = 0;
  e = cgm_tune(m, g, sm, b, w, nrhs, st, &shape);
  if (e != hipSuccess) return e;
  hipGraphExec_t exec;
  e = m->eng.graph(shape_kernels(g, shape), nrhs, &exec);
  if (e != hipSuccess) return e;
  e = cgm_begin(m, g, sm, b, x, tol, max_iter, start, w, nrhs, shape, st);
  if (e != hipSuccess) return e;
  e = m->eng.run(exec, nrhs, max_iter, st, nullptr);
  if (e != hipSuccess) return e;
  if (host_res) {
    e = hipMemcpyAsync(host_res, w.res, sizeof(CGDevResult) * nrhs, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
  }
  return e;
}

int cgm_shape_of(void* machine, int nrhs, int* spt, int* xcd) {
  CGMachine* m = (CGMachine*)machine;
  if (!m) return -1;
  for (int i = 0; i < 8; i++)
    if (m->eng.tuned_tag[i] >= 0 && m->eng.tuned_nrhs[i] == nrhs) {
      int sh = m->eng.tuned_tag[i];
      if (spt) *spt = kShapes[sh].spt;
      if (xcd) *xcd = kShapes[sh].xcd ? 1 : 0;
      return sh;
    }
  return -1;
}

}  // namespace gnx
