// gnx_cg.hip -- persistent FP64 conjugate gradient on A = M M^dag
// (conjugategradient.c:13-182), one cooperative launch per solve (or per
// batch of right-hand sides), two grid barriers per iteration.
//
// Schedule of iteration k (p_old/p_new double-buffered, beta_{-1} = 0):
//   phase A: p_new = r + beta*p_old at the site and, recomputed on the fly,
//            at its six neighbours; t = M^dag p_new; gamma_partial = |t|^2
//            -- barrier + deterministic reduction of gamma --
//   phase B: alpha = rho/gamma; x += alpha*p_new; r += (-alpha)*(M t);
//            rho_partial = |r|^2
//            -- barrier + deterministic reduction of rho --
//   exit when sqrt(rho) < tol (absolute, conjugategradient.c:87-89).
// The per-site arithmetic is the reference's (x += alpha p, r -= alpha M t,
// p = r + beta p, M and M^dag in reference order); only the global sums are
// taken in a fixed tree order instead of serially.  Every block reduces the
// per-block partials in the same order, so all blocks see bit-identical
// scalars and take the same branch.
#include <hip/hip_cooperative_groups.h>

#include "gnx_internal.h"

namespace cgrp = cooperative_groups;

namespace gnx {

constexpr int CG_BS = 256;
constexpr int CG_MAX_RHS = 64;
constexpr int NUM_XCD = 8;

struct CGParams {
  Geom g;
  const double* sm;
  const double2* b;
  double2* x;
  double2* r;
  double2* p0;
  double2* p1;
  double2* t;
  double* part;  // [2 slots][2 values][nrhs][nblocks]
  CGDevResult* res;
  double tol;
  long max_iter;
  int start;
  int nrhs;
};

// interleave the two checkerboard halves so that a contiguous chunk of j is a
// contiguous lex slab of both parities
GNX_DEV int f_of_j(const Geom& g, int j) { return (j >> 1) + (j & 1) * g.half; }

struct Chunk {
  int j0, j1;
};

GNX_DEV Chunk block_chunk(int V) {
  int nb = gridDim.x, b = blockIdx.x;
  // physical workgroups are dispatched round-robin over the XCDs; give each
  // XCD a contiguous range of chunks so a slab's halo stays in one L2
  int lb = (nb % NUM_XCD == 0) ? (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD : b;
  int chunk = (V + nb - 1) / nb;
  Chunk c;
  c.j0 = lb * chunk;
  c.j1 = min(V, c.j0 + chunk);
  return c;
}

GNX_DEV double* part_slot(const CGParams& p, int slot, int val, int rhs) {
  return p.part + ((size_t)((slot * 2 + val) * p.nrhs + rhs)) * gridDim.x;
}

// sum of the nblocks partials of (slot, val, rhs), fixed order, same in every block
GNX_DEV double gather_sum(const CGParams& p, int slot, int val, int rhs, double* lds) {
  const double* src = part_slot(p, slot, val, rhs);
  double acc = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += CG_BS) acc += src[i];
  return block_sum(acc, lds);
}

__global__ __launch_bounds__(CG_BS) void k_cg_f64(CGParams p) {
  cgrp::grid_group grid = cgrp::this_grid();
  __shared__ double lds[CG_BS / 64];
  __shared__ double s_rho[CG_MAX_RHS], s_beta[CG_MAX_RHS];
  __shared__ int s_active[CG_MAX_RHS];

  const Geom g = p.g;
  const int V = g.V;
  const Chunk ch = block_chunk(V);
  const int tid = threadIdx.x;
  const bool leader = (blockIdx.x == 0 && tid == 0);
  int slot = 0;

  // ------------------------------------------------------------ prologue
  if (p.start == GNX_START_NULL) {
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      const double2* xin = p.x + (size_t)rhs * V;
      double2* t = p.t + (size_t)rhs * V;
      for (int j = ch.j0 + tid; j < ch.j1; j += CG_BS) {
        int f = f_of_j(g, j);
        Site s = decode(g, f);
        t[f] = apply_Mdag(g, s, p.sm[f], [&](int k) { return xin[k]; });
      }
    }
    grid.sync();
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      const double2* t = p.t + (size_t)rhs * V;
      const double2* b = p.b + (size_t)rhs * V;
      double2* r = p.r + (size_t)rhs * V;
      double ar = 0.0, ab = 0.0;
      for (int j = ch.j0 + tid; j < ch.j1; j += CG_BS) {
        int f = f_of_j(g, j);
        Site s = decode(g, f);
        double2 bb = b[f];
        double2 rr = apply_sub_M(g, s, p.sm[f], bb, [&](int k) { return t[k]; });
        r[f] = rr;
        ar += cnorm2(rr);
        ab += cnorm2(bb);
      }
      ar = block_sum(ar, lds);
      ab = block_sum(ab, lds);
      if (tid == 0) {
        part_slot(p, slot, 0, rhs)[blockIdx.x] = ar;
        part_slot(p, slot, 1, rhs)[blockIdx.x] = ab;
      }
    }
    grid.sync();
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      double nr = gather_sum(p, slot, 0, rhs, lds);
      double nbb = gather_sum(p, slot, 1, rhs, lds);
      if (tid == 0) {
        bool early = sqrt(nr) < p.tol;
        s_active[rhs] = early ? 0 : 1;
        s_rho[rhs] = nbb;
        s_beta[rhs] = 0.0;
        if (leader) {
          p.res[rhs].early_exit = early ? 1 : 0;
          p.res[rhs].converged = early ? 1 : 0;
          p.res[rhs].iterations = 0;
          p.res[rhs].reliable_updates = 0;
          p.res[rhs].residual = sqrt(nr);
        }
      }
    }
    slot ^= 1;
    __syncthreads();
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      if (!s_active[rhs]) continue;
      const double2* b = p.b + (size_t)rhs * V;
      double2* x = p.x + (size_t)rhs * V;
      double2* r = p.r + (size_t)rhs * V;
      double2* p0 = p.p0 + (size_t)rhs * V;
      for (int j = ch.j0 + tid; j < ch.j1; j += CG_BS) {
        int f = f_of_j(g, j);
        double2 bb = b[f];
        x[f] = make_double2(0.0, 0.0);
        r[f] = bb;
        p0[f] = bb;
      }
    }
    grid.sync();
  } else {
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      const double2* b = p.b + (size_t)rhs * V;
      double2* x = p.x + (size_t)rhs * V;
      double2* r = p.r + (size_t)rhs * V;
      double2* p0 = p.p0 + (size_t)rhs * V;
      double ab = 0.0;
      for (int j = ch.j0 + tid; j < ch.j1; j += CG_BS) {
        int f = f_of_j(g, j);
        double2 bb = b[f];
        x[f] = make_double2(0.0, 0.0);
        r[f] = bb;
        p0[f] = bb;
        ab += cnorm2(bb);
      }
      ab = block_sum(ab, lds);
      if (tid == 0) part_slot(p, slot, 0, rhs)[blockIdx.x] = ab;
    }
    grid.sync();
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      double nbb = gather_sum(p, slot, 0, rhs, lds);
      if (tid == 0) {
        bool early = sqrt(nbb) < p.tol;
        s_active[rhs] = early ? 0 : 1;
        s_rho[rhs] = nbb;
        s_beta[rhs] = 0.0;
        if (leader) {
          p.res[rhs].early_exit = early ? 1 : 0;
          p.res[rhs].converged = early ? 1 : 0;
          p.res[rhs].iterations = 0;
          p.res[rhs].reliable_updates = 0;
          p.res[rhs].residual = sqrt(nbb);
        }
      }
    }
    slot ^= 1;
    __syncthreads();
  }

  // ------------------------------------------------------------ iterations
  int nactive = 0;
  for (int rhs = 0; rhs < p.nrhs; rhs++) nactive += s_active[rhs];

  for (long k = 0; k < p.max_iter && nactive > 0; k++) {
    double2* pold_base = (k & 1) ? p.p1 : p.p0;
    double2* pnew_base = (k & 1) ? p.p0 : p.p1;
    // phase A
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      if (!s_active[rhs]) continue;
      const double2* r = p.r + (size_t)rhs * V;
      const double2* pold = pold_base + (size_t)rhs * V;
      double2* pnew = pnew_base + (size_t)rhs * V;
      double2* t = p.t + (size_t)rhs * V;
      const double beta = s_beta[rhs];
      double acc = 0.0;
      for (int j = ch.j0 + tid; j < ch.j1; j += CG_BS) {
        int f = f_of_j(g, j);
        Site s = decode(g, f);
        auto ldp = [&](int q) { return cadd(r[q], rmul(beta, pold[q])); };
        pnew[f] = ldp(f);
        double2 tt = apply_Mdag(g, s, p.sm[f], ldp);
        t[f] = tt;
        acc += cnorm2(tt);
      }
      acc = block_sum(acc, lds);
      if (tid == 0) part_slot(p, slot, 0, rhs)[blockIdx.x] = acc;
    }
    grid.sync();
    const int slotA = slot;
    slot ^= 1;
    // phase B
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      if (!s_active[rhs]) continue;
      const double gamma = gather_sum(p, slotA, 0, rhs, lds);
      const double al = s_rho[rhs] / gamma;
      const double2* t = p.t + (size_t)rhs * V;
      const double2* pnew = pnew_base + (size_t)rhs * V;
      double2* x = p.x + (size_t)rhs * V;
      double2* r = p.r + (size_t)rhs * V;
      double acc = 0.0;
      for (int j = ch.j0 + tid; j < ch.j1; j += CG_BS) {
        int f = f_of_j(g, j);
        Site s = decode(g, f);
        x[f] = cadd(x[f], rmul(al, pnew[f]));
        double2 w = apply_M(g, s, p.sm[f], [&](int q) { return t[q]; });
        double2 rr = cadd(r[f], rmul(-al, w));
        r[f] = rr;
        acc += cnorm2(rr);
      }
      acc = block_sum(acc, lds);
      if (tid == 0) part_slot(p, slot, 0, rhs)[blockIdx.x] = acc;
    }
    grid.sync();
    const int slotB = slot;
    slot ^= 1;
    for (int rhs = 0; rhs < p.nrhs; rhs++) {
      if (!s_active[rhs]) continue;
      double rn = gather_sum(p, slotB, 0, rhs, lds);
      if (tid == 0) {
        if (sqrt(rn) < p.tol) {
          s_active[rhs] = 0;
          if (leader) {
            p.res[rhs].converged = 1;
            p.res[rhs].iterations = (int)(k + 1);
            p.res[rhs].residual = sqrt(rn);
          }
        } else {
          s_beta[rhs] = rn / s_rho[rhs];
          s_rho[rhs] = rn;
          if (leader) {
            p.res[rhs].iterations = (int)(k + 1);
            p.res[rhs].residual = sqrt(rn);
          }
        }
      }
    }
    __syncthreads();
    nactive = 0;
    for (int rhs = 0; rhs < p.nrhs; rhs++) nactive += s_active[rhs];
  }
}

static int cg_blocks(int device, int V) {
  static int cached_dev = -1, cached_max = 0;
  if (cached_dev != device) {
    int per_cu = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_cg_f64, CG_BS, 0) != hipSuccess)
      per_cu = 1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) ncu = 1;
    cached_max = per_cu * ncu;
    cached_dev = device;
  }
  int need = (V + CG_BS - 1) / CG_BS;
  int nb = need < cached_max ? need : cached_max;
  if (nb >= NUM_XCD) nb -= nb % NUM_XCD;
  return nb < 1 ? 1 : nb;
}

int cg_max_blocks_bound() { return 4096; }

hipError_t cg_f64_run(const Geom& g, const double* sm, const double2* b, double2* x, double tol, long max_iter,
                      int start, CGWork& w, int nrhs, hipStream_t st, int device, CGDevResult* host_res) {
  if (nrhs < 1 || nrhs > CG_MAX_RHS) return hipErrorInvalidValue;
  int nb = cg_blocks(device, g.V);
  if (nb > w.max_blocks) nb = w.max_blocks;
  CGParams p;
  p.g = g;
  p.sm = sm;
  p.b = b;
  p.x = x;
  p.r = w.r;
  p.p0 = w.p0;
  p.p1 = w.p1;
  p.t = w.t;
  p.part = w.part;
  p.res = w.res;
  p.tol = tol;
  p.max_iter = max_iter;
  p.start = start;
  p.nrhs = nrhs;
  void* args[] = {&p};
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_cg_f64, dim3(nb), dim3(CG_BS), args, 0, st);
  if (e != hipSuccess) return e;
  if (host_res) {
    e = hipMemcpyAsync(host_res, w.res, sizeof(CGDevResult) * nrhs, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
  }
  return e;
}

}  // namespace gnx
