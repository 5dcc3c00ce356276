// gnx_solver.h -- device helpers shared by the graph-replayed CG solvers.
#pragma once

#include "gnx_common.h"
#include "gnx_engine.h"

namespace gnx {

constexpr int NUM_XCD = 8;

// ---- launch shapes ----------------------------------------------------------
// SPT sites per thread; XCD = remap blocks so that each XCD (workgroups are
// dispatched round-robin over the 8 XCDs) sweeps one contiguous lex slab of
// both parities, keeping a slab's halo in its own L2.
GNX_DEVONLY int xcd_block(int b, int nb) {
  int q = nb / NUM_XCD, rem = nb % NUM_XCD, x = b % NUM_XCD;
  return x * q + (x < rem ? x : rem) + b / NUM_XCD;
}

template <int SPT, bool XCD>
struct SiteMap {
  int base;
  GNX_DEVONLY SiteMap(int nb) { base = (XCD ? xcd_block(blockIdx.x, nb) : (int)blockIdx.x) * SPT * ENG_BS; }
  // i-th site of this thread, or -1
  GNX_DEVONLY int f(const Geom& g, int i) const {
    int j = base + i * ENG_BS + (int)threadIdx.x;
    if (j >= g.V) return -1;
    return XCD ? (j >> 1) + (j & 1) * g.half : j;
  }
};

// ---- streaming stores ----------------------------------------------------------
// Vector outputs of a solver kernel are not re-read by the kernel that writes
// them, and the next kernel (on any XCD) reads them from memory anyway, so
// they are written non-temporally instead of occupying the L2 until the
// end-of-kernel write-back.
typedef double gnx_d2v __attribute__((ext_vector_type(2)));
typedef float gnx_f2v __attribute__((ext_vector_type(2)));
GNX_DEVONLY void st_nt(double2* p, double2 v) {
  gnx_d2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, (gnx_d2v*)p);
}
GNX_DEVONLY void st_nt(float2* p, float2 v) {
  gnx_f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, (gnx_f2v*)p);
}
GNX_DEVONLY void st_nt(unsigned int* p, unsigned int v) { __builtin_nontemporal_store(v, p); }

// per-thread share of n partials, fixed order; the loads of a round are
// issued together (one memory latency per 8*BS partials, not one per load)
GNX_DEVONLY double partial_share(const double* __restrict__ src, int n) {
  constexpr int K = 8;
  double acc = 0.0;
  for (int base = 0; base < n; base += K * ENG_BS) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int i = base + k * ENG_BS + (int)threadIdx.x;
      v[k] = i < n ? src[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; k++) acc += v[k];
  }
  return acc;
}

}  // namespace gnx
