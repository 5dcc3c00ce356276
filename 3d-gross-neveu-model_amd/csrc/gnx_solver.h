// gnx_solver.h -- device helpers shared by the graph-replayed CG solvers.
#pragma once

#include "gnx_common.h"
#include "gnx_engine.h"

namespace gnx {

constexpr int NUM_XCD = 8;

// ---- launch shapes ----------------------------------------------------------
// SPT sites per thread; XCD = remap blocks so that each XCD (workgroups are
// dispatched round-robin over the 8 XCDs) sweeps one contiguous lex slab of
// both parities, keeping a slab's halo in its own L2.  Inside the slab the
// thread-site index j runs in groups of 128: the first 64 are even sites, the
// next 64 the odd sites at the same positions, so every wave reads one
// parity's contiguous run (lanes alternating between the two halves, 2 MB
// apart at 64^3, cost the L1 half its throughput).
GNX_DEV int xcd_block(int b, int nb) {
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
    if (!XCD) return j;
    const int G = j >> 7, t = j & 127;
    const int r = min(64, g.half - 64 * G);  // sites per parity in this group (< 64 only in the last)
    return 64 * G + (t < r ? t : t - r + g.half);
  }
};

// ---- addressing ------------------------------------------------------------------
// Address of element i of a vector whose base is wave-uniform, as the base
// plus a 32-bit byte offset: the compiler then uses the scalar-base +
// vector-offset address form, one instruction per address instead of a
// 64-bit shift-and-add.  Volumes are capped at 2^26 sites, so offsets of
// vectors of up to 64 bytes per site stay below 2^32.
template <typename T>
GNX_DEVONLY T* elem(T* base, int i) {
  return (T*)((char*)base + (unsigned)i * (unsigned)sizeof(T));
}
template <typename T>
GNX_DEVONLY const T* elem(const T* base, int i) {
  return (const T*)((const char*)base + (unsigned)i * (unsigned)sizeof(T));
}

// ---- partial slots --------------------------------------------------------------
// Slot of block b's partial sum among nb.  Workgroups are dispatched
// round-robin over the XCDs, so with slot = b the 16 partials of one 128-byte
// line come from all eight XCDs and every XCD writes back a partial line;
// grouping the blocks of one XCD (slot = (b mod 8) * nb/8 + b/8, when 8
// divides nb) gives each line one writer.  Every reader sums all nb slots in
// slot order, so the sums stay deterministic.
GNX_DEVONLY int partial_slot(int b, int nb) {
  return (nb % NUM_XCD == 0) ? (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD : b;
}

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

// Per-thread share of n >= 1 partials, fixed order, in two steps so that a
// kernel can issue every load it needs before it waits for any:
// partial_issue loads the first round (up to 4*BS partials) with no branch
// -- every lane reads an in-bounds index, clamped to n-1 -- and
// partial_finish masks the surplus, sums, and continues in a loop when n is
// larger.
constexpr int PART_K = 4;
GNX_DEVONLY void partial_issue(const double* __restrict__ src, int n, double v[PART_K]) {
#pragma unroll
  for (int k = 0; k < PART_K; k++) {
    const int i = k * ENG_BS + (int)threadIdx.x;
    v[k] = src[i < n ? i : n - 1];
  }
}
GNX_DEVONLY double partial_finish(const double* __restrict__ src, int n, const double v[PART_K]) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < PART_K; k++) acc += (k * ENG_BS + (int)threadIdx.x < n) ? v[k] : 0.0;
  for (int base = PART_K * ENG_BS; base < n; base += PART_K * ENG_BS) {
    double w[PART_K];
#pragma unroll
    for (int k = 0; k < PART_K; k++) {
      const int i = base + k * ENG_BS + (int)threadIdx.x;
      w[k] = i < n ? src[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < PART_K; k++) acc += w[k];
  }
  return acc;
}
GNX_DEVONLY double partial_share(const double* __restrict__ src, int n) {
  double v[PART_K];
  partial_issue(src, n, v);
  return partial_finish(src, n, v);
}

// Head reduction of the iteration kernels, in two forms chosen per launch
// (a template argument: a branch between them would join two load paths,
// and the compiler then waits for every load at the join, before the
// operand loads).  WH (up to HEAD_W * 64 = 512 partials): every wave loads
// all of them itself (lane l: l, l + 64, ..., clamped, branch-free) and sums
// them with DPP -- no LDS and no barrier, the same order in every wave of
// every block; a block-wide sum of those partials costs a barrier that holds
// every wave to the slowest (11.34 against 11.17 us per FP64 iteration at
// 64^3, 11.74 against 11.11 with 4 sites per thread).  Otherwise the
// block-wide form above (8 loads per lane would leave a second round of
// loads behind the first wait).
constexpr int HEAD_W = 8;
GNX_DEV bool head_wave(int n) { return n <= HEAD_W * 64; }
template <bool WH>
GNX_DEVONLY void head_issue(const double* __restrict__ src, int n, double v[HEAD_W]) {
  if constexpr (WH) {
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int k = 0; k < HEAD_W; k++) {
      const int i = k * 64 + lane;
      v[k] = src[i < n ? i : n - 1];
    }
  } else {
    partial_issue(src, n, v);
  }
}
template <bool WH>
GNX_DEVONLY double head_finish(const double* __restrict__ src, int n, const double v[HEAD_W], double* lds) {
  if constexpr (WH) {
    const int lane = (int)(threadIdx.x & 63);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < HEAD_W; k++) acc += (k * 64 + lane < n) ? v[k] : 0.0;
    return wave_sum(acc);
  } else {
    return block_sum_once(partial_finish(src, n, v), lds);
  }
}

// The same for n pairs of partials (two sums taken at one point), one
// 16-byte load per pair; each component is summed in the order of
// head_issue/head_finish.  The block form's lds must hold 2 * ENG_BS / 64
// doubles.
template <bool WH>
GNX_DEVONLY void head_issue2(const double2* __restrict__ src, int n, double2 v[HEAD_W]) {
  if constexpr (WH) {
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int k = 0; k < HEAD_W; k++) {
      const int i = k * 64 + lane;
      v[k] = src[i < n ? i : n - 1];
    }
  } else {
#pragma unroll
    for (int k = 0; k < PART_K; k++) {
      const int i = k * ENG_BS + (int)threadIdx.x;
      v[k] = src[i < n ? i : n - 1];
    }
  }
}
template <bool WH>
GNX_DEVONLY double2 head_finish2(const double2* __restrict__ src, int n, const double2 v[HEAD_W], double* lds) {
  double ax = 0.0, ay = 0.0;
  if constexpr (WH) {
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int k = 0; k < HEAD_W; k++) {
      const bool in = k * 64 + lane < n;
      ax += in ? v[k].x : 0.0;
      ay += in ? v[k].y : 0.0;
    }
    return make_double2(wave_sum(ax), wave_sum(ay));
  } else {
#pragma unroll
    for (int k = 0; k < PART_K; k++) {
      const bool in = k * ENG_BS + (int)threadIdx.x < n;
      ax += in ? v[k].x : 0.0;
      ay += in ? v[k].y : 0.0;
    }
    for (int base = PART_K * ENG_BS; base < n; base += PART_K * ENG_BS) {
      double2 w[PART_K];
#pragma unroll
      for (int k = 0; k < PART_K; k++) {
        const int i = base + k * ENG_BS + (int)threadIdx.x;
        w[k] = i < n ? src[i] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int k = 0; k < PART_K; k++) {
        ax += w[k].x;
        ay += w[k].y;
      }
    }
    return block_sum2(ax, ay, lds);
  }
}
// Marks the end of a kernel's load phase: the scheduler moves no
// instruction across it, so no arithmetic on a loaded value (and no wait for
// one) lands between the loads.
GNX_DEVONLY void end_of_loads() { __builtin_amdgcn_sched_barrier(0); }

// Per-system solver state, written by the previous kernel, read with cached
// vector loads.  A uniform address would be read with a scalar load, waited
// on with lgkmcnt(0) -- a wait that also covers every later scalar load of a
// kernel argument -- so one more memory round trip would sit ahead of the
// operand loads.  mbcnt_lo(0, 0) is 0 in every lane, but the compiler cannot
// treat the index as uniform, so the fields come in as vector loads that
// return with the operands.  (Atomic loads would also be vector loads, but
// at device scope they bypass the caches and every wave of the grid would
// queue on one line.)
template <typename T>
GNX_DEVONLY T load_state(const T* p) {
  const int zero = (int)__builtin_amdgcn_mbcnt_lo(0u, 0u);
  return p[zero];
}

}  // namespace gnx
