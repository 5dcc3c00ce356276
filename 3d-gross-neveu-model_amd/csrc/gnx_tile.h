// gnx_tile.h -- tiles of the fused CG kernel (one kernel per iteration):
// which sites a block owns, the one-site halo it recomputes, and where each
// sits in the block's LDS box.  Host-callable, so that the indexing can be
// checked on the CPU.
#pragma once

#include "gnx_common.h"
#include "gnx_solver.h"

namespace gnx {

template <int T0, int T1, int T2>
struct FTile {
  static constexpr int N = T0 * T1 * T2;                 // tile sites
  static constexpr int HR = T2 / 2;                      // sites of one parity in a tile row
  static constexpr int NF0 = 2 * T1 * T2, NF1 = 2 * T0 * T2, NF2 = 2 * T0 * T1;
  static constexpr int NEXT = N + NF0 + NF1 + NF2;      // tile and halo
  static constexpr int B1 = T1 + 2, B2 = T2 + 2;         // LDS box: (T0+2) x (T1+2) x (T2+2)
  static constexpr int S0 = B1 * B2, S1 = B2;            // box strides of x0 and x1
  static constexpr int NBOX = (T0 + 2) * B1 * B2;
  static constexpr int SPT = N / ENG_BS;                 // tile sites per thread
  static constexpr int ROUNDS = (NEXT + ENG_BS - 1) / ENG_BS;
  static_assert(N % ENG_BS == 0 && T2 % 2 == 0, "tile shape");
};

GNX_DEV int wrapc(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// site of item e of a tile with origin (o0, o1, o2), and its LDS box slot.
// Items: the tile's sites, then the halo faces at x0 = -1, T0, at x1 = -1,
// T1 (all four in rows of HR sites of one parity, the parities alternating
// row by row), then at x2 = -1, T2.
template <int T0, int T1, int T2>
GNX_DEV Site ftile_site(const Geom& g, int o0, int o1, int o2, int e, int* box) {
  using TL = FTile<T0, T1, T2>;
  int a0, a1, a2, x0, x1, x2;
  if (e < TL::N + TL::NF0 + TL::NF1) {
    int r, j;
    if (e < TL::N) {
      j = e % TL::HR;
      r = e / TL::HR;
      a0 = (r >> 1) / T1;
      a1 = (r >> 1) % T1;
    } else if (e < TL::N + TL::NF0) {
      e -= TL::N;
      j = e % TL::HR;
      r = e / TL::HR;
      a0 = (r >> 1) / T1 ? T0 : -1;
      a1 = (r >> 1) % T1;
    } else {
      e -= TL::N + TL::NF0;
      j = e % TL::HR;
      r = e / TL::HR;
      a1 = (r >> 1) / T0 ? T1 : -1;
      a0 = (r >> 1) % T0;
    }
    x0 = wrapc(o0 + a0, g.L);
    x1 = wrapc(o1 + a1, g.L);
    a2 = 2 * j + ((x0 + x1 + (r & 1)) & 1);
    x2 = o2 + a2;
  } else {
    e -= TL::N + TL::NF0 + TL::NF1;
    const int side = e / (T0 * T1), rem = e % (T0 * T1);
    a0 = rem / T1;
    a1 = rem % T1;
    a2 = side ? T2 : -1;
    x0 = o0 + a0;
    x1 = o1 + a1;
    x2 = wrapc(o2 + a2, g.LT);
  }
  *box = ((a0 + 1) * TL::B1 + (a1 + 1)) * TL::B2 + (a2 + 1);
  Site s;
  s.x0 = x0;
  s.x1 = x1;
  s.x2 = x2;
  s.P = (x0 + x1 + x2) & 1;
  s.lex = (x0 * g.L + x1) * g.LT + x2;
  return s;
}

// tile origin of block b among nb (XCD slabs along x0: the blocks of one XCD
// take consecutive tiles, x0 slowest)
template <int T0, int T1, int T2>
GNX_DEV void ftile_origin(const Geom& g, int nb, int b, int* o0, int* o1, int* o2) {
  const int n2 = g.LT / T2, n12 = (g.L / T1) * n2;
  const int v = xcd_block(b, nb);
  const int t0 = v / n12, rem = v - t0 * n12;
  const int t1 = rem / n2;
  *o0 = t0 * T0;
  *o1 = t1 * T1;
  *o2 = (rem - t1 * n2) * T2;
}

}  // namespace gnx
