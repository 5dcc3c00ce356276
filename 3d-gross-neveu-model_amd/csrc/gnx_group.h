// gnx_group.h -- groups of four consecutive sites per thread (the four-site
// mixed-precision CG kernels); host-callable so the map and the operand
// assembly can be checked on the CPU.
#pragma once

#include "gnx_common.h"
#include "gnx_solver.h"

namespace gnx {

// ---- groups of four consecutive sites ------------------------------------------------
// With LT a multiple of 8, the f-sites 4h .. 4h+3 of one parity lie in one
// parity row (x0, x1 fixed, x2 = a, a+2, a+4, a+6 with a = 0 or 1 mod 8), so
// a thread owning such a group reads its own operands and those of the four
// sites' x0 and x1 neighbours as one aligned 16-byte load each (4-byte
// storage).  Along x2 the neighbours are the other parity's group g2 at the
// same offset in the row, for x2+1 when a is even and for x2-1 when a is
// odd, and the other direction is g2 shifted by one site: completed by one
// site of the adjacent group s2 (the previous one for a even, the next one
// for a odd, wrapping at the row's ends).
typedef unsigned int gnx_u4 __attribute__((ext_vector_type(4)));

// First site of thread t's group in block b of nb, or -1.  The order of
// SiteMap<4, true>: per 128 threads, 64 groups of the even parity, then the
// 64 odd groups at the same positions; blocks on XCD slabs, V / 1024 blocks
// per system.
GNX_DEV int group_first(const Geom& g, int nb, int b, int t) {
  const int jg = xcd_block(b, nb) * ENG_BS + t;
  const int halfg = g.half >> 2;
  if (jg >= 2 * halfg) return -1;
  const int G = jg >> 7, u = jg & 127;
  const int r = halfg - 64 * G < 64 ? halfg - 64 * G : 64;  // groups per parity here (< 64 only in the last)
  return 4 * (64 * G + (u < r ? u : u - r + halfg));
}

// where a group's operands are (first sites of aligned groups) and its four
// site codes
struct GroupGeo {
  int own, pp0, mm0, pp1, mm1;  // the group and its x0, x1 neighbour groups
  int g2, s2;                   // the x2 neighbours: aligned group, adjacent group
  bool even;                    // a even: g2 holds x2+1, s2's last site completes x2-1
  SiteCode k[4];
};

GNX_DEV GroupGeo group_geo(const Geom& g, int f0) {
  const Site s = decode(g, f0);
  const NbrF n = neighbours_f(g, s);
  const SiteCode k0 = site_code(g, s);
  GroupGeo q;
  q.own = f0;
  q.pp0 = n.pp[0];
  q.mm0 = n.mm[0];
  q.pp1 = n.pp[1];
  q.mm1 = n.mm[1];
  const int off = (1 - s.P) * g.half;  // the other parity
  const int h0 = s.lex >> 1;           // the group's offset within its parity
  const int row = g.LT >> 1;           // sites per parity row
  q.g2 = h0 + off;
  q.even = (s.x2 & 1) == 0;
  if (q.even)  // x2 - 1 of the first site: the previous group's last, or the row's last at x2 = 0
    q.s2 = (s.x2 == 0 ? h0 + row - 4 : h0 - 4) + off;
  else         // x2 + 1 of the last site: the next group's first, or the row's first at x2 = LT - 1
    q.s2 = (s.x2 + 6 == g.LT - 1 ? h0 + 4 - row : h0 + 4) + off;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int x2 = s.x2 + 2 * j;
    q.k[j].c = (k0.c & ~(SC_X2TOP | SC_X2BOT)) | (x2 == g.LT - 1 ? SC_X2TOP : 0u) | (x2 == 0 ? SC_X2BOT : 0u);
  }
  return q;
}

// The 4 x 7 operands of a group from its seven 16-byte loads (own, pp0, mm0,
// pp1, mm1, g2, s2): o[j] = {site, pp0, pp1, pp2, mm0, mm1, mm2} of site j,
// the order of the stencil helpers' (v0, vpp[3], vmm[3]).
template <typename T>
GNX_DEV void group_operands(const gnx_u4 w[7], bool even, T o[4][7]) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int jn = j < 3 ? j + 1 : 3, jp = j > 0 ? j - 1 : 0;
    const T up = j < 3 ? (T)w[5][jn] : (T)w[6].x;  // g2 shifted towards x2 + 1
    const T dn = j > 0 ? (T)w[5][jp] : (T)w[6].w;  // g2 shifted towards x2 - 1
    o[j][0] = w[0][j];
    o[j][1] = w[1][j];
    o[j][2] = w[3][j];
    o[j][3] = even ? (T)w[5][j] : up;
    o[j][4] = w[2][j];
    o[j][5] = w[4][j];
    o[j][6] = even ? dn : (T)w[5][j];
  }
}

}  // namespace gnx
