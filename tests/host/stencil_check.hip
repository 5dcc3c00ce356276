// CPU check of the device-side geometry and stencil code (the same inline
// functions the kernels use, instantiated for the host) against the oracle.
// Bit-exact: tables, sigma mass, M, M^dag, r-=Mv, r-=M^dag v, sigma force;
// the site code and the gather forms built on it (FP64 and FP32); the
// four-site group map and operand assembly.
// Build and run by tests/test_host_stencil.py; exit status 0 on success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../3d-gross-neveu-model_amd/csrc/gnx_stencil.h"
#include "../../3d-gross-neveu-model_amd/csrc/gnx_group.h"
#include "../../3d-gross-neveu-model_amd/csrc/gnx_tile.h"
#include "../../oracle/gnx_oracle.h"

using namespace gnx;

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails < 20) { printf(__VA_ARGS__); printf("\n"); } fails++; } } while (0)

static bool same(double a, double b) { return a == b; }

static void check(int L, int LT) {
  orc_geom og;
  if (orc_geom_init(&og, L, LT) != 0) { printf("bad size\n"); exit(2); }
  Geom g = make_geom(L, LT);
  int V = g.V;
  orc_rng rng;
  orc_rng_init(&rng, 1000 + L * 7 + LT);
  std::vector<double> sigma(V), sm(V), smo(V), b(2 * V), c(2 * V), out(2 * V);
  orc_draw_sigma(&rng, V, 2.0 / 3.4, sigma.data());
  orc_draw_z2_complex(&rng, V, b.data());
  for (int i = 0; i < 2 * V; i++) c[i] = orc_gauss1(&rng);
  orc_sigma_mass(&og, sigma.data(), smo.data());
  const double2* bv = (const double2*)b.data();
  auto ld = [&](int j) { return bv[j]; };
  for (int f = 0; f < V; f++) {
    Site s = decode(g, f);
    EXPECT(s.lex == og.f2lex[f], "L=%d LT=%d f=%d lex %d vs %d", L, LT, f, s.lex, og.f2lex[f]);
    EXPECT(f_of_lex(g, s.lex, s.P) == f, "f roundtrip %d", f);
    NbrF n = neighbours_f(g, s);
    HalfEta e = half_eta(g, s);
    for (int d = 0; d < 3; d++) {
      EXPECT(n.pp[d] == og.fpp[3 * f + d], "pp f=%d d=%d", f, d);
      EXPECT(n.mm[d] == og.fmm[3 * f + d], "mm f=%d d=%d", f, d);
      EXPECT(2.0 * e.pp[d] == og.eta_pp[3 * f + d], "eta_pp f=%d d=%d", f, d);
      EXPECT(2.0 * e.mm[d] == og.eta_mm[3 * f + d], "eta_mm f=%d d=%d", f, d);
    }
    sm[f] = sigma_mass_at(g, s, sigma.data());
    EXPECT(same(sm[f], smo[f]), "sigma mass f=%d", f);
  }
  std::vector<double> ref(2 * V);
  // M
  orc_dirac(&og, smo.data(), b.data(), ref.data());
  for (int f = 0; f < V; f++) {
    double2 y = apply_M(g, decode(g, f), smo[f], ld);
    EXPECT(same(y.x, ref[2 * f]) && same(y.y, ref[2 * f + 1]), "M L=%d f=%d", L, f);
  }
  orc_adj_dirac(&og, smo.data(), b.data(), ref.data());
  for (int f = 0; f < V; f++) {
    double2 y = apply_Mdag(g, decode(g, f), smo[f], ld);
    EXPECT(same(y.x, ref[2 * f]) && same(y.y, ref[2 * f + 1]), "Mdag L=%d LT=%d f=%d", L, LT, f);
  }
  ref = c;
  orc_sub_dirac(&og, smo.data(), b.data(), ref.data());
  for (int f = 0; f < V; f++) {
    double2 r = make_double2(c[2 * f], c[2 * f + 1]);
    double2 y = apply_sub_M(g, decode(g, f), smo[f], r, ld);
    EXPECT(same(y.x, ref[2 * f]) && same(y.y, ref[2 * f + 1]), "subM f=%d", f);
  }
  ref = c;
  orc_sub_adj_dirac(&og, smo.data(), b.data(), ref.data());
  for (int f = 0; f < V; f++) {
    double2 r = make_double2(c[2 * f], c[2 * f + 1]);
    double2 y = apply_sub_Mdag(g, decode(g, f), smo[f], r, ld);
    EXPECT(same(y.x, ref[2 * f]) && same(y.y, ref[2 * f + 1]), "subMdag f=%d", f);
  }
  // the site code from the load phase: its bits are the coordinates'
  // parities and boundary flags, and the gather forms that take it (hop
  // terms by fused multiply-add) equal the separate product-and-sum formula
  // with the oracle's eta tables, bit for bit, in FP64 and in FP32
  // (Gaussian operands, so that every sum rounds)
  const double2* cv = (const double2*)c.data();
  std::vector<float> c32(2 * V);
  for (int i = 0; i < 2 * V; i++) c32[i] = (float)c[i];
  const float2* bf = (const float2*)c32.data();
  for (int f = 0; f < V; f++) {
    const Site s = decode(g, f);
    const SiteGeo sg = site_geo(g, f);
    const NbrF n = neighbours_f(g, s);
    unsigned want = (unsigned)(s.x0 & 1) | ((unsigned)((s.x0 + s.x1) & 1) << 1);
    want |= (s.x2 == LT - 1 ? 4u : 0u) | (s.x2 == 0 ? 8u : 0u) | (s.x0 == L - 1 ? 16u : 0u) |
            (s.x0 == 0 ? 32u : 0u) | (s.x1 == L - 1 ? 64u : 0u) | (s.x1 == 0 ? 128u : 0u);
    EXPECT(sg.k.c == want, "site code f=%d: %u vs %u", f, sg.k.c, want);
    for (int d = 0; d < 3; d++)
      EXPECT(sg.n.pp[d] == n.pp[d] && sg.n.mm[d] == n.mm[d], "site_geo neighbours f=%d d=%d", f, d);
    double2 vpp[3], vmm[3];
    float2 wpp[3], wmm[3];
    for (int d = 0; d < 3; d++) {
      vpp[d] = cv[n.pp[d]];
      vmm[d] = cv[n.mm[d]];
      wpp[d] = bf[n.pp[d]];
      wmm[d] = bf[n.mm[d]];
    }
    double ex = smo[f] * cv[f].x, ey = smo[f] * cv[f].y;  // M^dag, gather form, separate products
    double mx = ex, my = ey;                               // M, the reference's order
    float fx = (float)smo[f] * bf[f].x, fy = (float)smo[f] * bf[f].y;
    float gx = fx, gy = fy;
    for (int d = 0; d < 3; d++) {
      const double hp = 0.5 * og.eta_pp[3 * f + d], hm = 0.5 * og.eta_mm[3 * f + d];
      ex = ex - hp * vpp[d].x; ey = ey - hp * vpp[d].y;
      ex = ex + hm * vmm[d].x; ey = ey + hm * vmm[d].y;
      mx = mx + hp * vpp[d].x; my = my + hp * vpp[d].y;
      mx = mx + (-hm) * vmm[d].x; my = my + (-hm) * vmm[d].y;
      const float hpf = (float)hp, hmf = (float)hm;
      fx = fx - hpf * wpp[d].x; fy = fy - hpf * wpp[d].y;
      fx = fx + hmf * wmm[d].x; fy = fy + hmf * wmm[d].y;
      gx = gx + hpf * wpp[d].x; gy = gy + hpf * wpp[d].y;
      gx = gx + (-hmf) * wmm[d].x; gy = gy + (-hmf) * wmm[d].y;
    }
    const double2 yd = apply_Mdag_fast_vals(sg.k, smo[f], cv[f], vpp, vmm);
    const double2 ym = apply_M_vals(sg.k, smo[f], cv[f], vpp, vmm);
    const float2 zd = apply_Mdag_fast_vals(sg.k, (float)smo[f], bf[f], wpp, wmm);
    const float2 zm = apply_M_vals(sg.k, (float)smo[f], bf[f], wpp, wmm);
    EXPECT(same(yd.x, ex) && same(yd.y, ey), "gather Mdag (FP64) f=%d", f);
    EXPECT(same(ym.x, mx) && same(ym.y, my), "M from the code (FP64) f=%d", f);
    EXPECT(zd.x == fx && zd.y == fy, "gather Mdag (FP32) f=%d", f);
    EXPECT(zm.x == gx && zm.y == gy, "M (FP32) f=%d", f);
  }
  // force from arbitrary X, Y
  std::vector<double> X0(2 * V), Y0(2 * V), X1(2 * V), Y1(2 * V), F(V), Fo(V);
  for (int i = 0; i < 2 * V; i++) { X0[i] = orc_gauss1(&rng); Y0[i] = orc_gauss1(&rng); X1[i] = orc_gauss1(&rng); Y1[i] = orc_gauss1(&rng); }
  double* Xs[2] = {X0.data(), X1.data()};
  double* Ys[2] = {Y0.data(), Y1.data()};
  orc_sigma_force_from_solutions(&og, sigma.data(), 0.85, Xs, Ys, Fo.data());
  for (int f = 0; f < V; f++) {
    Site s = decode(g, f);
    double Fv = 0.0;
    Fv = force_flavour(g, s, (const double2*)X0.data(), (const double2*)Y0.data(), Fv);
    Fv = force_flavour(g, s, (const double2*)X1.data(), (const double2*)Y1.data(), Fv);
    Fv -= (4 * 0.85 / 2.0) * sigma[f];
    EXPECT(same(Fv, Fo[f]), "force f=%d", f);
  }
  orc_geom_free(&og);
}

// decode() against 64-bit division at every site (no oracle tables: these
// volumes are too large for them); returns the number of wrong sites.
static long decode_scan(int L, int LT) {
  Geom g = make_geom(L, LT);
  long bad = 0;
  for (long f = 0; f < (long)g.V; f++) {
    long P = f >= g.half ? 1 : 0;
    long n0 = 2 * (f - P * g.half);
    long q = n0 / LT, x0 = q / L, x1 = q % L, x2 = n0 % LT + ((x0 + x1 + P) & 1);
    Site s = decode(g, (int)f);
    if (s.x0 != x0 || s.x1 != x1 || s.x2 != x2 || s.P != P || s.lex != q * LT + x2) bad++;
  }
  return bad;
}

// Tiles of the fused CG kernel on an L x L x LT lattice: the blocks' tiles
// cover every site exactly once; a block's items (tile, then halo) take
// distinct LDS box slots inside the box; every tile site decodes to the
// coordinates its item names; and the box slots at +-1 along each axis of a
// tile site hold exactly the site's six neighbours (f-indices of
// neighbours_f), so M^dag from LDS reads what M^dag from memory would.
template <int T0, int T1, int T2>
static void check_tiles(int L, int LT) {
  using TL = FTile<T0, T1, T2>;
  if (L % T0 || L % T1 || LT % T2) { printf("tile does not divide %d^2 x %d\n", L, LT); exit(2); }
  Geom g = make_geom(L, LT);
  const int nb = g.V / TL::N;
  std::vector<int> owner(g.V, 0);
  std::vector<int> slot_f(TL::NBOX);
  for (int b = 0; b < nb; b++) {
    int o0, o1, o2;
    ftile_origin<T0, T1, T2>(g, nb, b, &o0, &o1, &o2);
    EXPECT(o0 % T0 == 0 && o1 % T1 == 0 && o2 % T2 == 0 && o0 < L && o1 < L && o2 < LT,
           "tile %dx%dx%d L=%d LT=%d block %d origin %d %d %d", T0, T1, T2, L, LT, b, o0, o1, o2);
    std::fill(slot_f.begin(), slot_f.end(), -1);
    for (int e = 0; e < TL::NEXT; e++) {
      int bx = -1;
      const Site st = ftile_site<T0, T1, T2>(g, o0, o1, o2, e, &bx);
      const bool ok = st.x0 >= 0 && st.x0 < L && st.x1 >= 0 && st.x1 < L && st.x2 >= 0 && st.x2 < LT &&
                      bx >= 0 && bx < TL::NBOX;
      EXPECT(ok, "tile %dx%dx%d L=%d block %d item %d out of range", T0, T1, T2, L, b, e);
      if (!ok) continue;
      EXPECT(st.P == ((st.x0 + st.x1 + st.x2) & 1) && st.lex == (st.x0 * L + st.x1) * LT + st.x2, "site fields");
      const int f = f_of_lex(g, st.lex, st.P);
      EXPECT(slot_f[bx] < 0, "block %d item %d: box slot %d taken twice", b, e, bx);
      slot_f[bx] = f;
      if (e < TL::N) {
        const int a0 = st.x0 - o0, a1 = st.x1 - o1, a2 = st.x2 - o2;
        EXPECT(a0 >= 0 && a0 < T0 && a1 >= 0 && a1 < T1 && a2 >= 0 && a2 < T2, "item %d outside its tile", e);
        EXPECT(bx == ((a0 + 1) * TL::B1 + (a1 + 1)) * TL::B2 + (a2 + 1), "tile slot");
        owner[f]++;
        const Site d = decode(g, f);
        EXPECT(d.x0 == st.x0 && d.x1 == st.x1 && d.x2 == st.x2 && d.P == st.P && d.lex == st.lex, "decode");
      }
    }
    for (int e = 0; e < TL::N; e++) {
      int bx = -1;
      const Site st = ftile_site<T0, T1, T2>(g, o0, o1, o2, e, &bx);
      const NbrF n = neighbours_f(g, st);
      const int sd[3] = {TL::S0, TL::S1, 1};
      for (int k = 0; k < 3; k++) {
        EXPECT(slot_f[bx + sd[k]] == n.pp[k], "L=%d block %d item %d: +%d neighbour in LDS %d, expected %d", L, b,
               e, k, slot_f[bx + sd[k]], n.pp[k]);
        EXPECT(slot_f[bx - sd[k]] == n.mm[k], "L=%d block %d item %d: -%d neighbour in LDS %d, expected %d", L, b,
               e, k, slot_f[bx - sd[k]], n.mm[k]);
      }
    }
  }
  for (int f = 0; f < g.V; f++) EXPECT(owner[f] == 1, "tile %dx%dx%d L=%d: site %d owned %d times", T0, T1, T2, L, f, owner[f]);
}

// Four-site groups (LT a multiple of 8): the thread-to-group map covers every
// group once with aligned, in-range first sites, and each group's seven
// 16-byte loads and operand assembly give every site the operand indices and
// site code of its own decode.
static void check_groups(int L, int LT) {
  const Geom g = make_geom(L, LT);
  const int nb = (g.V + 4 * ENG_BS - 1) / (4 * ENG_BS);
  std::vector<int> seen(g.V / 4, 0);
  for (int b = 0; b < nb; b++)
    for (int t = 0; t < ENG_BS; t++) {
      const int f0 = group_first(g, nb, b, t);
      if (f0 < 0) continue;
      EXPECT(f0 % 4 == 0 && f0 + 3 < g.V, "groups %dx%d: block %d thread %d first site %d", L, LT, b, t, f0);
      if (f0 % 4 == 0 && f0 >= 0 && f0 + 3 < g.V) seen[f0 / 4]++;
    }
  for (int h = 0; h < g.V / 4; h++) EXPECT(seen[h] == 1, "groups %dx%d: group %d owned %d times", L, LT, h, seen[h]);
  for (int f0 = 0; f0 < g.V; f0 += 4) {
    const GroupGeo q = group_geo(g, f0);
    const int gi[7] = {q.own, q.pp0, q.mm0, q.pp1, q.mm1, q.g2, q.s2};
    gnx_u4 w[7];
    for (int k = 0; k < 7; k++) {
      EXPECT(gi[k] % 4 == 0 && gi[k] >= 0 && gi[k] + 3 < g.V, "groups %dx%d: group %d load %d at %d", L, LT, f0, k,
             gi[k]);
      for (int e = 0; e < 4; e++) w[k][e] = (unsigned)(gi[k] + e);
    }
    unsigned o[4][7];
    group_operands<unsigned>(w, q.even, o);
    for (int j = 0; j < 4; j++) {
      const Site s = decode(g, f0 + j);
      const NbrF n = neighbours_f(g, s);
      const int want[7] = {f0 + j, n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
      for (int k = 0; k < 7; k++)
        EXPECT((int)o[j][k] == want[k], "groups %dx%d: site %d operand %d is %u, expected %d", L, LT, f0 + j, k,
               o[j][k], want[k]);
      EXPECT(q.k[j].c == site_code(g, s).c, "groups %dx%d: site %d code %u, expected %u", L, LT, f0 + j, q.k[j].c,
             site_code(g, s).c);
    }
  }
}

int main() {
  int gsizes[][2] = {{4, 8}, {8, 8}, {12, 8}, {6, 24}, {16, 16}, {16, 32}, {10, 40}, {64, 64}, {96, 96}};
  for (auto& s : gsizes) check_groups(s[0], s[1]);
  check_tiles<4, 8, 16>(16, 16);
  check_tiles<4, 8, 16>(8, 16);
  check_tiles<4, 8, 16>(32, 32);
  check_tiles<4, 8, 16>(24, 48);
  check_tiles<4, 8, 16>(64, 64);
  check_tiles<4, 8, 16>(96, 96);
  check_tiles<8, 8, 8>(8, 8);
  check_tiles<8, 8, 8>(16, 24);
  check_tiles<8, 8, 8>(48, 48);
  check_tiles<8, 8, 8>(64, 64);
  int sizes[][2] = {{4, 4}, {4, 6}, {6, 4}, {6, 10}, {8, 8}, {10, 4}, {12, 12}, {16, 8}, {16, 16}, {24, 24}, {32, 32}};
  for (auto& s : sizes) check(s[0], s[1]);
  // the exact range of decode(): every accepted shape decodes exactly, and
  // shapes past the range (accepted by the 2^26-site cap alone) do not
  int accepted[][2] = {{96, 96}, {256, 1024}, {382, 382}, {388, 200}, {1354, 16}, {2188, 4}, {2048, 16}};
  for (auto& s : accepted) {
    EXPECT(geom_decodable(s[0], s[1]), "%d^2 x %d should be accepted", s[0], s[1]);
    long bad = decode_scan(s[0], s[1]);
    EXPECT(bad == 0, "%d^2 x %d: %ld sites decode wrong", s[0], s[1], bad);
  }
  int rejected[][2] = {{128, 4000}, {256, 1000}, {220, 1364}, {4094, 4}, {2046, 16}};
  for (auto& s : rejected) {
    EXPECT(!geom_decodable(s[0], s[1]), "%d^2 x %d should be rejected", s[0], s[1]);
    EXPECT(decode_scan(s[0], s[1]) > 0, "%d^2 x %d decodes exactly: the range check is too strict here",
           s[0], s[1]);
  }
  if (fails) { printf("%d failures\n", fails); return 1; }
  printf("stencil_check: all sizes bit-exact\n");
  return 0;
}
