// gnx_stencil.h -- the staggered Dirac operator at one site, in the exact
// accumulation order of the reference (bit-exact in FP64):
//
//   M   (dirac.c:46-82):  y = sm*v; for d=0..2: y += 0.5*eta_pp*v[pp]; y -= 0.5*eta_mm*v[mm]
//   M^+ (dirac.c:88-123): scatter form in the reference; its per-site order is
//       the ascending f-index of the six source sites.  Within the
//       opposite-parity half f-order is lex order, so the sources sort as
//       [below along x0][below x1][below x2][above x2][above x1][above x0],
//       where "below/above" compares the neighbour's coordinate with the
//       site's; at a wrap both neighbours along that axis fall on one side.
//   r -= M v (dirac.c:213-247): same order as M with flipped signs.
//
// Each hop term 0.5*eta*v is exact (a power of two times +-v), so only the
// order of the additions matters; the mass product sm*v is rounded first
// (the library is compiled with -ffp-contract=off).  In the gather forms a
// hop term enters by one explicit fused multiply-add (hop()), which rounds
// the exact product's sum once, as the separate product and sum do; the
// reference-order M^+ keeps its selected terms as separate products.
// Absent slots add -0.0, the exact additive identity.  The operators take
// the site's SiteCode (gnx_common.h), which a kernel forms in its load
// phase; the Geom/Site overloads form it on the spot.
#pragma once

#include <utility>

#include "gnx_common.h"

namespace gnx {

// Complex-type traits: the operator code is shared by the FP64 drop-in
// (double2, bit-exact) and the FP32-arithmetic inner solvers (float2).
template <typename T2>
struct CT;
template <>
struct CT<double2> {
  using R = double;
  GNX_DEV static double2 nzero() { return make_double2(-0.0, -0.0); }
};
template <>
struct CT<float2> {
  using R = float;
  GNX_DEV static float2 nzero() { return make_float2(-0.0f, -0.0f); }
};
GNX_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
GNX_DEV float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
GNX_DEV float2 rmul(float s, float2 a) { return make_float2(s * a.x, s * a.y); }
GNX_DEV float2 rmul(double s, float2 a) { return rmul((float)s, a); }
GNX_DEV float cnorm2(float2 a) { return a.x * a.x + a.y * a.y; }

// acc + c v for a hop coefficient c = +-0.5.  The product is exact (a power
// of two times v, for every v above the subnormal range), so one fused
// multiply-add rounds exactly like the separate product and sum of the
// reference, in one instruction per component.
GNX_DEV double2 hop(double c, double2 v, double2 acc) { return make_double2(fma(c, v.x, acc.x), fma(c, v.y, acc.y)); }
GNX_DEV float2 hop(float c, float2 v, float2 acc) { return make_float2(fmaf(c, v.x, acc.x), fmaf(c, v.y, acc.y)); }

// M v from operand values already in registers, the order of apply_M:
// v0 = v[site], vpp[d] = v[pp(z,d)], vmm[d] = v[mm(z,d)]
template <typename T2>
GNX_DEV T2 apply_M_vals(SiteCode k, typename CT<T2>::R sm, T2 v0, const T2 vpp[3], const T2 vmm[3]) {
  using R = typename CT<T2>::R;
  const HalfEtaT<R> e = half_eta_t<R>(k);
  T2 acc = rmul(sm, v0);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    acc = hop(e.pp[d], vpp[d], acc);
    acc = hop(-e.mm[d], vmm[d], acc);
  }
  return acc;
}
template <typename T2>
GNX_DEV T2 apply_M_vals(const Geom& g, const Site& s, typename CT<T2>::R sm, T2 v0, const T2 vpp[3],
                        const T2 vmm[3]) {
  return apply_M_vals(site_code(g, s), sm, v0, vpp, vmm);
}

// loads in the order the sum consumes them (the site first), so the
// arithmetic starts on the first value back
template <typename LoadF, typename T2 = decltype(std::declval<LoadF>()(0))>
GNX_DEV T2 apply_M(const Geom& g, const Site& s, typename CT<T2>::R sm, LoadF load) {
  NbrF n = neighbours_f(g, s);
  const T2 v0 = load(f_of_lex(g, s.lex, s.P));
  T2 vpp[3], vmm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    vpp[d] = load(n.pp[d]);
    vmm[d] = load(n.mm[d]);
  }
  return apply_M_vals(site_code(g, s), sm, v0, vpp, vmm);
}

// r - M v, reference order (r -= sm*v, then -= pp, += mm per direction)
template <typename T2>
GNX_DEV T2 apply_sub_M_vals(SiteCode k, typename CT<T2>::R sm, T2 r, T2 v0, const T2 vpp[3], const T2 vmm[3]) {
  using R = typename CT<T2>::R;
  const HalfEtaT<R> e = half_eta_t<R>(k);
  T2 acc = csub(r, rmul(sm, v0));
#pragma unroll
  for (int d = 0; d < 3; d++) {
    acc = hop(-e.pp[d], vpp[d], acc);
    acc = hop(e.mm[d], vmm[d], acc);
  }
  return acc;
}
template <typename LoadF, typename T2 = decltype(std::declval<LoadF>()(0))>
GNX_DEV T2 apply_sub_M(const Geom& g, const Site& s, typename CT<T2>::R sm, T2 r, LoadF load) {
  NbrF n = neighbours_f(g, s);
  const T2 v0 = load(f_of_lex(g, s.lex, s.P));
  T2 vpp[3], vmm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    vpp[d] = load(n.pp[d]);
    vmm[d] = load(n.mm[d]);
  }
  return apply_sub_M_vals(site_code(g, s), sm, r, v0, vpp, vmm);
}

// Ordered accumulation of the six M^+ (or -M^+) hop terms onto acc.
// tpp[d]: term from source pp(z,d); tmm[d]: term from source mm(z,d).
template <typename T2>
GNX_DEV T2 accumulate_adj_order(SiteCode k, T2 acc, const T2 tpp[3], const T2 tmm[3]) {
  const T2 z = CT<T2>::nzero();
  const unsigned topb[3] = {SC_X0TOP, SC_X1TOP, SC_X2TOP};
  const unsigned botb[3] = {SC_X0BOT, SC_X1BOT, SC_X2BOT};
  // below-slots: d = 0, 1, 2
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool top = (k.c & topb[d]) != 0, bot = (k.c & botb[d]) != 0;
    T2 s1 = top ? tpp[d] : (bot ? z : tmm[d]);
    T2 s2 = top ? tmm[d] : z;
    acc = cadd(acc, s1);
    acc = cadd(acc, s2);
  }
  // above-slots: d = 2, 1, 0
#pragma unroll
  for (int dd = 0; dd < 3; dd++) {
    const int d = 2 - dd;
    const bool top = (k.c & topb[d]) != 0, bot = (k.c & botb[d]) != 0;
    T2 s3 = top ? z : tpp[d];
    T2 s4 = bot ? tmm[d] : z;
    acc = cadd(acc, s3);
    acc = cadd(acc, s4);
  }
  return acc;
}
template <typename T2>
GNX_DEV T2 accumulate_adj_order(const Geom& g, const Site& s, T2 acc, const T2 tpp[3], const T2 tmm[3]) {
  return accumulate_adj_order(site_code(g, s), acc, tpp, tmm);
}

// M^+ v at site s.  Source mm(z,d) contributes +0.5*eta_pp[mm(z,d),d]*v, source
// pp(z,d) contributes -0.5*eta_mm[pp(z,d),d]*v.  eta_d depends only on
// coordinates below d, so eta at the source equals eta at z except for the
// antiperiodic time flips: eta_pp[mm(z,2),2] flips when z2 == 0 and
// eta_mm[pp(z,2),2] flips when z2 == LT-1.
GNX_DEV void adj_terms(SiteCode k, double hp[3], double hm[3]) {
  const double e1 = (k.c & SC_X0ODD) ? -0.5 : 0.5;
  const double e2 = (k.c & SC_X01ODD) ? -0.5 : 0.5;
  hp[0] = 0.5;  hm[0] = -0.5;
  hp[1] = e1;   hm[1] = -e1;
  hp[2] = (k.c & SC_X2BOT) ? -e2 : e2;   // coefficient of v[mm(z,2)]
  hm[2] = (k.c & SC_X2TOP) ? e2 : -e2;   // coefficient of v[pp(z,2)]
}
GNX_DEV void adj_terms(const Geom& g, const Site& s, double hp[3], double hm[3]) { adj_terms(site_code(g, s), hp, hm); }

// M^+ from operand values already in registers:
// v0 = v[site], vpp[d] = v[pp(z,d)], vmm[d] = v[mm(z,d)]
template <typename T2>
GNX_DEV T2 apply_Mdag_vals(SiteCode k, typename CT<T2>::R sm, T2 v0, const T2 vpp[3], const T2 vmm[3]) {
  double cm[3], cp[3];
  adj_terms(k, cm, cp);
  T2 tpp[3], tmm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    tmm[d] = rmul(cm[d], vmm[d]);
    tpp[d] = rmul(cp[d], vpp[d]);
  }
  return accumulate_adj_order(k, rmul(sm, v0), tpp, tmm);
}
template <typename T2>
GNX_DEV T2 apply_Mdag_vals(const Geom& g, const Site& s, typename CT<T2>::R sm, T2 v0, const T2 vpp[3],
                           const T2 vmm[3]) {
  return apply_Mdag_vals(site_code(g, s), sm, v0, vpp, vmm);
}

// M^+ in the plain gather form, M^+ = sm - D with D real antisymmetric:
//   y = sm v - sum_d [0.5 eta_pp(x,d) v(x+d) - 0.5 eta_mm(x,d) v(x-d)]
// For the solvers, which keep no reference order (the exact-order form
// above costs about 30 selects per site).
template <typename T2>
GNX_DEV T2 apply_Mdag_fast_vals(SiteCode k, typename CT<T2>::R sm, T2 v0, const T2 vpp[3], const T2 vmm[3]) {
  using R = typename CT<T2>::R;
  const HalfEtaT<R> e = half_eta_t<R>(k);
  T2 acc = rmul(sm, v0);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    acc = hop(-e.pp[d], vpp[d], acc);
    acc = hop(e.mm[d], vmm[d], acc);
  }
  return acc;
}
template <typename T2>
GNX_DEV T2 apply_Mdag_fast_vals(const Geom& g, const Site& s, typename CT<T2>::R sm, T2 v0, const T2 vpp[3],
                                const T2 vmm[3]) {
  return apply_Mdag_fast_vals(site_code(g, s), sm, v0, vpp, vmm);
}

// fused multiply-add of a complex vector by a real, for the reduced-precision
// inner solvers
GNX_DEV float2 cfma(float a, float2 x, float2 y) { return make_float2(fmaf(a, x.x, y.x), fmaf(a, x.y, y.y)); }

template <typename LoadF, typename T2 = decltype(std::declval<LoadF>()(0))>
GNX_DEV T2 apply_Mdag(const Geom& g, const Site& s, typename CT<T2>::R sm, LoadF load) {
  NbrF n = neighbours_f(g, s);
  T2 vpp[3], vmm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    vmm[d] = load(n.mm[d]);
    vpp[d] = load(n.pp[d]);
  }
  return apply_Mdag_vals(site_code(g, s), sm, load(f_of_lex(g, s.lex, s.P)), vpp, vmm);
}

// r - M^+ v (dirac.c:253-287): r -= sm*v, then the scatter terms negated, in
// the same ascending-source order.
template <typename LoadF, typename T2 = decltype(std::declval<LoadF>()(0))>
GNX_DEV T2 apply_sub_Mdag(const Geom& g, const Site& s, typename CT<T2>::R sm, T2 r, LoadF load) {
  NbrF n = neighbours_f(g, s);
  const SiteCode k = site_code(g, s);
  double cm[3], cp[3];
  adj_terms(k, cm, cp);
  int f = f_of_lex(g, s.lex, s.P);
  T2 tpp[3], tmm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    tmm[d] = rmul(-cm[d], load(n.mm[d]));
    tpp[d] = rmul(-cp[d], load(n.pp[d]));
  }
  return accumulate_adj_order(k, csub(r, rmul(sm, load(f))), tpp, tmm);
}

// ------------------------------------------------------------------ link-coupled operator
// The link-coupled staggered operator of the reference's commented-out
// copy_dirac.c:46-123 (the live dirac.c ignores the links):
//   M_U v(x)   = sm v(x) + sum_d [ (0.5 eta_pp U_d(x)) v(x+d) - (0.5 eta_mm conj U_d(x-d)) v(x-d) ]
//   M_U^+ v(x) = sm v(x) - sum_d [ same two terms ]   (gather form, d = 0,1,2)
// Links are U(1) phases stored in f-order, three per site: U[3 f + d] =
// U_d(site f).  Complex products use the term order of C99 complex
// multiplication, (ac - bd, ad + bc), with no fused multiply-add.
GNX_DEV double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
GNX_DEV double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// hop terms of site s onto acc: SIGN = +1 for M_U, -1 for M_U^+ and r -= M_U v
// vpp[d] = v(x+d), vmm[d] = v(x-d), Uo[d] = U_d(x), Um[d] = U_d(x-d)
template <int SIGN>
GNX_DEV double2 link_hops(SiteCode k, double2 acc, const double2 vpp[3], const double2 vmm[3], const double2 Uo[3],
                          const double2 Um[3]) {
  const HalfEta e = half_eta(k);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double2 tp = cmul(rmul(e.pp[d], Uo[d]), vpp[d]);
    const double2 tm = cmul(rmul(e.mm[d], cconj(Um[d])), vmm[d]);
    if (SIGN > 0) {
      acc = cadd(acc, tp);
      acc = csub(acc, tm);
    } else {
      acc = csub(acc, tp);
      acc = cadd(acc, tm);
    }
  }
  return acc;
}

template <int SIGN>
GNX_DEV double2 link_hops(const Geom& g, const Site& s, double2 acc, const double2 vpp[3], const double2 vmm[3],
                          const double2 Uo[3], const double2 Um[3]) {
  return link_hops<SIGN>(site_code(g, s), acc, vpp, vmm, Uo, Um);
}

// links of site s: own three and the three of its mm neighbours
GNX_DEV void load_links(const double2* __restrict__ U, int f, const NbrF& n, double2 Uo[3], double2 Um[3]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    Uo[d] = U[3 * f + d];
    Um[d] = U[3 * n.mm[d] + d];
  }
}

// op: 0 = M_U v, 1 = M_U^+ v, 2 = r - M_U v
template <typename LoadF>
GNX_DEV double2 apply_link(const Geom& g, const Site& s, int op, double sm, double2 r, const double2* __restrict__ U,
                           LoadF load) {
  NbrF n = neighbours_f(g, s);
  const int f = f_of_lex(g, s.lex, s.P);
  double2 vpp[3], vmm[3], Uo[3], Um[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    vpp[d] = load(n.pp[d]);
    vmm[d] = load(n.mm[d]);
  }
  load_links(U, f, n, Uo, Um);
  const double2 m0 = rmul(sm, load(f));
  if (op == 0) return link_hops<1>(g, s, m0, vpp, vmm, Uo, Um);
  if (op == 1) return link_hops<-1>(g, s, m0, vpp, vmm, Uo, Um);
  return link_hops<-1>(g, s, csub(r, m0), vpp, vmm, Uo, Um);
}

// ------------------------------------------------------------------ sigma sector
// f-index of the site with coordinates (c0,c1,c2)
GNX_DEV int f_of_coords(const Geom& g, int c0, int c1, int c2) {
  int lex = (c0 * g.L + c1) * g.LT + c2;
  return (lex >> 1) + ((c0 + c1 + c2) & 1) * g.half;
}

GNX_DEV double sigma_mass_at(const Geom& g, const Site& s, const double* __restrict__ sigma) {
  int a0 = s.x0, a1 = s.x1, a2 = s.x2;
  int b0 = a0 == 0 ? g.L - 1 : a0 - 1;
  int b1 = a1 == 0 ? g.L - 1 : a1 - 1;
  int b2 = a2 == 0 ? g.LT - 1 : a2 - 1;
  double m = sigma[f_of_coords(g, a0, a1, a2)];
  m += sigma[f_of_coords(g, b0, a1, a2)];
  m += sigma[f_of_coords(g, a0, b1, a2)];
  m += sigma[f_of_coords(g, a0, a1, b2)];
  m += sigma[f_of_coords(g, b0, b1, a2)];
  m += sigma[f_of_coords(g, b0, a1, b2)];
  m += sigma[f_of_coords(g, a0, b1, b2)];
  m += sigma[f_of_coords(g, b0, b1, b2)];
  return m / 8.0;
}

// sigma force of one flavour at dual site s (moleculardynamics.c:606-620):
// corners {x, x+0+1, x+0+2, x+1+2, x+0, x+1, x+2, x+0+1+2}, in that order.
GNX_DEV double force_flavour(const Geom& g, const Site& s, const double2* __restrict__ X,
                             const double2* __restrict__ Y, double F) {
  int a0 = s.x0, a1 = s.x1, a2 = s.x2;
  int u0 = a0 == g.L - 1 ? 0 : a0 + 1;
  int u1 = a1 == g.L - 1 ? 0 : a1 + 1;
  int u2 = a2 == g.LT - 1 ? 0 : a2 + 1;
  int c[8];
  c[0] = f_of_coords(g, a0, a1, a2);
  c[1] = f_of_coords(g, u0, u1, a2);
  c[2] = f_of_coords(g, u0, a1, u2);
  c[3] = f_of_coords(g, a0, u1, u2);
  c[4] = f_of_coords(g, u0, a1, a2);
  c[5] = f_of_coords(g, a0, u1, a2);
  c[6] = f_of_coords(g, a0, a1, u2);
  c[7] = f_of_coords(g, u0, u1, u2);
#pragma unroll
  for (int i = 0; i < 8; i++) F += (0.125) * cdot_re(X[c[i]], Y[c[i]]);
  return F;
}

}  // namespace gnx
