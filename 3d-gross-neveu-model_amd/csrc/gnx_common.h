// gnx_common.h -- device-side building blocks shared by the gnx kernels.
//
// Lattice geometry is computed in-kernel from the f-index (no neighbour or
// phase tables are read from memory):
//   lex n = (x0*L + x1)*LT + x2                  (coordinate.c:25-65)
//   f     = n/2 + parity*V/2, parity=(x0+x1+x2)&1 (coordinate.c:301-313)
//   eta_0 = 1, eta_1 = (-1)^x0, eta_2 = (-1)^(x0+x1); the antiperiodic time
//   boundary flips the forward hop leaving x2=LT-1 and the backward hop
//   leaving x2=0 (coordinate.c:133-200, macro.h:100).
// Complex vectors are double2 (re, im) per site in f-order, the layout of the
// reference's Fermion (fermion.h:49-52), so no permutation at the boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// geometry and stencil helpers are host-callable too, so they can be checked
// on the CPU against the oracle before any GPU run
#define GNX_DEV __host__ __device__ __forceinline__
#define GNX_DEVONLY __device__ __forceinline__

namespace gnx {

struct Geom {
  int L, LT, V, half;
  int LLT;            // L*LT   (stride of x0 in lex)
  unsigned magic_LT;  // ceil(2^32/LT): n/LT = umulhi(n, magic_LT) (exact range: geom_decodable)
  unsigned magic_L;
};

// Geometry of an L x L x LT lattice (host side).
inline Geom make_geom(int L, int LT) {
  Geom g;
  g.L = L;
  g.LT = LT;
  g.V = L * L * LT;
  g.half = g.V / 2;
  g.LLT = L * LT;
  g.magic_LT = (unsigned)((((unsigned long long)1 << 32) + LT - 1) / LT);
  g.magic_L = (unsigned)((((unsigned long long)1 << 32) + L - 1) / L);
  return g;
}

// decode() divides by LT and L with umulhi(n, m), m = ceil(2^32/d).  With
// e = m d - 2^32 (0 <= e < d), umulhi(n, m) = floor(n/d) exactly while
// n e < 2^32: n = q d + r gives n m / 2^32 = q + (r + n e / 2^32) / d, and
// r <= d - 1.  decode() divides lex indices n < V by LT and quotients
// q < L^2 by L; both bounds are checked here (powers of two have e = 0).
inline bool geom_decodable(int L, int LT) {
  const unsigned long long two32 = 1ull << 32;
  const unsigned long long V = (unsigned long long)L * L * LT;
  if (V > (1ull << 26)) return false;
  const unsigned long long eLT = (unsigned long long)make_geom(L, LT).magic_LT * LT - two32;
  const unsigned long long eL = (unsigned long long)make_geom(L, LT).magic_L * L - two32;
  return (V - 1) * eLT < two32 && ((unsigned long long)L * L - 1) * eL < two32;
}

// Site decode: coordinates, parity and lex index of f-site f.
struct Site {
  int x0, x1, x2, P, lex;
};

GNX_DEV unsigned umulhi(unsigned a, unsigned b) { return (unsigned)(((unsigned long long)a * b) >> 32); }

GNX_DEV Site decode(const Geom& g, int f) {
  Site s;
  s.P = f >= g.half ? 1 : 0;
  unsigned h = (unsigned)(f - s.P * g.half);
  unsigned n0 = 2u * h;
  unsigned q = umulhi(n0, g.magic_LT);         // n0 / LT
  int x2e = (int)(n0 - q * (unsigned)g.LT);      // even
  unsigned x0 = umulhi(q, g.magic_L);            // q / L
  s.x0 = (int)x0;
  s.x1 = (int)(q - x0 * (unsigned)g.L);
  s.x2 = x2e + ((s.x0 + s.x1 + s.P) & 1);
  s.lex = (int)(q * (unsigned)g.LT) + s.x2;
  return s;
}

// lex indices of the 6 neighbours
struct Nbr {
  int pp[3], mm[3];
};

GNX_DEV Nbr neighbours_lex(const Geom& g, const Site& s) {
  Nbr n;
  n.pp[2] = (s.x2 == g.LT - 1) ? s.lex - (g.LT - 1) : s.lex + 1;
  n.mm[2] = (s.x2 == 0) ? s.lex + (g.LT - 1) : s.lex - 1;
  n.pp[1] = (s.x1 == g.L - 1) ? s.lex - (g.L - 1) * g.LT : s.lex + g.LT;
  n.mm[1] = (s.x1 == 0) ? s.lex + (g.L - 1) * g.LT : s.lex - g.LT;
  n.pp[0] = (s.x0 == g.L - 1) ? s.lex - (g.L - 1) * g.LLT : s.lex + g.LLT;
  n.mm[0] = (s.x0 == 0) ? s.lex + (g.L - 1) * g.LLT : s.lex - g.LLT;
  return n;
}

// f-index of a lex site of parity P
GNX_DEV int f_of_lex(const Geom& g, int lex, int P) { return (lex >> 1) + P * g.half; }

// f-indices of the 6 neighbours (all of parity 1-P)
struct NbrF {
  int pp[3], mm[3];
};

GNX_DEV NbrF neighbours_f(const Geom& g, const Site& s) {
  Nbr n = neighbours_lex(g, s);
  NbrF r;
  int off = (1 - s.P) * g.half;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    r.pp[d] = (n.pp[d] >> 1) + off;
    r.mm[d] = (n.mm[d] >> 1) + off;
  }
  return r;
}

// The parities and boundary flags the operator needs at a site, in one word,
// so that a kernel decodes a site once, in its load phase, and keeps one
// register for its arithmetic:
//   bit 0: x0 odd          bit 1: x0 + x1 odd
//   bit 2: x2 == LT - 1    bit 3: x2 == 0
//   bit 4: x0 == L - 1     bit 5: x0 == 0
//   bit 6: x1 == L - 1     bit 7: x1 == 0
struct SiteCode {
  unsigned c;
};
enum : unsigned {
  SC_X0ODD = 1u, SC_X01ODD = 2u, SC_X2TOP = 4u, SC_X2BOT = 8u,
  SC_X0TOP = 16u, SC_X0BOT = 32u, SC_X1TOP = 64u, SC_X1BOT = 128u
};

GNX_DEV SiteCode site_code(const Geom& g, const Site& s) {
  unsigned c = (unsigned)(s.x0 & 1) | ((unsigned)((s.x0 + s.x1) & 1) << 1);
  c |= (s.x2 == g.LT - 1 ? SC_X2TOP : 0u) | (s.x2 == 0 ? SC_X2BOT : 0u);
  c |= (s.x0 == g.L - 1 ? SC_X0TOP : 0u) | (s.x0 == 0 ? SC_X0BOT : 0u);
  c |= (s.x1 == g.L - 1 ? SC_X1TOP : 0u) | (s.x1 == 0 ? SC_X1BOT : 0u);
  SiteCode k = {c};
  return k;
}

// Half-phases 0.5*eta (sign folded): hp_pp[d] = 0.5*eta_pp[x,d], hp_mm[d] = 0.5*eta_mm[x,d]
template <typename R>
struct HalfEtaT {
  R pp[3], mm[3];
};
using HalfEta = HalfEtaT<double>;

template <typename R>
GNX_DEV HalfEtaT<R> half_eta_t(SiteCode k) {
  const R h = (R)0.5;
  const R e1 = (k.c & SC_X0ODD) ? -h : h;
  const R e2 = (k.c & SC_X01ODD) ? -h : h;
  HalfEtaT<R> e;
  e.pp[0] = h;  e.mm[0] = h;
  e.pp[1] = e1; e.mm[1] = e1;
  e.pp[2] = (k.c & SC_X2TOP) ? -e2 : e2;
  e.mm[2] = (k.c & SC_X2BOT) ? -e2 : e2;
  return e;
}
GNX_DEV HalfEta half_eta(SiteCode k) { return half_eta_t<double>(k); }
GNX_DEV HalfEta half_eta(const Geom& g, const Site& s) { return half_eta(site_code(g, s)); }

// neighbours and code of f-site f from one decode (a kernel's load phase)
struct SiteGeo {
  NbrF n;
  SiteCode k;
};
GNX_DEV SiteGeo site_geo(const Geom& g, int f) {
  const Site s = decode(g, f);
  SiteGeo r = {neighbours_f(g, s), site_code(g, s)};
  return r;
}

// ---------------------------------------------------------------- complex
GNX_DEV double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
GNX_DEV double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
GNX_DEV double2 rmul(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
GNX_DEV double cnorm2(double2 a) { return a.x * a.x + a.y * a.y; }
// Re(conj(a) * b)
GNX_DEV double cdot_re(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
GNX_DEV double cdot_im(double2 a, double2 b) { return a.x * b.y - a.y * b.x; }


// ---------------------------------------------------------------- reductions
// Deterministic wavefront sum.  Within each 16-lane row, four DPP exchanges
// (quad xor 1, quad xor 2, half-row mirror, row mirror) each add a lane's
// partial sum to its partner's; the partner adds the same two values in the
// other operand order, so both hold the same bits.  The four row sums are
// then read out and added as (r0 + r1) + (r2 + r3): the result is
// wave-uniform.  DPP moves and lane reads are VALU operations, so no step
// waits on the LDS crossbar (ds_bpermute), which __shfl_xor would use.
template <int CTRL>
GNX_DEVONLY double dpp_move(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
GNX_DEVONLY float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
// DPP controls: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror
constexpr int DPP_QXOR1 = 0xB1, DPP_QXOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

GNX_DEVONLY double lane_read(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}
GNX_DEVONLY float lane_read(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <typename T>
GNX_DEVONLY T wave_sum_t(T v) {
  v += dpp_move<DPP_QXOR1>(v);
  v += dpp_move<DPP_QXOR2>(v);
  v += dpp_move<DPP_HALF_MIRROR>(v);
  v += dpp_move<DPP_MIRROR>(v);
  return (lane_read(v, 0) + lane_read(v, 16)) + (lane_read(v, 32) + lane_read(v, 48));
}
GNX_DEVONLY double wave_sum(double v) { return wave_sum_t(v); }
GNX_DEVONLY float wave_sumf(float v) { return wave_sum_t(v); }

// The nw wave sums in lds added in wave order.  For the block sizes the
// kernels use (4, 8 or 16 waves) the count is a constant, so every read
// issues before the first add; a loop over a run-time count reads, waits
// and adds one value per trip (3 to 15 LDS round trips in a row).
template <int NW>
GNX_DEVONLY double sum_waves_n(const double* lds) {
  double w[NW];
#pragma unroll
  for (int k = 0; k < NW; k++) w[k] = lds[k];
  double t = w[0];
#pragma unroll
  for (int k = 1; k < NW; k++) t += w[k];
  return t;
}
GNX_DEVONLY double sum_waves(const double* lds, int nw) {
  switch (nw) {
    case 4: return sum_waves_n<4>(lds);
    case 8: return sum_waves_n<8>(lds);
    case 16: return sum_waves_n<16>(lds);
    default: {
      double t = lds[0];
      for (int w = 1; w < nw; w++) t += lds[w];
      return t;
    }
  }
}

// Deterministic block sum; every thread gets the same bits: the wave sums
// added in wave order.  lds must hold blockDim.x/64 doubles.  Two
// __syncthreads(): the first keeps a previous call's readers ahead of this
// call's writes to lds.
GNX_DEVONLY double block_sum(double v, double* lds) {
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[wid] = v;
  __syncthreads();
  return sum_waves(lds, nw);
}

// block_sum for an lds array that the kernel uses only once: no leading
// barrier (no earlier reader of lds can be overtaken).
GNX_DEVONLY double block_sum_once(double v, double* lds) {
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) lds[wid] = v;
  __syncthreads();
  return sum_waves(lds, nw);
}

// Two deterministic block sums at once (each bit-identical to block_sum of
// its input); lds must hold 2*blockDim.x/64 doubles.  Contains __syncthreads().
GNX_DEVONLY double2 block_sum2(double a, double b, double* lds) {
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    lds[wid] = a;
    lds[nw + wid] = b;
  }
  __syncthreads();
  return make_double2(sum_waves(lds, nw), sum_waves(lds + nw, nw));
}

}  // namespace gnx
