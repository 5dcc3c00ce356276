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
  unsigned magic_LT;  // ceil(2^32/LT): n/LT = umulhi(n, magic_LT) for n*LT < 2^32
  unsigned magic_L;
};

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

// Half-phases 0.5*eta (sign folded): hp_pp[d] = 0.5*eta_pp[x,d], hp_mm[d] = 0.5*eta_mm[x,d]
struct HalfEta {
  double pp[3], mm[3];
};

GNX_DEV HalfEta half_eta(const Geom& g, const Site& s) {
  HalfEta e;
  double e1 = (s.x0 & 1) ? -0.5 : 0.5;
  double e2 = ((s.x0 + s.x1) & 1) ? -0.5 : 0.5;
  e.pp[0] = 0.5; e.mm[0] = 0.5;
  e.pp[1] = e1;  e.mm[1] = e1;
  e.pp[2] = (s.x2 == g.LT - 1) ? -e2 : e2;
  e.mm[2] = (s.x2 == 0) ? -e2 : e2;
  return e;
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
// Deterministic wavefront sum (fixed butterfly order, 64 lanes).
GNX_DEVONLY double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

GNX_DEVONLY float wave_sumf(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic block sum; every thread gets the result. lds must hold
// blockDim.x/64 doubles. Contains __syncthreads().
GNX_DEVONLY double block_sum(double v, double* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double t = (lane < nw) ? lds[lane] : 0.0;
  t = wave_sum(t);
  return t;
}

// Two deterministic block sums at once (each bit-identical to block_sum of
// its input); lds must hold 2*blockDim.x/64 doubles.  Contains __syncthreads().
GNX_DEVONLY double2 block_sum2(double a, double b, double* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if (lane == 0) {
    lds[wid] = a;
    lds[nw + wid] = b;
  }
  __syncthreads();
  const double ta = (lane < nw) ? lds[lane] : 0.0;
  const double tb = (lane < nw) ? lds[nw + lane] : 0.0;
  return make_double2(wave_sum(ta), wave_sum(tb));
}

}  // namespace gnx
