// gnx_kernels.hip -- single-pass kernels of the fermion hot path:
// sigma mass, Dirac operator applications, sigma force, HMC field updates,
// deterministic FP64 reductions for energies and condensate dot products.
// Every kernel is a memory-bound stencil or sweep: one site per thread,
// coalesced double2 loads in f-order, no tables read from memory.
#include "gnx_internal.h"
#include "gnx_solver.h"  // elem()
#include "gnx_tile.h"

namespace gnx {

// ------------------------------------------------------------------ sigma mass
// sigma.c:75-94.  The eight dual-lattice corners below site x, summed in the
// reference's order, then / 8.0.
__global__ void k_sigma_mass(Geom g, const double* __restrict__ sigma, double* __restrict__ sm) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  Site s = decode(g, f);
  sm[f] = sigma_mass_at(g, s, sigma);
}

// sigma += eps*mom (moleculardynamics.c:828-829); the sigma mass is recomputed
// by a second launch (it reads neighbours of the updated field).
__global__ void k_evolve_sigma(int V, double* __restrict__ sigma, const double* __restrict__ mom, double eps) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= V) return;
  sigma[f] += (eps * mom[f]);
}

// mom += force*eps (moleculardynamics.c:843-848)
__global__ void k_evolve_momenta(int V, double* __restrict__ mom, const double* __restrict__ force, double eps) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= V) return;
  mom[f] += (force[f] * eps);
}

// ------------------------------------------------------------------ Dirac
// blockIdx.y: vector of a batch (the same sigma mass for all).  Sites are
// mapped as the CG kernels' XCD slab shape: each XCD sweeps one lex slab of
// both parities, so a site's x0 neighbours are mostly in its own L2 (with
// the plain mapping they sit L*LT/2 sites = L*LT/512 blocks away, on another
// XCD unless that count is a multiple of 8, and every element is fetched by
// three L2s once the vectors outgrow the Infinity Cache).
template <int OP>
__global__ __launch_bounds__(256) void k_dirac(Geom g, const double* __restrict__ sm, const double2* __restrict__ in,
                                               double2* __restrict__ out) {
  static_assert(ENG_BS == 256, "the site map assumes 256-thread blocks");
  const int f = SiteMap<1, true>((int)gridDim.x).f(g, 0);
  if (f < 0) return;
  in += (size_t)blockIdx.y * g.V;
  out += (size_t)blockIdx.y * g.V;
  const Site s = decode(g, f);
  auto ld = [&](int j) { return *elem(in, j); };
  const double smv = *elem(sm, f);
  double2 y;
  if (OP == GNX_OP_M) y = apply_M(g, s, smv, ld);
  else if (OP == GNX_OP_MDAG) y = apply_Mdag(g, s, smv, ld);
  else if (OP == GNX_OP_SUB_M) y = apply_sub_M(g, s, smv, *elem(out, f), ld);
  else y = apply_sub_Mdag(g, s, smv, *elem(out, f), ld);
  *elem(out, f) = y;
}

// ------------------------------------------------------------------ normal operators
// out = M M^dag in (OP2 = GNX_OP_MMDAG) or M^dag M in (GNX_OP_MDAGM) in one
// launch, the intermediate vector kept in LDS: each block computes the first
// operator on its tile and on the tile's one-site halo (gnx_tile.h), into an
// LDS box, then the second operator on the tile from the box.  Every value is
// the one the two single-operator kernels compute (the same operations on the
// same operands: a halo site again by this block), so the result is
// bit-identical to the two launches through a staged vector, which read and
// write the intermediate in memory (80 B/site against 40 B/site here, plus a
// halo of 0.875 first-operator sites per tile site, mostly L2 hits).
template <int OP2, int T0, int T1, int T2>
__global__ __launch_bounds__(ENG_BS) void k_normal(Geom g, int nb, const double* __restrict__ sm,
                                                   const double2* __restrict__ in, double2* __restrict__ out) {
  using TL = FTile<T0, T1, T2>;
  constexpr int R = TL::ROUNDS, SPT = TL::SPT;
  __shared__ double2 box[TL::NBOX];
  in += (size_t)blockIdx.y * g.V;
  out += (size_t)blockIdx.y * g.V;
  int o0, o1, o2;
  ftile_origin<T0, T1, T2>(g, nb, (int)blockIdx.x, &o0, &o1, &o2);
  int fs[R], bx[R];
  SiteCode kc[R];
  double smv[R];
  double2 v[R][7];
#pragma unroll
  for (int i = 0; i < R; i++) {
    // an item past the last one loads the tile's first site and stores nothing
    const int e = i * ENG_BS + (int)threadIdx.x;
    const Site st = ftile_site<T0, T1, T2>(g, o0, o1, o2, e < TL::NEXT ? e : 0, &bx[i]);
    fs[i] = f_of_lex(g, st.lex, st.P);
    kc[i] = site_code(g, st);
    const NbrF n = neighbours_f(g, st);
    const int q[7] = {fs[i], n.pp[0], n.pp[1], n.pp[2], n.mm[0], n.mm[1], n.mm[2]};
    smv[i] = *elem(sm, fs[i]);
#pragma unroll
    for (int k = 0; k < 7; k++) v[i][k] = *elem(in, q[k]);
  }
  end_of_loads();
#pragma unroll
  for (int i = 0; i < R; i++) {
    if (i >= SPT && i * ENG_BS + (int)threadIdx.x >= TL::NEXT) continue;
    const SiteCode st = kc[i];
    box[bx[i]] = OP2 == GNX_OP_MMDAG ? apply_Mdag_vals(st, smv[i], v[i][0], &v[i][1], &v[i][4])
                                     : apply_M_vals(st, smv[i], v[i][0], &v[i][1], &v[i][4]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const SiteCode st = kc[i];
    const int b = bx[i];
    const double2 vpp[3] = {box[b + TL::S0], box[b + TL::S1], box[b + 1]};
    const double2 vmm[3] = {box[b - TL::S0], box[b - TL::S1], box[b - 1]};
    *elem(out, fs[i]) = OP2 == GNX_OP_MMDAG ? apply_M_vals(st, smv[i], box[b], vpp, vmm)
                                            : apply_Mdag_vals(st, smv[i], box[b], vpp, vmm);
  }
}

// whether the one-launch normal operator's 4 x 8 x 16 tiles divide g
bool normal_fused_fits(const Geom& g) { return g.L % 4 == 0 && g.L % 8 == 0 && g.LT % 16 == 0; }

// One application: the one launch where it is faster than two (host-launched
// at 16^3-48^3 4.3-4.8 against 8.3 us, 64^3 6.8 against 8.3; at 96^3, where
// the stencils are bound by bytes, its halo costs more: 17.6 against 15.2).
// A batch always takes it (the two launches stage a single vector).
static bool normal_one_launch(const Geom& g, int nvec) {
  return normal_fused_fits(g) && (nvec > 1 || g.V <= 300000);
}

// ------------------------------------------------------------------ sigma force
// moleculardynamics.c:606-620 (fermion part, flavour-outer, corner order
// {x, x+0+1, x+0+2, x+1+2, x+0, x+1, x+2, x+0+1+2}) then :576-581.
// accumulate: add onto the incoming force instead of starting from 0 (the
// reference's AddForceFromFermionSectorRespectToSigma adds to a caller-zeroed
// force); sigma_term: finish with AddForceFromSigmaSector
__global__ void k_sigma_force(Geom g, const double* __restrict__ sigma, double beta_s, int nflav,
                              const double2* __restrict__ X0, const double2* __restrict__ Y0,
                              const double2* __restrict__ X1, const double2* __restrict__ Y1,
                              double* __restrict__ force, int accumulate, int sigma_term) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.V) return;
  Site s = decode(g, f);
  double F = accumulate ? force[f] : 0.0;
  F = force_flavour(g, s, X0, Y0, F);
  F = force_flavour(g, s, X1, Y1, F);
  if (sigma_term) F -= (nflav * beta_s / 2.0) * sigma[f];
  force[f] = F;
}

// ------------------------------------------------------------------ reductions
// Deterministic two-pass reduction: per-block partials in a fixed tree, then a
// single block sums the partials in index order.  out[k] for k < nval.
// kind selects the summand:
//   RED_NORM2    |a|^2                      (complex a)
//   RED_DOT_RE   Re(conj a * b)
//   RED_DOT_IM   Im(conj a * b)
//   RED_RSQ      a_r^2 (real field)
//   RED_SUM      a_r (real field)
enum RedKind { RED_NORM2 = 0, RED_DOT_RE = 1, RED_DOT_IM = 2, RED_RSQ = 3, RED_SUM = 4 };

template <int BS>
__global__ __launch_bounds__(BS) void k_reduce_partial(int n, int kind, const double2* __restrict__ a,
                                                        const double2* __restrict__ b,
                                                        const double* __restrict__ ar,
                                                        double* __restrict__ part) {
  __shared__ double lds[BS / 64];
  double acc = 0.0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    switch (kind) {
      case RED_NORM2: acc += cnorm2(a[i]); break;
      case RED_DOT_RE: acc += cdot_re(a[i], b[i]); break;
      case RED_DOT_IM: acc += cdot_im(a[i], b[i]); break;
      case RED_RSQ: acc += ar[i] * ar[i]; break;
      default: acc += ar[i]; break;
    }
  }
  double s = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_reduce_final(int nparts, const double* __restrict__ part,
                                                      double scale, double* __restrict__ out) {
  __shared__ double lds[BS / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += BS) acc += part[i];
  double s = block_sum(acc, lds);
  if (threadIdx.x == 0) *out = s * scale;
}

// ------------------------------------------------------------------ batched trace terms
// Trace term of the condensate for a batch of noise vectors
// (condensate.c:140-145 for each vector): y = M^dag X_k at the site (gather
// form, reference add order) and the block's partial sum of Re(conj(v_k) y);
// blockIdx.y = k.  part is [n][gridDim.x].
template <int BS>
__global__ __launch_bounds__(BS) void k_mdag_dot(Geom g, const double* __restrict__ sm,
                                                  const double2* __restrict__ X, const double2* __restrict__ v,
                                                  double* __restrict__ part) {
  __shared__ double lds[BS / 64];
  const size_t off = (size_t)blockIdx.y * g.V;
  const int f = blockIdx.x * BS + threadIdx.x;
  double acc = 0.0;
  if (f < g.V) {
    const double2* x = X + off;
    const double2 y = apply_Mdag(g, decode(g, f), sm[f], [&](int q) { return x[q]; });
    acc = cdot_re(v[off + f], y);
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// out[k] = scale * (sum of row k of part), one block per row, fixed order
template <int BS>
__global__ __launch_bounds__(BS) void k_reduce_rows(int nb, const double* __restrict__ part, double scale,
                                                     double* __restrict__ out) {
  __shared__ double lds[BS / 64];
  const double* p = part + (size_t)blockIdx.x * nb;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += BS) acc += p[i];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = s * scale;
}

// ------------------------------------------------------------------ elementwise
__global__ void k_axpy_real(int V, double alpha, const double2* __restrict__ xv, double2* __restrict__ y) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= V) return;
  y[f] = cadd(y[f], rmul(alpha, xv[f]));
}

// y = x * sm (complex times real field, condensate.c:221,225)
__global__ void k_scale_by_field(int V, const double* __restrict__ sm, const double2* __restrict__ xv,
                                 double2* __restrict__ y) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= V) return;
  double2 v = xv[f];
  double m = sm[f];
  y[f] = make_double2(v.x * m, v.y * m);
}

// ------------------------------------------------------------------ host launchers
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

hipError_t launch_sigma_mass(const Geom& g, const double* sigma, double* sm, hipStream_t st) {
  hipLaunchKernelGGL(k_sigma_mass, dim3(nblk(g.V, 256)), dim3(256), 0, st, g, sigma, sm);
  return hipGetLastError();
}

hipError_t launch_evolve_sigma(const Geom& g, double* sigma, double* sm, const double* mom, double eps,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_evolve_sigma, dim3(nblk(g.V, 256)), dim3(256), 0, st, g.V, sigma, mom, eps);
  hipLaunchKernelGGL(k_sigma_mass, dim3(nblk(g.V, 256)), dim3(256), 0, st, g, (const double*)sigma, sm);
  return hipGetLastError();
}

hipError_t launch_evolve_momenta(const Geom& g, double* mom, const double* force, double eps, hipStream_t st) {
  hipLaunchKernelGGL(k_evolve_momenta, dim3(nblk(g.V, 256)), dim3(256), 0, st, g.V, mom, force, eps);
  return hipGetLastError();
}

hipError_t launch_dirac(const Geom& g, int op, const double* sm, const double* in, double* out, double* tmp,
                        hipStream_t st, int nvec) {
  const double2* vin = (const double2*)in;
  double2* vout = (double2*)out;
  dim3 grid(nblk(g.V, 256), nvec), blk(256);
  if ((op == GNX_OP_MMDAG || op == GNX_OP_MDAGM) && normal_one_launch(g, nvec)) {
    // one launch, the intermediate in LDS
    const int nb = g.V / (4 * 8 * 16);
    if (op == GNX_OP_MMDAG)
      hipLaunchKernelGGL((k_normal<GNX_OP_MMDAG, 4, 8, 16>), dim3(nb, nvec), blk, 0, st, g, nb, sm, vin, vout);
    else
      hipLaunchKernelGGL((k_normal<GNX_OP_MDAGM, 4, 8, 16>), dim3(nb, nvec), blk, 0, st, g, nb, sm, vin, vout);
    return hipGetLastError();
  }
  // otherwise the two-operator products stage one vector in tmp
  if (nvec != 1 && (op == GNX_OP_MMDAG || op == GNX_OP_MDAGM)) return hipErrorInvalidValue;
  switch (op) {
    case GNX_OP_M:
      hipLaunchKernelGGL(k_dirac<GNX_OP_M>, grid, blk, 0, st, g, sm, vin, vout);
      break;
    case GNX_OP_MDAG:
      hipLaunchKernelGGL(k_dirac<GNX_OP_MDAG>, grid, blk, 0, st, g, sm, vin, vout);
      break;
    case GNX_OP_SUB_M:
      hipLaunchKernelGGL(k_dirac<GNX_OP_SUB_M>, grid, blk, 0, st, g, sm, vin, vout);
      break;
    case GNX_OP_SUB_MDAG:
      hipLaunchKernelGGL(k_dirac<GNX_OP_SUB_MDAG>, grid, blk, 0, st, g, sm, vin, vout);
      break;
    case GNX_OP_MMDAG:
      hipLaunchKernelGGL(k_dirac<GNX_OP_MDAG>, grid, blk, 0, st, g, sm, vin, (double2*)tmp);
      hipLaunchKernelGGL(k_dirac<GNX_OP_M>, grid, blk, 0, st, g, sm, (const double2*)tmp, vout);
      break;
    case GNX_OP_MDAGM:
      hipLaunchKernelGGL(k_dirac<GNX_OP_M>, grid, blk, 0, st, g, sm, vin, (double2*)tmp);
      hipLaunchKernelGGL(k_dirac<GNX_OP_MDAG>, grid, blk, 0, st, g, sm, (const double2*)tmp, vout);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sigma_force(const Geom& g, const double* sigma, double beta_s, int nflav, const double* X0,
                              const double* Y0, const double* X1, const double* Y1, double* force,
                              hipStream_t st, int accumulate, int sigma_term) {
  hipLaunchKernelGGL(k_sigma_force, dim3(nblk(g.V, 256)), dim3(256), 0, st, g, sigma, beta_s, nflav,
                     (const double2*)X0, (const double2*)Y0, (const double2*)X1, (const double2*)Y1, force,
                     accumulate, sigma_term);
  return hipGetLastError();
}

hipError_t launch_reduce(int n, int kind, const void* a, const void* b, double scale, double* part, int nparts,
                         double* out, hipStream_t st) {
  int g = nblk(n, 256);
  if (g > nparts) g = nparts;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_reduce_partial<256>, dim3(g), dim3(256), 0, st, n, kind, (const double2*)a,
                     (const double2*)b, (const double*)a, part);
  hipLaunchKernelGGL(k_reduce_final<256>, dim3(1), dim3(256), 0, st, g, (const double*)part, scale, out);
  return hipGetLastError();
}

hipError_t launch_axpy_real(int V, double alpha, const double* xv, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_real, dim3(nblk(V, 256)), dim3(256), 0, st, V, alpha, (const double2*)xv, (double2*)y);
  return hipGetLastError();
}

hipError_t launch_scale_by_field(int V, const double* sm, const double* xv, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_by_field, dim3(nblk(V, 256)), dim3(256), 0, st, V, sm, (const double2*)xv,
                     (double2*)y);
  return hipGetLastError();
}

// part must hold n * ceil(V/256) doubles, out n doubles
hipError_t launch_mdag_dot_batch(const Geom& g, const double* sm, const double* X, const double* v, int n,
                                 double* part, double* out, hipStream_t st) {
  constexpr int BS = 256;
  const int nb = (g.V + BS - 1) / BS;
  k_mdag_dot<BS><<<dim3(nb, n), BS, 0, st>>>(g, sm, (const double2*)X, (const double2*)v, part);
  k_reduce_rows<BS><<<n, BS, 0, st>>>(nb, part, 1.0 / g.V, out);
  return hipGetLastError();
}

}  // namespace gnx
