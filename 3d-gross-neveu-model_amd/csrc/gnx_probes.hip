// gnx_probes.hip -- measurement probes behind the design of the CG schedule
// (lib/libgnx_probes.so; the benchmark's `probes` leg).  Not part of the
// solver library: no product code calls these.
//
//   gnx_probe_graph_nodes   cost of one kernel node of a replayed hipGraph,
//                           as a function of the bytes the node streams: an
//                           empty kernel, and streaming kernels moving the
//                           bytes of a Dirac application (2 vectors + a
//                           real field in, one vector out: 40 B/site) and of the
//                           one-reduction CG's second kernel (152 B/site),
//                           each node reading the previous node's output;
//                           and the 152 B/site node's reads alone (88 B/site,
//                           one partial sum per block) and writes alone
//                           (64 B/site).
//   gnx_probe_grid_barrier  cost of a grid-wide barrier inside one
//                           persistent kernel (atomic arrival counter and a
//                           generation word, device-scope release before the
//                           arrival and acquire after the wait), with each
//                           block writing a given number of bytes between
//                           barriers.  Every wait is bounded: a block that
//                           spins past the limit raises a flag, and every
//                           block leaves the loop at the next barrier once
//                           the flag is up, so the kernel always ends.
//   gnx_probe_halo          the communication of a persistent CG iteration
//   gnx_probe_halo_tag      with no grid barrier: neighbour-face exchanges
//                           through per-block flags (or round-tagged data
//                           words) and an all-block gather of partial sums,
//                           one block per 16 x 8 x 8 tile, every wait
//                           bounded, every received value checked.
//
// C ABI (host functions; all return hipError_t as int, 0 on success):
//   int gnx_probe_graph_nodes(int kind, long sites, int nodes, int reps, double* us_per_node);
//   int gnx_probe_empty_nodes(int blocks, int threads, int nodes, int reps, double* us_per_node);
//   int gnx_probe_width_nodes(int width, long bytes, int nodes, int reps, double* us_per_node);
//   int gnx_probe_gather_nodes(int width, long elems, int nodes, int reps, double* us_per_node);
//   int gnx_probe_grid_barrier(int mode, int blocks, int iters, long bytes_per_block,
//                              double* us_per_barrier, int* max_blocks, int* failed);
//   mode 0: every poll an acquire load; mode 1: relaxed polls, one acquire fence
//   int gnx_probe_halo(int mode, int nt0, int nt1, int nt2, int iters, long lds_pad,
//                      double* us_per_round, int* failed, int* stale, int* per_cu);
//   int gnx_probe_halo_tag(int red, int nt0, int nt1, int nt2, int iters, long lds_pad,
//                          double* us_per_round, int* failed, int* stale, int* per_cu);
//   (lds_pad: dynamic LDS per block, which sets how many blocks share a CU)
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gnx_probe {

constexpr int kBS = 256;

__global__ __launch_bounds__(kBS) void k_empty(int* sink) {
  if (sink && threadIdx.x == 0 && blockIdx.x == 0x7fffffff) *sink = 1;  // never true: keeps the kernel
}

// the same for any block size up to 1024
__global__ __launch_bounds__(1024) void k_empty_any(int* sink) {
  if (sink && threadIdx.x == 0 && blockIdx.x == 0x7fffffff) *sink = 1;
}

// 40 B/site: out = a + w*b (two vectors and a field in, one vector out)
__global__ __launch_bounds__(kBS) void k_stream40(long n, const double2* __restrict__ a, const double2* __restrict__ b,
                                                  const double* __restrict__ w, double2* __restrict__ out) {
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const double2 x = a[i], y = b[i];
  const double s = w[i];
  out[i] = make_double2(x.x + s * y.x, x.y + s * y.y);
}

// 152 B/site: five vectors and a field in, four vectors out (the byte count
// of the one-reduction CG's second kernel at one site)
__global__ __launch_bounds__(kBS) void k_stream152(long n, const double2* __restrict__ v0, const double* __restrict__ w,
                                                   double2* __restrict__ v1, double2* __restrict__ v2,
                                                   double2* __restrict__ v3, double2* __restrict__ v4) {
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const double2 q = v0[i], a = v1[i], b = v2[i], c = v3[i], d = v4[i];
  const double s = w[i];
  v1[i] = make_double2(a.x + s * q.x, a.y + s * q.y);
  v2[i] = make_double2(b.x - s * a.x, b.y - s * a.y);
  v3[i] = make_double2(c.x + s * b.x, c.y + s * b.y);
  v4[i] = make_double2(d.x - s * c.x, d.y - s * c.y);
}

// 88 B/site read, nothing written but one partial sum per block: the
// reads of k_stream152 alone
__global__ __launch_bounds__(kBS) void k_read88(long n, const double2* __restrict__ v0, const double* __restrict__ w,
                                                const double2* __restrict__ v1, const double2* __restrict__ v2,
                                                const double2* __restrict__ v3, const double2* __restrict__ v4,
                                                double* __restrict__ part) {
  __shared__ double lds[kBS / 64];
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
    const double2 q = v0[i], a = v1[i], b = v2[i], c = v3[i], d = v4[i];
    acc = w[i] * (q.x + a.y) + b.x * c.y + d.x * d.y;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// 64 B/site written, nothing read: the writes of k_stream152 alone
__global__ __launch_bounds__(kBS) void k_write64(long n, double s, double2* __restrict__ v1, double2* __restrict__ v2,
                                                 double2* __restrict__ v3, double2* __restrict__ v4) {
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  const double t = s + (double)i;
  v1[i] = make_double2(t, s);
  v2[i] = make_double2(s, t);
  v3[i] = make_double2(t, t);
  v4[i] = make_double2(s, s);
}

// The same bytes through lanes of W bytes: three arrays read and one
// written, each of `bytes` bytes, element type T (4, 8 or 16 bytes) -- the
// load/store instruction count per byte scales as 1/W.
template <typename T>
__global__ __launch_bounds__(kBS) void k_width(long n, const T* __restrict__ a, const T* __restrict__ b,
                                              const T* __restrict__ c, T* __restrict__ out) {
  const long i = (long)blockIdx.x * kBS + threadIdx.x;
  if (i >= n) return;
  T x = a[i], y = b[i], z = c[i];
  unsigned* px = (unsigned*)&x;
  const unsigned* py = (const unsigned*)&y;
  const unsigned* pz = (const unsigned*)&z;
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) px[k] ^= py[k] + pz[k];
  out[i] = x;
}

// A seven-point gather over two 4-byte-element arrays laid out like one
// parity of a 64^3 FP16 vector (rows of 32 elements, planes of 2048), one
// element per lane (W = 4) or four consecutive elements per lane loaded as
// 16 bytes (W = 16; the +-1 neighbours are then the same 16 bytes shifted
// by one element, read as two aligned 16-byte loads) -- the same bytes and
// the same neighbour pattern, 1/4 of the load instructions.  One 4-byte
// output element per site.
template <int W>
__global__ __launch_bounds__(kBS) void k_gather7(long n, const unsigned* __restrict__ a, const unsigned* __restrict__ b,
                                                unsigned* __restrict__ out) {
  constexpr long ROW = 32, PLANE = 2048;
  const long t = (long)blockIdx.x * kBS + threadIdx.x;
  if (W == 4) {
    const long i = t;
    if (i >= n) return;
    auto at = [&](const unsigned* v, long j) { return v[(j + n) % n]; };
    unsigned s = 0;
    const long o[7] = {0, 1, -1, ROW, -ROW, PLANE, -PLANE};
#pragma unroll
    for (int k = 0; k < 7; k++) s += at(a, i + o[k]) ^ at(b, i + o[k]);
    out[i] = s;
  } else {
    const long i = 4 * t;  // four consecutive elements
    if (i >= n) return;
    const uint4* a4 = (const uint4*)a;
    const uint4* b4 = (const uint4*)b;
    const long q = t, nq = n / 4;
    auto ld = [&](const uint4* v, long j) { return v[(j + nq) % nq]; };
    uint4 acc = make_uint4(0, 0, 0, 0);
    auto add = [&](uint4 x, uint4 y) {
      acc.x += x.x ^ y.x; acc.y += x.y ^ y.y; acc.z += x.z ^ y.z; acc.w += x.w ^ y.w;
    };
    const uint4 a0 = ld(a4, q), b0 = ld(b4, q), ap = ld(a4, q + 1), bp = ld(b4, q + 1), am = ld(a4, q - 1),
                bm = ld(b4, q - 1);
    add(a0, b0);
    add(make_uint4(a0.y, a0.z, a0.w, ap.x), make_uint4(b0.y, b0.z, b0.w, bp.x));  // +1
    add(make_uint4(am.w, a0.x, a0.y, a0.z), make_uint4(bm.w, b0.x, b0.y, b0.z));  // -1
    const long o[4] = {ROW / 4, -ROW / 4, PLANE / 4, -PLANE / 4};
#pragma unroll
    for (int k = 0; k < 4; k++) add(ld(a4, q + o[k]), ld(b4, q + o[k]));
    ((uint4*)out)[q] = acc;
  }
}

// The same gather with each block's tile staged in LDS first: a tile of 2
// planes x 16 rows x 32 elements (1024 sites, 4 per thread) and its halo, a
// box of 4 planes x 18 rows x 32 (2304 elements per array), loaded with
// whole 128-byte rows, then the seven neighbours of every site read from
// the box.  Flat index arithmetic wraps like k_gather7's.
__global__ __launch_bounds__(kBS) void k_gather7_lds(long n, const unsigned* __restrict__ a,
                                                    const unsigned* __restrict__ b, unsigned* __restrict__ out) {
  constexpr int ROW = 32, PLANE = 2048, BR = 18, BP = 4, SLICE = BR * ROW;  // 576
  __shared__ unsigned box[2][BP * SLICE];
  const long tile0 = (long)(blockIdx.x / 4) * 2 * PLANE + (long)(blockIdx.x % 4) * 16 * ROW;
  // box element (dp, dr, c) <-> flat tile0 + (dp - 1) PLANE + (dr - 1) ROW + c
  for (int e = threadIdx.x; e < BP * SLICE; e += kBS) {
    const int dp = e / SLICE, rest = e % SLICE;
    const long j = ((tile0 + (long)(dp - 1) * PLANE + (long)(rest - ROW)) % n + n) % n;
    box[0][e] = a[j];
    box[1][e] = b[j];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int st = k * kBS + (int)threadIdx.x;  // tile site: planes 1-2, rows 1-16 of the box
    const int dp = 1 + st / 512, dr = 1 + (st % 512) / 32, c = st % 32;
    const int e = dp * SLICE + dr * ROW + c;
    const int o[7] = {0, 1, -1, ROW, -ROW, SLICE, -SLICE};
    unsigned sum = 0;
#pragma unroll
    for (int q = 0; q < 7; q++) sum += box[0][e + o[q]] ^ box[1][e + o[q]];
    const long i = tile0 + (long)(dp - 1) * PLANE + (long)(dr - 1) * ROW + c;
    if (i < n) out[i] = sum;
  }
}

struct BarrierState {
  unsigned count;
  unsigned gen;
  int failed;
  int pad;
};

constexpr long kSpinLimit = 1L << 22;

// mode 0 (POLL_ACQUIRE): every poll of the generation word is a
// device-scope acquire load (each one invalidates the polling XCD's L2);
// mode 1: relaxed polls and one acquire fence after the wait.  The arrival
// is a device-scope release either way (the L2 write-back of the block's
// data).
template <bool POLL_ACQUIRE>
__device__ void grid_barrier(BarrierState* bs, unsigned nblocks) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = __hip_atomic_load(&bs->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(&bs->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblocks - 1) {
      __hip_atomic_store(&bs->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bs->gen, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(&bs->gen, POLL_ACQUIRE ? __ATOMIC_ACQUIRE : __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > kSpinLimit) {
          __hip_atomic_store(&bs->failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (!POLL_ACQUIRE) __atomic_thread_fence(__ATOMIC_ACQUIRE);
    }
  }
  __syncthreads();
}

// each block writes its bytes (vector stores over its own slice), then waits
// at the barrier; iters rounds.  Once a wait has timed out, every block
// leaves at its next round.
template <bool POLL_ACQUIRE>
__global__ __launch_bounds__(kBS) void k_barrier_loop(BarrierState* bs, int iters, double2* slice, long elems_per_block) {
  __shared__ int stop;
  double2* mine = slice + (long)blockIdx.x * elems_per_block;
  for (int it = 0; it < iters; it++) {
    for (long i = threadIdx.x; i < elems_per_block; i += kBS) mine[i] = make_double2((double)it, (double)i);
    grid_barrier<POLL_ACQUIRE>(bs, gridDim.x);
    if (threadIdx.x == 0) stop = __hip_atomic_load(&bs->failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (stop) return;
  }
}

// ---- halo-exchange probe: the communication of a persistent CG iteration --
// One block per tile of a periodic nt0 x nt1 x nt2 grid of 16 x 8 x 8-site
// tiles (1024 threads, one per site).  Per round k, as a persistent CG on
// M M^dag would do it: publish the tile's six faces of r (640 complex
// values, two 8-byte device-scope stores each) and raise rflag = k; wait
// for the six neighbours' rflag >= k and read their faces; publish the q
// faces and a partial pair and raise qflag = k; wait for the neighbours'
// qflag, read their q faces, then wait for every block's qflag and read all
// partials.  Every value carries (k, block, index), and the reader counts
// values that are not the ones of round k (stale).  Double-buffered by the
// parity of k.  Flags are only written by their block and only grow.
//   mode bit 0: the writer's ordering before a flag is a device-scope
//               release fence (L2 write-back and wait) instead of a plain
//               wait for its stores (s_waitcnt vmcnt(0));
//   mode bit 1: the reader issues a device-scope acquire fence after a
//               wait (L2 invalidate) before it reads the faces.
// Every wait is bounded as in the barrier probe.
constexpr int kHB = 1024;
constexpr int kFace = 640;  // face sites of a 16 x 8 x 8 tile: 2 (64 + 128 + 128)
constexpr long kHaloSpin = 1L << 20;

struct HaloArgs {
  unsigned* rflag;           // [blocks]
  unsigned* qflag;           // [blocks]
  unsigned long long* rface;  // [2][blocks][kFace][2] bit patterns of (re, im)
  unsigned long long* qface;
  unsigned long long* part;   // [2][blocks][2]
  int* failed;
  int* stale;
  int nt0, nt1, nt2;
};

// face of halo slot i: 0/1 = x0 low/high (64 sites), 2/3 = x1 (128), 4/5 = x2 (128)
__device__ __forceinline__ int face_of(int i) { return i < 128 ? (i >> 6) : 2 + (i - 128) / 128; }
__device__ __forceinline__ int face_off(int f) { return f < 2 ? 64 * f : 128 + 128 * (f - 2); }

__device__ __forceinline__ void st_dev(unsigned long long* p, double v) {
  __hip_atomic_store(p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_dev(const unsigned long long* p) {
  return __longlong_as_double((long long)__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

template <bool WB>
__device__ __forceinline__ void publish_order() {
  if (WB)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  else
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// bounded wait for *f >= k; false once this or another block gave up
__device__ bool wait_ge(const unsigned* f, unsigned k, int* failed) {
  long spins = 0;
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
    if (++spins > kHaloSpin) {
      __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

__device__ __forceinline__ double tagv(int k, int b, int i, int kind) {
  return (double)k * 1e7 + (double)b * 2000.0 + (double)i + (kind ? 0.5 : 0.0);
}

template <bool WB, bool INV>
__global__ __launch_bounds__(kHB) void k_halo_loop(HaloArgs a, int iters) {
  __shared__ double2 halo[2 * kFace];
  __shared__ int stop;
  __shared__ int bad;
  const int b = blockIdx.x, tid = threadIdx.x, nblk = gridDim.x;
  const int c0 = b / (a.nt1 * a.nt2), c1 = (b / a.nt2) % a.nt1, c2 = b % a.nt2;
  auto id = [&](int x, int y, int z) {
    x = (x + a.nt0) % a.nt0;
    y = (y + a.nt1) % a.nt1;
    z = (z + a.nt2) % a.nt2;
    return (x * a.nt1 + y) * a.nt2 + z;
  };
  // neighbour across face f: the block whose face f^1 touches this tile
  const int nb[6] = {id(c0 - 1, c1, c2), id(c0 + 1, c1, c2), id(c0, c1 - 1, c2),
                     id(c0, c1 + 1, c2), id(c0, c1, c2 - 1), id(c0, c1, c2 + 1)};
  if (tid == 0) bad = 0;
  int mybad = 0;
  for (int k = 1; k <= iters; k++) {
    const int s = k & 1;
    for (int i = tid; i < kFace; i += kHB) {
      unsigned long long* d = a.rface + (((size_t)s * nblk + b) * kFace + i) * 2;
      st_dev(d, tagv(k, b, i, 0));
      st_dev(d + 1, -tagv(k, b, i, 0));
    }
    publish_order<WB>();
    __syncthreads();
    if (tid == 0) {
      stop = 0;
      __hip_atomic_store(a.rflag + b, (unsigned)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (tid < 6 && !wait_ge(a.rflag + nb[tid], (unsigned)k, a.failed)) stop = 1;
    __syncthreads();
    if (stop) break;
    if (INV) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int i = tid; i < kFace; i += kHB) {
      const int f = face_of(i), src = nb[f], j = face_off(f ^ 1) + (i - face_off(f));
      const unsigned long long* p = a.rface + (((size_t)s * nblk + src) * kFace + j) * 2;
      const double re = ld_dev(p), im = ld_dev(p + 1);
      if (re != tagv(k, src, j, 0) || im != -tagv(k, src, j, 0)) mybad++;
      halo[i] = make_double2(re, im);
    }
    // q faces and the partial pair, behind one flag
    for (int i = tid; i < kFace; i += kHB) {
      unsigned long long* d = a.qface + (((size_t)s * nblk + b) * kFace + i) * 2;
      st_dev(d, tagv(k, b, i, 1) + halo[i].x * 0.0);
      st_dev(d + 1, -tagv(k, b, i, 1));
    }
    if (tid == 0) {
      st_dev(a.part + ((size_t)s * nblk + b) * 2, tagv(k, b, 0, 1));
      st_dev(a.part + ((size_t)s * nblk + b) * 2 + 1, tagv(k, b, 1, 1));
    }
    publish_order<WB>();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(a.qflag + b, (unsigned)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid < 6 && !wait_ge(a.qflag + nb[tid], (unsigned)k, a.failed)) stop = 1;
    __syncthreads();
    if (stop) break;
    if (INV) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int i = tid; i < kFace; i += kHB) {
      const int f = face_of(i), src = nb[f], j = face_off(f ^ 1) + (i - face_off(f));
      const unsigned long long* p = a.qface + (((size_t)s * nblk + src) * kFace + j) * 2;
      const double re = ld_dev(p), im = ld_dev(p + 1);
      if (re != tagv(k, src, j, 1) || im != -tagv(k, src, j, 1)) mybad++;
      halo[kFace + i] = make_double2(re, im);
    }
    if (tid < nblk && !wait_ge(a.qflag + tid, (unsigned)k, a.failed)) stop = 1;
    __syncthreads();
    if (stop) break;
    if (INV) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (tid < nblk) {
      const unsigned long long* p = a.part + ((size_t)s * nblk + tid) * 2;
      if (ld_dev(p) != tagv(k, tid, 0, 1) || ld_dev(p + 1) != tagv(k, tid, 1, 1)) mybad++;
    }
    __syncthreads();
  }
  if (mybad) atomicAdd(&bad, mybad);
  __syncthreads();
  if (tid == 0 && bad) atomicAdd(a.stale, bad);
}

// ---- tagged-word exchange: no flags and no fences --------------------------
// The same rounds as k_halo_loop, with every double sent as two 64-bit words
// (tag << 32 | high half, tag << 32 | low half), tag = round.  A 64-bit
// device-scope store is single-copy atomic, so a reader that polls a word
// until it carries the round's tag has that word's half of the value: no
// flag, no wait for store completion and no separate data load.  The
// partial pairs: RED 0 none; RED 1 every block polls all blocks' pairs; RED
// 2 two levels, the first block of each group of 16 polls its group's
// pairs and publishes a group pair, and every block polls the group pairs.
constexpr int kGroup = 16;

__device__ __forceinline__ void st_tag(unsigned long long* p, unsigned tag, double v) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned long long t = (unsigned long long)tag << 32;
  __hip_atomic_store(p, t | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, t | (bits & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// bounded poll of n tagged doubles at p (2n words) until all carry tag
template <int N>
__device__ bool ld_tag(const unsigned long long* p, unsigned tag, double* v, int* failed) {
  unsigned long long w[2 * N];
  long spins = 0;
  while (true) {
#pragma unroll
    for (int j = 0; j < 2 * N; j++) w[j] = __hip_atomic_load(p + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 2 * N; j++) ok = ok && (unsigned)(w[j] >> 32) == tag;
    if (ok) break;
    if (++spins > kHaloSpin) {
      __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int j = 0; j < N; j++) v[j] = __hiloint2double((int)(unsigned)w[2 * j], (int)(unsigned)w[2 * j + 1]);
  return true;
}

struct TagArgs {
  unsigned long long* rface;  // [2][blocks][kFace][4]
  unsigned long long* qface;
  unsigned long long* part;   // [2][blocks][4]
  unsigned long long* gpart;  // [2][groups][4]
  int* failed;
  int* stale;
  int nt0, nt1, nt2;
};

template <int RED>
__global__ __launch_bounds__(kHB) void k_halo_tag(TagArgs a, int iters) {
  __shared__ double2 halo[2 * kFace];
  __shared__ int stop;
  __shared__ int bad;
  const int b = blockIdx.x, tid = threadIdx.x, nblk = gridDim.x;
  const int ngroups = (nblk + kGroup - 1) / kGroup;
  const int c0 = b / (a.nt1 * a.nt2), c1 = (b / a.nt2) % a.nt1, c2 = b % a.nt2;
  auto id = [&](int x, int y, int z) {
    x = (x + a.nt0) % a.nt0;
    y = (y + a.nt1) % a.nt1;
    z = (z + a.nt2) % a.nt2;
    return (x * a.nt1 + y) * a.nt2 + z;
  };
  const int nb[6] = {id(c0 - 1, c1, c2), id(c0 + 1, c1, c2), id(c0, c1 - 1, c2),
                     id(c0, c1 + 1, c2), id(c0, c1, c2 - 1), id(c0, c1, c2 + 1)};
  if (tid == 0) {
    bad = 0;
    stop = 0;
  }
  __syncthreads();
  int mybad = 0;
  for (int k = 1; k <= iters; k++) {
    const int s = k & 1;
    const unsigned tag = (unsigned)k;
    for (int i = tid; i < kFace; i += kHB) {
      unsigned long long* d = a.rface + (((size_t)s * nblk + b) * kFace + i) * 4;
      st_tag(d, tag, tagv(k, b, i, 0));
      st_tag(d + 2, tag, -tagv(k, b, i, 0));
    }
    for (int i = tid; i < kFace; i += kHB) {
      const int f = face_of(i), src = nb[f], j = face_off(f ^ 1) + (i - face_off(f));
      double v[2];
      if (!ld_tag<2>(a.rface + (((size_t)s * nblk + src) * kFace + j) * 4, tag, v, a.failed)) {
        stop = 1;
        break;
      }
      if (v[0] != tagv(k, src, j, 0) || v[1] != -tagv(k, src, j, 0)) mybad++;
      halo[i] = make_double2(v[0], v[1]);
    }
    __syncthreads();
    if (stop) break;
    for (int i = tid; i < kFace; i += kHB) {
      unsigned long long* d = a.qface + (((size_t)s * nblk + b) * kFace + i) * 4;
      st_tag(d, tag, tagv(k, b, i, 1) + halo[i].x * 0.0);
      st_tag(d + 2, tag, -tagv(k, b, i, 1));
    }
    if (RED && tid == 0) {
      st_tag(a.part + ((size_t)s * nblk + b) * 4, tag, tagv(k, b, 0, 1));
      st_tag(a.part + ((size_t)s * nblk + b) * 4 + 2, tag, tagv(k, b, 1, 1));
    }
    for (int i = tid; i < kFace; i += kHB) {
      const int f = face_of(i), src = nb[f], j = face_off(f ^ 1) + (i - face_off(f));
      double v[2];
      if (!ld_tag<2>(a.qface + (((size_t)s * nblk + src) * kFace + j) * 4, tag, v, a.failed)) {
        stop = 1;
        break;
      }
      if (v[0] != tagv(k, src, j, 1) || v[1] != -tagv(k, src, j, 1)) mybad++;
      halo[kFace + i] = make_double2(v[0], v[1]);
    }
    if (RED == 1 && tid < nblk) {
      double v[2];
      if (!ld_tag<2>(a.part + ((size_t)s * nblk + tid) * 4, tag, v, a.failed))
        stop = 1;
      else if (v[0] != tagv(k, tid, 0, 1) || v[1] != tagv(k, tid, 1, 1))
        mybad++;
    }
    if (RED == 2) {
      const int g = b / kGroup;
      if (b % kGroup == 0) {
        // group leader: its members' pairs, then the group pair
        __shared__ double gs[2];
        if (tid == 0) gs[0] = gs[1] = 0.0;
        __syncthreads();
        const int m = g * kGroup + tid;
        if (tid < kGroup && m < nblk) {
          double v[2];
          if (!ld_tag<2>(a.part + ((size_t)s * nblk + m) * 4, tag, v, a.failed))
            stop = 1;
          else if (v[0] != tagv(k, m, 0, 1) || v[1] != tagv(k, m, 1, 1))
            mybad++;
        }
        __syncthreads();
        if (tid == 0 && !stop) {
          st_tag(a.gpart + ((size_t)s * ngroups + g) * 4, tag, tagv(k, g, 0, 0));
          st_tag(a.gpart + ((size_t)s * ngroups + g) * 4 + 2, tag, tagv(k, g, 1, 0));
        }
      }
      if (tid < ngroups) {
        double v[2];
        if (!ld_tag<2>(a.gpart + ((size_t)s * ngroups + tid) * 4, tag, v, a.failed))
          stop = 1;
        else if (v[0] != tagv(k, tid, 0, 0) || v[1] != tagv(k, tid, 1, 0))
          mybad++;
      }
    }
    __syncthreads();
    if (stop) break;
  }
  if (mybad) atomicAdd(&bad, mybad);
  __syncthreads();
  if (tid == 0 && bad) atomicAdd(a.stale, bad);
}

int elapsed_us(hipEvent_t a, hipEvent_t b, double* us) {
  float ms = 0.f;
  hipError_t e = hipEventElapsedTime(&ms, a, b);
  *us = 1e3 * ms;
  return (int)e;
}

}  // namespace gnx_probe

using namespace gnx_probe;

extern "C" {

int gnx_probe_graph_nodes(int kind, long sites, int nodes, int reps, double* us_per_node) {
  if (kind < 0 || kind > 4 || sites < 1 || nodes < 2 || reps < 1 || !us_per_node) return (int)hipErrorInvalidValue;
  hipError_t e;
  hipStream_t st = nullptr;
  double2* v[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* w = nullptr;
  double* part = nullptr;  // one partial sum per block (k_read88)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e30;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return (int)e;
  for (int k = 0; k < 5 && e == hipSuccess; k++) {
    e = hipMalloc(&v[k], sizeof(double2) * sites);
    if (e == hipSuccess) e = hipMemsetAsync(v[k], 0, sizeof(double2) * sites, st);
  }
  if (e == hipSuccess) e = hipMalloc(&w, sizeof(double) * sites);
  if (e == hipSuccess) e = hipMemsetAsync(w, 0, sizeof(double) * sites, st);
  if (e == hipSuccess) e = hipMalloc(&part, sizeof(double) * ((sites + kBS - 1) / kBS + 1));
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  const unsigned grid = kind == 0 ? 1024u : (unsigned)((sites + kBS - 1) / kBS);
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < nodes; i++) {
      if (kind == 0) {
        hipLaunchKernelGGL(k_empty, dim3(grid), dim3(kBS), 0, st, (int*)nullptr);
      } else if (kind == 1) {
        // each node reads the previous node's output
        hipLaunchKernelGGL(k_stream40, dim3(grid), dim3(kBS), 0, st, sites, (const double2*)v[i & 1],
                           (const double2*)v[2], (const double*)w, v[(i + 1) & 1]);
      } else if (kind == 2) {
        hipLaunchKernelGGL(k_stream152, dim3(grid), dim3(kBS), 0, st, sites, (const double2*)v[i & 1], (const double*)w,
                           v[2], v[3], v[4], v[(i + 1) & 1]);
      } else if (kind == 3) {
        hipLaunchKernelGGL(k_read88, dim3(grid), dim3(kBS), 0, st, sites, (const double2*)v[0], (const double*)w,
                           (const double2*)v[1], (const double2*)v[2], (const double2*)v[3], (const double2*)v[4],
                           part);
      } else {
        hipLaunchKernelGGL(k_write64, dim3(grid), dim3(kBS), 0, st, sites, (double)i, v[1], v[2], v[3], v[4]);
      }
    }
    e = hipStreamEndCapture(st, &graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipGraphLaunch(exec, st);  // warm-up
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    double us = 0.0;
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us);
    if (e == hipSuccess && us < best) best = us;
  }
  if (e == hipSuccess) *us_per_node = best / nodes;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (int k = 0; k < 5; k++)
    if (v[k]) (void)hipFree(v[k]);
  if (w) (void)hipFree(w);
  if (part) (void)hipFree(part);
  (void)hipStreamDestroy(st);
  return (int)e;
}

// nodes of k_width: 3 x bytes read and 1 x bytes written per node, through
// W-byte lanes (W = 4, 8 or 16), each node reading the previous node's output
int gnx_probe_width_nodes(int width, long bytes, int nodes, int reps, double* us_per_node) {
  if ((width != 4 && width != 8 && width != 16) || bytes < 16 || nodes < 2 || reps < 1 || !us_per_node)
    return (int)hipErrorInvalidValue;
  hipError_t e;
  hipStream_t st = nullptr;
  char* v[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e30;
  const long n = bytes / width;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return (int)e;
  for (int k = 0; k < 4 && e == hipSuccess; k++) {
    e = hipMalloc(&v[k], (size_t)n * width);
    if (e == hipSuccess) e = hipMemsetAsync(v[k], 0, (size_t)n * width, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  const unsigned grid = (unsigned)((n + kBS - 1) / kBS);
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < nodes; i++) {
      char* in0 = v[i & 1];
      char* o = v[(i + 1) & 1];
      if (width == 4)
        hipLaunchKernelGGL(k_width<unsigned>, dim3(grid), dim3(kBS), 0, st, n, (const unsigned*)in0,
                           (const unsigned*)v[2], (const unsigned*)v[3], (unsigned*)o);
      else if (width == 8)
        hipLaunchKernelGGL(k_width<uint2>, dim3(grid), dim3(kBS), 0, st, n, (const uint2*)in0, (const uint2*)v[2],
                           (const uint2*)v[3], (uint2*)o);
      else
        hipLaunchKernelGGL(k_width<uint4>, dim3(grid), dim3(kBS), 0, st, n, (const uint4*)in0, (const uint4*)v[2],
                           (const uint4*)v[3], (uint4*)o);
    }
    e = hipStreamEndCapture(st, &graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipGraphLaunch(exec, st);  // warm-up
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    double us = 0.0;
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us);
    if (e == hipSuccess && us < best) best = us;
  }
  if (e == hipSuccess) *us_per_node = best / nodes;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (int k = 0; k < 4; k++)
    if (v[k]) (void)hipFree(v[k]);
  (void)hipStreamDestroy(st);
  return (int)e;
}

// nodes of k_gather7 (width 4 or 16) or k_gather7_lds (width 0) over arrays
// of `elems` 4-byte elements (a multiple of 8192), each node reading the
// previous node's output
int gnx_probe_gather_nodes(int width, long elems, int nodes, int reps, double* us_per_node) {
  if ((width != 0 && width != 4 && width != 16) || elems < 4096 || elems % 8192 || nodes < 2 || reps < 1 ||
      !us_per_node)
    return (int)hipErrorInvalidValue;
  hipError_t e;
  hipStream_t st = nullptr;
  unsigned* v[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e30;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return (int)e;
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    e = hipMalloc(&v[k], sizeof(unsigned) * elems);
    if (e == hipSuccess) e = hipMemsetAsync(v[k], 0, sizeof(unsigned) * elems, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  const long lanes = width == 4 ? elems : elems / 4;
  // width 0: the LDS-staged form, one block per 1024-site tile
  const unsigned grid = width == 0 ? (unsigned)(elems / 1024) : (unsigned)((lanes + kBS - 1) / kBS);
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < nodes; i++) {
      const unsigned* in0 = v[i & 1];
      unsigned* o = v[(i + 1) & 1];
      if (width == 0)
        hipLaunchKernelGGL(k_gather7_lds, dim3(grid), dim3(kBS), 0, st, elems, in0, (const unsigned*)v[2], o);
      else if (width == 4)
        hipLaunchKernelGGL(k_gather7<4>, dim3(grid), dim3(kBS), 0, st, elems, in0, (const unsigned*)v[2], o);
      else
        hipLaunchKernelGGL(k_gather7<16>, dim3(grid), dim3(kBS), 0, st, elems, in0, (const unsigned*)v[2], o);
    }
    e = hipStreamEndCapture(st, &graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipGraphLaunch(exec, st);  // warm-up
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    double us = 0.0;
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us);
    if (e == hipSuccess && us < best) best = us;
  }
  if (e == hipSuccess) *us_per_node = best / nodes;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (int k = 0; k < 3; k++)
    if (v[k]) (void)hipFree(v[k]);
  (void)hipStreamDestroy(st);
  return (int)e;
}

// empty kernel nodes of a given grid and block size (the boundary's cost as
// a function of the workgroups it dispatches)
int gnx_probe_empty_nodes(int blocks, int threads, int nodes, int reps, double* us_per_node) {
  if (blocks < 1 || (threads != 64 && threads != 128 && threads != 256 && threads != 512 && threads != 1024) ||
      nodes < 2 || reps < 1 || !us_per_node)
    return (int)hipErrorInvalidValue;
  hipError_t e;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 1e30;
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return (int)e;
  e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < nodes; i++)
      hipLaunchKernelGGL(k_empty_any, dim3(blocks), dim3(threads), 0, st, (int*)nullptr);
    e = hipStreamEndCapture(st, &graph);
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipGraphLaunch(exec, st);  // warm-up
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    double us = 0.0;
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us);
    if (e == hipSuccess && us < best) best = us;
  }
  if (e == hipSuccess) *us_per_node = best / nodes;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(st);
  return (int)e;
}

int gnx_probe_grid_barrier(int mode, int blocks, int iters, long bytes_per_block, double* us_per_barrier,
                           int* max_blocks, int* failed) {
  if (mode < 0 || mode > 1 || blocks < 1 || iters < 1 || bytes_per_block < 0 || !us_per_barrier || !max_blocks ||
      !failed)
    return (int)hipErrorInvalidValue;
  const void* kern = mode == 0 ? (const void*)k_barrier_loop<true> : (const void*)k_barrier_loop<false>;
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBS, 0);
  if (e != hipSuccess) return (int)e;
  *max_blocks = cus * per_cu;
  // every block must be resident at once, or the barrier could never fill
  if (blocks > *max_blocks) return (int)hipErrorInvalidConfiguration;
  const long elems = (bytes_per_block + (long)sizeof(double2) - 1) / (long)sizeof(double2);
  BarrierState* bs = nullptr;
  double2* slice = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t st = nullptr;
  e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&bs, sizeof(BarrierState));
  if (e == hipSuccess) e = hipMalloc(&slice, sizeof(double2) * (elems > 0 ? elems : 1) * blocks);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double us[3] = {0.0, 0.0, 0.0};
  int rounds[3] = {1, 1, iters + 1};
  const BarrierState zero = {0u, 0u, 0, 0};
  int fail = 0;
  // the time of iters barriers is the difference of two launches (1 and
  // iters + 1 rounds), which removes the launch and the first round; the
  // first one-round launch is a warm-up
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    e = hipMemcpyAsync(bs, &zero, sizeof zero, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    void* args[] = {(void*)&bs, (void*)&rounds[k], (void*)&slice, (void*)&elems};
    if (e == hipSuccess) e = hipLaunchCooperativeKernel(kern, dim3(blocks), dim3(kBS), args, 0, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us[k]);
    BarrierState got = zero;
    if (e == hipSuccess) e = hipMemcpy(&got, bs, sizeof got, hipMemcpyDeviceToHost);
    fail |= got.failed;
  }
  if (e == hipSuccess) {
    *failed = fail;
    *us_per_barrier = (us[2] - us[1]) / iters;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (bs) (void)hipFree(bs);
  if (slice) (void)hipFree(slice);
  if (st) (void)hipStreamDestroy(st);
  return (int)e;
}

int gnx_probe_halo(int mode, int nt0, int nt1, int nt2, int iters, long lds_pad, double* us_per_round, int* failed,
                   int* stale, int* per_cu) {
  if (mode < 0 || mode > 3 || nt0 < 1 || nt1 < 1 || nt2 < 1 || iters < 1 || lds_pad < 0 || !us_per_round ||
      !failed || !stale || !per_cu)
    return (int)hipErrorInvalidValue;
  const int blocks = nt0 * nt1 * nt2;
  const void* kerns[4] = {(const void*)k_halo_loop<false, false>, (const void*)k_halo_loop<true, false>,
                          (const void*)k_halo_loop<false, true>, (const void*)k_halo_loop<true, true>};
  const void* kern = kerns[mode];
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess && lds_pad > 0)
    e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pad);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, kern, kHB, (size_t)lds_pad);
  if (e != hipSuccess) return (int)e;
  // every block must be resident at once (one per CU), or the waits could
  // never end; the partial wait has one thread per block
  if (blocks > cus * *per_cu || blocks > kHB) return (int)hipErrorInvalidConfiguration;
  HaloArgs a = {};
  a.nt0 = nt0;
  a.nt1 = nt1;
  a.nt2 = nt2;
  const size_t face_words = (size_t)2 * blocks * kFace * 2;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&a.rflag, sizeof(unsigned) * blocks);
  if (e == hipSuccess) e = hipMalloc(&a.qflag, sizeof(unsigned) * blocks);
  if (e == hipSuccess) e = hipMalloc(&a.rface, sizeof(unsigned long long) * face_words);
  if (e == hipSuccess) e = hipMalloc(&a.qface, sizeof(unsigned long long) * face_words);
  if (e == hipSuccess) e = hipMalloc(&a.part, sizeof(unsigned long long) * 4 * blocks);
  if (e == hipSuccess) e = hipMalloc(&a.failed, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a.stale, sizeof(int));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double us[3] = {0.0, 0.0, 0.0};
  int rounds[3] = {1, 1, iters + 1};
  int fail = 0, nstale = 0;
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    e = hipMemsetAsync(a.rflag, 0, sizeof(unsigned) * blocks, st);
    if (e == hipSuccess) e = hipMemsetAsync(a.qflag, 0, sizeof(unsigned) * blocks, st);
    if (e == hipSuccess) e = hipMemsetAsync(a.failed, 0, sizeof(int), st);
    if (e == hipSuccess) e = hipMemsetAsync(a.stale, 0, sizeof(int), st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    void* args[] = {(void*)&a, (void*)&rounds[k]};
    if (e == hipSuccess) e = hipLaunchKernel(kern, dim3(blocks), dim3(kHB), args, (size_t)lds_pad, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us[k]);
    int f = 0, sv = 0;
    if (e == hipSuccess) e = hipMemcpy(&f, a.failed, sizeof f, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&sv, a.stale, sizeof sv, hipMemcpyDeviceToHost);
    fail |= f;
    nstale += sv;
  }
  if (e == hipSuccess) {
    *failed = fail;
    *stale = nstale;
    *us_per_round = (us[2] - us[1]) / iters;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a.rflag);
  (void)hipFree(a.qflag);
  (void)hipFree(a.rface);
  (void)hipFree(a.qface);
  (void)hipFree(a.part);
  (void)hipFree(a.failed);
  (void)hipFree(a.stale);
  if (st) (void)hipStreamDestroy(st);
  return (int)e;
}

int gnx_probe_halo_tag(int red, int nt0, int nt1, int nt2, int iters, long lds_pad, double* us_per_round,
                       int* failed, int* stale, int* per_cu) {
  if (red < 0 || red > 2 || nt0 < 1 || nt1 < 1 || nt2 < 1 || iters < 1 || lds_pad < 0 || !us_per_round ||
      !failed || !stale || !per_cu)
    return (int)hipErrorInvalidValue;
  const int blocks = nt0 * nt1 * nt2;
  const int ngroups = (blocks + kGroup - 1) / kGroup;
  const void* kerns[3] = {(const void*)k_halo_tag<0>, (const void*)k_halo_tag<1>, (const void*)k_halo_tag<2>};
  const void* kern = kerns[red];
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess && lds_pad > 0)
    e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pad);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, kern, kHB, (size_t)lds_pad);
  if (e != hipSuccess) return (int)e;
  if (blocks > cus * *per_cu || blocks > kHB || ngroups > kHB) return (int)hipErrorInvalidConfiguration;
  TagArgs a = {};
  a.nt0 = nt0;
  a.nt1 = nt1;
  a.nt2 = nt2;
  const size_t face_words = (size_t)2 * blocks * kFace * 4;
  const size_t part_words = (size_t)2 * blocks * 4, gpart_words = (size_t)2 * ngroups * 4;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&a.rface, sizeof(unsigned long long) * face_words);
  if (e == hipSuccess) e = hipMalloc(&a.qface, sizeof(unsigned long long) * face_words);
  if (e == hipSuccess) e = hipMalloc(&a.part, sizeof(unsigned long long) * part_words);
  if (e == hipSuccess) e = hipMalloc(&a.gpart, sizeof(unsigned long long) * gpart_words);
  if (e == hipSuccess) e = hipMalloc(&a.failed, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&a.stale, sizeof(int));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double us[3] = {0.0, 0.0, 0.0};
  int rounds[3] = {1, 1, iters + 1};
  int fail = 0, nstale = 0;
  for (int k = 0; k < 3 && e == hipSuccess; k++) {
    // tag 0 matches no round
    e = hipMemsetAsync(a.rface, 0, sizeof(unsigned long long) * face_words, st);
    if (e == hipSuccess) e = hipMemsetAsync(a.qface, 0, sizeof(unsigned long long) * face_words, st);
    if (e == hipSuccess) e = hipMemsetAsync(a.part, 0, sizeof(unsigned long long) * part_words, st);
    if (e == hipSuccess) e = hipMemsetAsync(a.gpart, 0, sizeof(unsigned long long) * gpart_words, st);
    if (e == hipSuccess) e = hipMemsetAsync(a.failed, 0, sizeof(int), st);
    if (e == hipSuccess) e = hipMemsetAsync(a.stale, 0, sizeof(int), st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    void* args[] = {(void*)&a, (void*)&rounds[k]};
    if (e == hipSuccess) e = hipLaunchKernel(kern, dim3(blocks), dim3(kHB), args, (size_t)lds_pad, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = (hipError_t)elapsed_us(e0, e1, &us[k]);
    int f = 0, sv = 0;
    if (e == hipSuccess) e = hipMemcpy(&f, a.failed, sizeof f, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&sv, a.stale, sizeof sv, hipMemcpyDeviceToHost);
    fail |= f;
    nstale += sv;
  }
  if (e == hipSuccess) {
    *failed = fail;
    *stale = nstale;
    *us_per_round = (us[2] - us[1]) / iters;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a.rface);
  (void)hipFree(a.qface);
  (void)hipFree(a.part);
  (void)hipFree(a.gpart);
  (void)hipFree(a.failed);
  (void)hipFree(a.stale);
  if (st) (void)hipStreamDestroy(st);
  return (int)e;
}

}  // extern "C"
