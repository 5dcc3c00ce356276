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
//
// C ABI (host functions; all return hipError_t as int, 0 on success):
//   int gnx_probe_graph_nodes(int kind, long sites, int nodes, int reps, double* us_per_node);
//   int gnx_probe_grid_barrier(int mode, int blocks, int iters, long bytes_per_block,
//                              double* us_per_barrier, int* max_blocks, int* failed);
//   mode 0: every poll an acquire load; mode 1: relaxed polls, one acquire fence
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gnx_probe {

constexpr int kBS = 256;

__global__ __launch_bounds__(kBS) void k_empty(int* sink) {
  if (sink && threadIdx.x == 0 && blockIdx.x == 0x7fffffff) *sink = 1;  // never true: keeps the kernel
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

}  // extern "C"
