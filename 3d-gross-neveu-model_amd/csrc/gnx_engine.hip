// gnx_engine.hip -- graph capture, replay and polling for the solvers.
#include <string.h>

#include "gnx_engine.h"

namespace gnx {

GraphEngine::~GraphEngine() { release(); }

void GraphEngine::invalidate_graphs() {
  for (auto& g : graphs_)
    if (g.exec) {
      (void)hipGraphExecDestroy(g.exec);
      g = Entry();
    }
}

void GraphEngine::release() {
  invalidate_graphs();
  if (dargs_) (void)hipFree(dargs_);
  if (hargs_) (void)hipHostFree(hargs_);
  if (state_) (void)hipFree(state_);
  if (hstate_) (void)hipHostFree(hstate_);
  if (part_) (void)hipFree(part_);
  if (capture_) (void)hipStreamDestroy(capture_);
  dargs_ = hargs_ = state_ = nullptr;
  hstate_ = nullptr;
  part_ = nullptr;
  capture_ = nullptr;
  cap_ = 0;
  args_bytes_ = state_bytes_ = part_doubles_ = 0;
}

hipError_t GraphEngine::ensure(int nrhs, size_t args_bytes, size_t state_bytes, size_t part_doubles) {
  if (cap_ >= nrhs && args_bytes_ >= args_bytes && state_bytes_ == state_bytes && part_doubles_ >= part_doubles)
    return hipSuccess;
  release();
  int cap = nrhs < 2 ? 2 : nrhs;
  hipError_t e = hipMalloc(&dargs_, args_bytes);
  if (e == hipSuccess) e = hipHostMalloc(&hargs_, args_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&state_, state_bytes * 2 * cap);
  if (e == hipSuccess) e = hipHostMalloc((void**)&hstate_, state_bytes * cap, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&part_, sizeof(double) * part_doubles);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&capture_, hipStreamNonBlocking);
  if (e != hipSuccess) {
    release();
    return e;
  }
  memset(hargs_, 0, args_bytes);
  cap_ = cap;
  args_bytes_ = args_bytes;
  state_bytes_ = state_bytes;
  part_doubles_ = part_doubles;
  return hipSuccess;
}

hipError_t GraphEngine::stage_args(hipStream_t st) {
  // the previous solve's staging may still be read by an in-flight copy
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(dargs_, hargs_, args_bytes_, hipMemcpyHostToDevice, st);
}

hipError_t GraphEngine::graph(int tag, int nrhs, const ChunkBody& body, hipGraphExec_t* exec) {
  for (auto& g : graphs_)
    if (g.exec && g.tag == tag && g.nrhs == nrhs) {
      *exec = g.exec;
      return hipSuccess;
    }
  Entry* slot = nullptr;
  for (auto& g : graphs_)
    if (!g.exec) {
      slot = &g;
      break;
    }
  if (!slot) {
    slot = &graphs_[next_evict_];
    next_evict_ = (next_evict_ + 1) % (int)(sizeof(graphs_) / sizeof(graphs_[0]));
    (void)hipGraphExecDestroy(slot->exec);
    *slot = Entry();
  }
  hipGraph_t graph;
  hipError_t e = hipStreamBeginCapture(capture_, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  body(capture_);
  e = hipStreamEndCapture(capture_, &graph);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    slot->exec = nullptr;
    return e;
  }
  slot->tag = tag;
  slot->nrhs = nrhs;
  *exec = slot->exec;
  return hipSuccess;
}

hipError_t GraphEngine::run(hipGraphExec_t exec, int nrhs, long max_slots, hipStream_t st, int* max_iters) {
  long done = 0;
  int burst = (last_iters + ENG_CHUNK - 1) / ENG_CHUNK;
  if (burst < 1) burst = 1;
  hipError_t e;
  for (;;) {
    for (int i = 0; i < burst; i++) {
      e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) return e;
    }
    done += (long)burst * ENG_CHUNK;
    e = hipMemcpyAsync(hstate_, state_, state_bytes_ * nrhs, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    bool any = false;
    int mx = 0;
    for (int k = 0; k < nrhs; k++) {
      const int* s = (const int*)(hstate_ + state_bytes_ * k);
      any |= s[0] != 0;
      mx = s[1] > mx ? s[1] : mx;
    }
    if (!any) {
      last_iters = mx;
      if (max_iters) *max_iters = mx;
      return hipSuccess;
    }
    if (done > max_slots + 2 * ENG_CHUNK) return hipErrorUnknown;  // the state machine failed to stop
    burst = 1;
  }
}

hipError_t GraphEngine::time_replays(hipGraphExec_t exec, int chunks, hipStream_t st, float* ms) {
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return e;
  }
  (void)hipEventRecord(e0, st);
  for (int k = 0; k < chunks && e == hipSuccess; k++) e = hipGraphLaunch(exec, st);
  (void)hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return e;
}

}  // namespace gnx
