// gnx_engine.hip -- graph capture, replay and polling for the solvers.
#include <string.h>

#include "gnx_engine.h"
#include "gnx_replay_plan.h"

namespace gnx {

GraphEngine::~GraphEngine() { release(); }

void GraphEngine::invalidate_graphs() {
  for (auto& g : graphs_)
    if (g.g.chunk) {
      (void)hipGraphExecDestroy(g.g.chunk);
      for (hipGraphExec_t t : g.g.tail)
        if (t) (void)hipGraphExecDestroy(t);
      g = Entry();
    }
}

unsigned long long eng_key(int tag, int nrhs, double tol, int start) {
  unsigned long long t;
  memcpy(&t, &tol, sizeof t);
  unsigned long long h = 1469598103934665603ull;  // FNV-1a over the four fields
  const unsigned long long w[4] = {(unsigned long long)(unsigned)tag, (unsigned long long)(unsigned)nrhs, t,
                                   (unsigned long long)(unsigned)start};
  for (unsigned long long v : w)
    for (int b = 0; b < 8; b++) {
      h ^= (v >> (8 * b)) & 0xFFu;
      h *= 1099511628211ull;
    }
  return h;
}

GraphEngine::Pred* GraphEngine::pred_find(unsigned long long key) {
  for (auto& p : pred_)
    if (p.valid && p.key == key) {
      p.used = ++pred_clock_;
      return &p;
    }
  return nullptr;
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

// capture `iters` iterations of `body` into an executable graph
static hipError_t capture_graph(hipStream_t cs, const ChunkBody& body, int iters, hipGraphExec_t* exec) {
  hipGraph_t graph;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  body(cs, iters);
  e = hipStreamEndCapture(cs, &graph);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) *exec = nullptr;
  return e;
}

hipError_t GraphEngine::graph(int tag, int nrhs, const ChunkBody& body, EngGraph* out) {
  for (auto& g : graphs_)
    if (g.g.chunk && g.tag == tag && g.nrhs == nrhs) {
      *out = g.g;
      return hipSuccess;
    }
  Entry* slot = nullptr;
  for (auto& g : graphs_)
    if (!g.g.chunk) {
      slot = &g;
      break;
    }
  if (!slot) {
    slot = &graphs_[next_evict_];
    next_evict_ = (next_evict_ + 1) % (int)(sizeof(graphs_) / sizeof(graphs_[0]));
    (void)hipGraphExecDestroy(slot->g.chunk);
    for (hipGraphExec_t t : slot->g.tail)
      if (t) (void)hipGraphExecDestroy(t);
    *slot = Entry();
  }
  EngGraph g;
  hipError_t e = capture_graph(capture_, body, ENG_CHUNK, &g.chunk);
  for (int k = 0; k < ENG_TAILS && e == hipSuccess; k++)
    e = capture_graph(capture_, body, (k + 1) * ENG_TAIL, &g.tail[k]);
  if (e != hipSuccess) {
    if (g.chunk) (void)hipGraphExecDestroy(g.chunk);
    for (hipGraphExec_t t : g.tail)
      if (t) (void)hipGraphExecDestroy(t);
    return e;
  }
  slot->tag = tag;
  slot->nrhs = nrhs;
  slot->g = g;
  *out = g;
  return hipSuccess;
}

hipError_t GraphEngine::run(const EngGraph& g, int nrhs, long max_slots, hipStream_t st, int* max_iters,
                            unsigned long long key, SlotsOf slots_of, const FinishFn* finish) {
  Pred* p = pred_find(key);
  ReplayKind kind;
  if (p) {
    kind.last = p->last;
    kind.it_last = p->it_last;
    kind.it_prev = p->it_prev;
    kind.exits = p->exits;
    kind.runs = p->runs;
  }
  ReplayBurst b = replay_first(p ? &kind : nullptr, last_slots_);
  long done = 0;
  int misses = 0;
  hipError_t e;
  for (;;) {
    for (int i = 0; i < b.full; i++) {
      e = hipGraphLaunch(g.chunk, st);
      if (e != hipSuccess) return e;
    }
    if (b.tails > 0) {
      e = hipGraphLaunch(g.tail[b.tails - 1], st);
      if (e != hipSuccess) return e;
    }
    done += b.slots();
    launched_slots += b.slots();
    polls++;
    e = hipMemcpyAsync(hstate_, state_, state_bytes_ * nrhs, hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    if (finish) {
      e = (*finish)(st);
      if (e != hipSuccess) return e;
    }
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    bool any = false;
    int mx = 0, slots = 0;
    for (int k = 0; k < nrhs; k++) {
      const char* sk = hstate_ + state_bytes_ * k;
      const int* s = (const int*)sk;
      any |= s[0] != 0;
      mx = s[1] > mx ? s[1] : mx;
      const int u = slots_of ? slots_of(sk) : s[1] + 2;
      slots = u > slots ? u : slots;
    }
    if (!any) {
      if (!p) {  // a new kind: take the least recently used entry
        p = &pred_[0];
        for (auto& q : pred_)
          if (q.used < p->used) p = &q;
        *p = Pred();
        p->key = key;
        p->valid = true;
        p->used = ++pred_clock_;
      }
      replay_record(&kind, slots, mx);
      p->last = kind.last;
      p->it_last = kind.it_last;
      p->it_prev = kind.it_prev;
      p->exits = kind.exits;
      p->runs = kind.runs;
      last_slots_ = slots;
      if (max_iters) *max_iters = mx;
      return hipSuccess;
    }
    if (done > max_slots + 2 * ENG_CHUNK) return hipErrorUnknown;  // the state machine failed to stop
    b = replay_after_miss(p ? &kind : nullptr, done, misses++);
  }
}

hipError_t GraphEngine::time_replays(const EngGraph& g, int chunks, hipStream_t st, float* ms) {
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return e;
  }
  (void)hipEventRecord(e0, st);
  for (int k = 0; k < chunks && e == hipSuccess; k++) e = hipGraphLaunch(g.chunk, st);
  (void)hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return e;
}

}  // namespace gnx
