// gnx_engine.h -- host machinery shared by the graph-replayed solvers.
//
// A solver is a pair of kernels (X, Y) that together perform one iteration
// of every active system of a batch, plus start kernels launched before the
// replay.  The kernels read their arguments from a device copy of a plain
// struct and keep per-system state in a device array state[2][nrhs]; the
// first two ints of each state entry are `active` and `iters`.  The engine
// captures CHUNK (X, Y) pairs into one hipGraph per (kernel pair, batch
// size), and 1, 2 and 3 TAILs of pairs into three more, replays them until
// no system is active, and polls only the state between replays.  A
// replayed slot of a finished system still costs its two graph nodes, and
// every replay a graph launch, so each run first launches about as many
// slots as the last run of the same kind needed, as whole chunks and at most
// one tail graph, and polls once.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

namespace gnx {

constexpr int ENG_BS = 256;     // threads per block of every solver kernel
constexpr int ENG_CHUNK = 32;   // iterations per graph replay
constexpr int ENG_TAIL = 8;     // tail replays are 1, 2 or 3 of these (even, and a multiple of the deferred form's ring)
constexpr int ENG_TAILS = 3;
static_assert(ENG_CHUNK == (ENG_TAILS + 1) * ENG_TAIL, "a chunk is one tail more than the longest tail graph");

// launches `iters` iterations (every X and Y kernel) into the capture stream;
// iters is ENG_CHUNK or a multiple of ENG_TAIL, so every replay starts at
// slot parity 0
using ChunkBody = std::function<void(hipStream_t, int iters)>;

// the captured graphs of a (kernel pair, batch size): a chunk, and tails of
// ENG_TAIL, 2 ENG_TAIL and 3 ENG_TAIL iterations
struct EngGraph {
  hipGraphExec_t chunk = nullptr;
  hipGraphExec_t tail[ENG_TAILS] = {nullptr, nullptr, nullptr};
};

// Slots (X, Y pairs) a finished system used, from its polled state entry, at
// least: the engine launches this many for the next run of the same kind.
// The default, for states whose every slot but the last counts an
// iteration, is iters + 2.
using SlotsOf = int (*)(const char* state);

// what identifies runs of one kind (expected to take similar slot counts):
// the graph, the batch size, the tolerance and the start
unsigned long long eng_key(int tag, int nrhs, double tol, int start);

// work a solve enqueues once its systems have stopped (copies of the
// solutions and results out of the solver's buffers): the engine enqueues it
// after every burst, ahead of the poll's wait, so that a run that ends at a
// poll needs no second wait; a copy made after a burst that fell short is
// made again after the next one (the solvers read only their own buffers
// while they run)
using FinishFn = std::function<hipError_t(hipStream_t)>;

// Launch shapes chosen per key (batch size with the operator or storage
// format folded in).  When every slot is taken the least recently used
// entry is replaced, so a context that sees many keys re-chooses only the
// keys it evicted, never every key on every call.
struct ShapeCache {
  static constexpr int N = 16;
  int key[N];
  int shape[N];
  unsigned long long used[N];
  unsigned long long clock = 0;
  ShapeCache() { clear(); }
  void clear() {
    for (int i = 0; i < N; i++) {
      key[i] = -1;
      shape[i] = -1;
      used[i] = 0;
    }
  }
  // the shape stored for k (and marks it used), or -1
  int find(int k) {
    for (int i = 0; i < N; i++)
      if (key[i] == k) {
        used[i] = ++clock;
        return shape[i];
      }
    return -1;
  }
  // the shape stored for k without marking it used, or -1
  int peek(int k) const {
    for (int i = 0; i < N; i++)
      if (key[i] == k) return shape[i];
    return -1;
  }
  void put(int k, int s) {
    int slot = 0;
    for (int i = 0; i < N; i++) {
      if (key[i] == k || key[i] < 0) {
        slot = i;
        break;
      }
      if (used[i] < used[slot]) slot = i;
    }
    key[slot] = k;
    shape[slot] = s;
    used[slot] = ++clock;
  }
};

class GraphEngine {
 public:
  ~GraphEngine();
  // (re)allocate for nrhs systems; args/state sizes in bytes, partials in doubles
  hipError_t ensure(int nrhs, size_t args_bytes, size_t state_bytes, size_t part_doubles);
  void* host_args() { return hargs_; }
  const void* dev_args() const { return dargs_; }
  void* dev_state() { return state_; }
  double* dev_part() { return part_; }
  // wait for the stream, then copy the staged host arguments to the device
  hipError_t stage_args(hipStream_t st);
  // graphs of one chunk and the tails for (tag, nrhs), captured from
  // `body` on first use
  hipError_t graph(int tag, int nrhs, const ChunkBody& body, EngGraph* g);
  // drop every captured graph (their kernel arguments point at old buffers)
  void invalidate_graphs();
  // replay until every system is inactive; returns the largest `iters`.
  // key: eng_key() of the run; slots_of: null for the default; finish:
  // null, or the solve's closing work, complete on a successful return
  hipError_t run(const EngGraph& g, int nrhs, long max_slots, hipStream_t st, int* max_iters,
                 unsigned long long key, SlotsOf slots_of, const FinishFn* finish = nullptr);
  // time `chunks` chunk replays (after the caller staged a start), in ms
  hipError_t time_replays(const EngGraph& g, int chunks, hipStream_t st, float* ms);
  // slots launched and polls made by run() since construction
  long launched_slots = 0, polls = 0;
  ShapeCache shapes;   // chosen launch shape per key
  long tune_runs = 0;  // timing-tuner runs (one per key chosen by timing)

 private:
  void release();
  struct Entry {
    int tag = -1;
    int nrhs = 0;
    EngGraph g;
  };
  // what the runs of a kind needed (gnx_replay_plan.h), least recently
  // used replaced
  struct Pred {
    unsigned long long key = 0;
    bool valid = false;
    int last = -1, it_last = -1, it_prev = -1;
    unsigned exits = 0;
    int runs = 0;
    unsigned long long used = 0;
  };
  void* dargs_ = nullptr;
  void* hargs_ = nullptr;
  void* state_ = nullptr;
  char* hstate_ = nullptr;
  double* part_ = nullptr;
  size_t args_bytes_ = 0, state_bytes_ = 0, part_doubles_ = 0;
  int cap_ = 0;
  hipStream_t capture_ = nullptr;
  Entry graphs_[12];
  int next_evict_ = 0;
  Pred pred_[16];
  unsigned long long pred_clock_ = 0;
  int last_slots_ = 0;  // the last run's, for a kind not seen before
  Pred* pred_find(unsigned long long key);
};

}  // namespace gnx
