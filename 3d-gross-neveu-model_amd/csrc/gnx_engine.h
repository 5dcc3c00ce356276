// gnx_engine.h -- host machinery shared by the graph-replayed solvers.
//
// A solver is a pair of kernels (X, Y) that together perform one iteration
// of every active system of a batch, plus start kernels launched before the
// replay.  The kernels read their arguments from a device copy of a plain
// struct and keep per-system state in a device array state[2][nrhs]; the
// first two ints of each state entry are `active` and `iters`.  The engine
// captures CHUNK (X, Y) pairs into one hipGraph per (kernel pair, batch
// size), replays it until no system is active, and polls only the `active`
// flags between replays.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

namespace gnx {

constexpr int ENG_BS = 256;     // threads per block of every solver kernel
constexpr int ENG_CHUNK = 32;   // iterations per graph replay

// launches ENG_CHUNK iterations (every X and Y kernel) into the capture stream
using ChunkBody = std::function<void(hipStream_t)>;

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
  // graph of one chunk for (tag, nrhs), captured from `body` on first use
  hipError_t graph(int tag, int nrhs, const ChunkBody& body, hipGraphExec_t* exec);
  // drop every captured graph (their kernel arguments point at old buffers)
  void invalidate_graphs();
  // replay until every system is inactive; returns the largest `iters`
  hipError_t run(hipGraphExec_t exec, int nrhs, long max_slots, hipStream_t st, int* max_iters);
  // time `chunks` replays (after the caller staged a start), in ms
  hipError_t time_replays(hipGraphExec_t exec, int chunks, hipStream_t st, float* ms);

  int last_iters = 0;
  ShapeCache shapes;   // chosen launch shape per key
  long tune_runs = 0;  // timing-tuner runs (one per key chosen by timing)

 private:
  void release();
  struct Entry {
    int tag = -1;
    int nrhs = 0;
    hipGraphExec_t exec = nullptr;
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
};

}  // namespace gnx
