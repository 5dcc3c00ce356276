// gnx_replay_plan.h -- how many iterations the replay engine launches per
// burst (host-only and pure, so the rules can be checked on the CPU).
//
// A burst is whole chunks (ENG_CHUNK iterations) and at most one tail graph
// of 1..ENG_TAILS times ENG_TAIL iterations; the engine polls the systems'
// state after each burst.  A run's first burst is what the last run of the
// same kind needed, plus, when that run iterated, the change between the
// kind's last two iterating runs (at most half a chunk).  A run that stops
// in its prologue (a null start already below the tolerance) needs 2 slots
// and says nothing about the next run, so the kind also keeps its last
// iterating runs: after a first burst that falls short, the engine launches
// what the last iterating run used beyond what is already launched (else a
// tail of two), and after that whole chunks.  A kind whose recent runs
// often stopped in the prologue starts with one short tail instead of its
// predicted need when the slots that need would waste on an exit outweigh
// the poll it costs an iterating run (REPLAY_POLL_SLOTS empty slots).
#pragma once

#include "gnx_engine.h"

namespace gnx {

// a poll (host round trip, then the next launch) costs about as much GPU time
// as this many slots of stopped systems (two empty graph nodes each)
constexpr int REPLAY_POLL_SLOTS = 10;
constexpr int REPLAY_HISTORY = 16;  // runs per kind whose prologue exits are counted

struct ReplayBurst {
  int full;   // whole chunks
  int tails;  // one tail graph of tails * ENG_TAIL iterations (0: none)
  long slots() const { return (long)full * ENG_CHUNK + (long)tails * ENG_TAIL; }
};

// the slots needed by a kind's last run and by its last two iterating runs
// (-1: none yet), and which of its last runs stopped in the prologue (bit 0
// the last run)
struct ReplayKind {
  int last = -1;
  int it_last = -1, it_prev = -1;
  unsigned exits = 0;
  int runs = 0;
};

// at least w slots as whole chunks and at most one tail (a remainder past
// the longest tail takes a whole chunk)
inline ReplayBurst replay_split(long w) {
  ReplayBurst b;
  b.full = (int)(w / ENG_CHUNK);
  b.tails = (int)((w % ENG_CHUNK + ENG_TAIL - 1) / ENG_TAIL);
  if (b.tails > ENG_TAILS) {
    b.full++;
    b.tails = 0;
  }
  return b;
}

// first burst of a run of kind k (null: a kind not seen yet, which goes by
// last_any, the engine's last run of any kind; 0 before any run)
inline ReplayBurst replay_first(const ReplayKind* k, int last_any) {
  const int need = k ? k->last : last_any;
  int margin = 0;
  if (k && k->last == k->it_last && k->it_prev >= 0)
    margin = k->it_last > k->it_prev ? k->it_last - k->it_prev : k->it_prev - k->it_last;
  if (margin > ENG_CHUNK / 2) margin = ENG_CHUNK / 2;
  const long want = need > 0 ? (long)need + margin : ENG_CHUNK;
  if (k && k->runs > 0 && want > ENG_TAIL) {
    // expected waste of the full burst on a prologue exit against the poll
    // a short first burst costs an iterating run
    const int n = k->runs < REPLAY_HISTORY ? k->runs : REPLAY_HISTORY;
    const int e = __builtin_popcount(k->exits & ((1u << n) - 1u));
    if ((long)e * (want - ENG_TAIL) > (long)REPLAY_POLL_SLOTS * (n - e)) return ReplayBurst{0, 1};
  }
  return replay_split(want);
}

// the burst after a poll found a system still active; `done` slots are
// launched so far and `misses` bursts fell short before this one
inline ReplayBurst replay_after_miss(const ReplayKind* k, long done, int misses) {
  if (misses > 0) return ReplayBurst{1, 0};
  const long rest = (k && k->it_last > 0) ? k->it_last - done : 0;
  if (rest >= 2 * ENG_TAIL) return replay_split(rest);
  return ReplayBurst{0, 2};
}

// a finished run of kind k needed `slots`; `iters` > 0 if it iterated
inline void replay_record(ReplayKind* k, int slots, int iters) {
  k->last = slots;
  k->exits = (k->exits << 1) | (iters > 0 ? 0u : 1u);
  if (k->runs < REPLAY_HISTORY) k->runs++;
  if (iters > 0) {
    k->it_prev = k->it_last;
    k->it_last = slots;
  }
}

}  // namespace gnx
