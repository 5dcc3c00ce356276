// CPU check of the replay engine's burst sizes (gnx_replay_plan.h, the
// functions GraphEngine::run calls).  Build and run by
// tests/test_host_replay.py; exit status 0 on success.
#include <cstdio>
#include <map>
#include <vector>

#include "../../3d-gross-neveu-model_amd/csrc/gnx_replay_plan.h"

using namespace gnx;

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails < 20) { printf(__VA_ARGS__); printf("\n"); } fails++; } } while (0)

struct Run {
  int kind, slots, iters;
};
struct Totals {
  long launched = 0, needed = 0, polls = 0;
  int worst_polls = 0;
};

// replays a sequence of runs through the rules as GraphEngine::run does:
// first burst, then bursts after each miss until the need is launched
static Totals simulate(const std::vector<Run>& seq, std::vector<int>* polls_of = nullptr) {
  std::map<int, ReplayKind> kinds;
  int last_any = 0;
  Totals t;
  for (const Run& r : seq) {
    auto it = kinds.find(r.kind);
    const ReplayKind* k = it == kinds.end() ? nullptr : &it->second;
    ReplayBurst b = replay_first(k, last_any);
    long done = b.slots();
    int polls = 1, misses = 0;
    while (done < r.slots) {
      b = replay_after_miss(k, done, misses++);
      EXPECT(b.slots() > 0, "a burst after a miss launches nothing");
      if (b.slots() <= 0) break;
      done += b.slots();
      polls++;
    }
    replay_record(&kinds[r.kind], r.slots, r.iters);
    last_any = r.slots;
    t.launched += done;
    t.needed += r.slots;
    t.polls += polls;
    if (polls > t.worst_polls) t.worst_polls = polls;
    if (polls_of) polls_of->push_back(polls);
  }
  return t;
}

int main() {
  // splits: at least w slots, whole chunks and at most one tail, less than
  // one tail of waste
  for (long w = 1; w <= 2000; w++) {
    const ReplayBurst b = replay_split(w);
    EXPECT(b.tails >= 0 && b.tails <= ENG_TAILS, "split(%ld): tails %d", w, b.tails);
    EXPECT(b.slots() >= w && b.slots() - w < ENG_TAIL, "split(%ld) = %ld slots", w, b.slots());
  }
  EXPECT(replay_split(0).slots() == 0, "split(0)");

  // a kind not seen: the engine's last run of any kind, or one chunk
  EXPECT(replay_first(nullptr, 0).slots() == ENG_CHUNK, "first run of an engine");
  EXPECT(replay_first(nullptr, 407).slots() == 408, "a new kind after a 407-slot run");

  // a steady kind: the same need, no margin
  ReplayKind k;
  replay_record(&k, 407, 405);
  EXPECT(replay_first(&k, 0).slots() == 408, "one run of 407: %ld", replay_first(&k, 0).slots());
  replay_record(&k, 407, 405);
  EXPECT(replay_first(&k, 0).slots() == 408, "two runs of 407");
  // a moving kind: the change as headroom, at most half a chunk
  ReplayKind m;
  replay_record(&m, 100, 98);
  replay_record(&m, 110, 108);
  EXPECT(replay_first(&m, 0).slots() == 120, "100 then 110: %ld", replay_first(&m, 0).slots());
  replay_record(&m, 60, 58);
  EXPECT(replay_first(&m, 0).slots() == 80, "110 then 60: %ld (60 + 16)", replay_first(&m, 0).slots());
  // a prologue exit: 2 slots next, no margin from it, and the miss
  // recovers the last iterating run's need in one burst
  ReplayKind q;
  replay_record(&q, 40, 38);
  replay_record(&q, 40, 38);
  replay_record(&q, 2, 0);
  EXPECT(q.it_last == 40 && q.it_prev == 40, "a prologue exit changed the iterating history");
  EXPECT(replay_first(&q, 0).slots() == ENG_TAIL, "after a prologue exit: %ld", replay_first(&q, 0).slots());
  const ReplayBurst r1 = replay_after_miss(&q, ENG_TAIL, 0);
  EXPECT(ENG_TAIL + r1.slots() == 40, "recovery after a prologue exit: %ld", ENG_TAIL + r1.slots());
  EXPECT(replay_after_miss(&q, 40, 1).slots() == ENG_CHUNK, "a second miss takes a chunk");
  // a near miss with no iterating history: a tail of two
  EXPECT(replay_after_miss(nullptr, 32, 0).slots() == 2 * ENG_TAIL, "near miss");

  // the pattern of the HMC's solves on one context (logged from the bench's
  // 16-chain ensembles): per trajectory an energy solve (kind 0, ~105
  // iterations), a force solve that stops in its prologue, then 19 force
  // solves of ~38-46 iterations (kind 1); a link trajectory's first force
  // solves iterate (~104), then its 11 stop in the prologue (kind 2)
  std::vector<Run> seq;
  for (int traj = 0; traj < 6; traj++) {
    seq.push_back({0, 105 + (traj % 3) * 3 + 2, 105 + (traj % 3) * 3});
    seq.push_back({1, 2, 0});
    for (int s = 0; s < 19; s++) {
      const int it = 38 + (traj % 2) * 4 + (s > 12 ? 1 : 0);
      seq.push_back({1, it + 2, it});
    }
    seq.push_back({2, 106, 104});
    seq.push_back({2, 105, 103});
    for (int s = 0; s < 11; s++) seq.push_back({2, 2, 0});
  }
  std::vector<int> polls;
  const Totals all = simulate(seq, &polls);
  // after the first trajectory (each kind's first runs have no history)
  const size_t warm = 34;
  const std::vector<Run> later(seq.begin() + warm, seq.end());
  Totals t;
  {
    const Totals first = simulate(std::vector<Run>(seq.begin(), seq.begin() + warm));
    t.launched = all.launched - first.launched;
    t.needed = all.needed - first.needed;
    t.polls = all.polls - first.polls;
    for (size_t i = warm; i < polls.size(); i++) t.worst_polls = polls[i] > t.worst_polls ? polls[i] : t.worst_polls;
  }
  const double waste = (double)(t.launched - t.needed) / (double)t.launched;
  printf("HMC pattern: %zu runs, %ld polls, %ld launched for %ld needed; after the first trajectory %zu runs, "
         "%ld polls (at most %d per run), %.1f %% of the launched slots after the systems stopped\n",
         seq.size(), all.polls, all.launched, all.needed, later.size(), t.polls, t.worst_polls, 100.0 * waste);
  EXPECT(t.worst_polls <= 2, "a run polled %d times", t.worst_polls);
  EXPECT(t.polls <= (long)later.size() * 6 / 5, "%ld polls for %zu runs", t.polls, later.size());
  EXPECT(waste < 0.17, "%.1f %% of the launched slots after the systems stopped", 100.0 * waste);
  // the steady force solves each poll once
  for (size_t i = warm; i < seq.size(); i++)
    if (seq[i].kind == 1 && seq[i - 1].kind == 1 && seq[i - 1].iters > 0 && seq[i].slots == seq[i - 1].slots)
      EXPECT(polls[i] == 1, "steady force solve %zu polled %d times", i, polls[i]);

  if (fails) {
    printf("replay_plan_check: %d failures\n", fails);
    return 1;
  }
  printf("replay_plan_check: ok\n");
  return 0;
}
