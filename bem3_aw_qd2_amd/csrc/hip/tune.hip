// tune.hip -- online autotuner of launch shapes (declarations in tune.h).
//
// Exploration runs on the library's real calls, so it costs no extra work: a
// call of a class still being explored launches the candidate with the fewest
// timed calls so far, bracketed by two HIP events, until each candidate has
// kSamples of them; later calls of the class take candidate 0 untimed until
// the timings are in.  Event pairs are read with
// hipEventQuery at later tuner calls (never waiting).  A class spans sizes
// within a factor sqrt(2), so a timing counts as a rate: the caller's measure
// of the call's work (flops, bytes) per millisecond.  Once every candidate of a
// class holds kSamples rates the class is fixed to the candidate with the
// highest one (the maximum discards launches slowed by first-use code loading
// or by other streams).  Every candidate gives the same bits, so a result never
// depends on the tuner's state.  All state is process-wide and guarded by one
// mutex (the serving loop factors and solves on several threads).
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include <vector>
#include "common.h"
#include "tune.h"

namespace {

const int kCands[TUNE_NSITES] = {2, 4};
const int kSamples = 3;
const int kMaxPending = 512;

struct Entry {
  int dev, site;
  long long key;
  int chosen;          // -1 while exploring
  int issued[4];       // calls launched per candidate
  int n[4];            // timings read per candidate
  float best[4];       // highest rate per candidate (work per ms)
};

struct Pending {
  int entry, cand, gen;
  double work;
  hipEvent_t e0, e1;
};

std::mutex g_mu;
std::vector<Entry> g_entries;
std::vector<Pending> g_pending;      // indexed by slot; e0 == nullptr when free
std::vector<hipEvent_t> g_events;    // recycled events
int g_mode = 1;                      // 0: default candidates, 1: online tuning
int g_force[TUNE_NSITES] = {-1, -1};
int g_gen = 0;                       // table generation (bqd2_autotune_reset)

hipEvent_t take_event() {
  if (!g_events.empty()) {
    hipEvent_t e = g_events.back();
    g_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void settle(Entry &E) {
  if (E.chosen >= 0) return;
  const int nc = kCands[E.site];
  int best = 0;
  for (int c = 0; c < nc; c++) {
    if (E.n[c] < kSamples) return;
    if (E.best[c] > E.best[best]) best = c;
  }
  E.chosen = best;
}

// read finished timings (caller holds g_mu); wait = true blocks on each pair
void drain(bool wait) {
  for (Pending &p : g_pending) {
    if (!p.e0 || !p.e1) continue;
    hipError_t q = wait ? hipEventSynchronize(p.e1) : hipEventQuery(p.e1);
    if (q == hipErrorNotReady) continue;
    float ms = 0.0f;
    if (q == hipSuccess && p.gen == g_gen && hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
      // timings of a settled class (calls issued while it was explored) are
      // dropped: the choice stays the fastest of the timings it was made on
      Entry &E = g_entries[p.entry];
      const float rate = (float)(p.work / (ms > 1e-6f ? ms : 1e-6f));
      if (E.chosen < 0) {
        if (E.n[p.cand] == 0 || rate > E.best[p.cand]) E.best[p.cand] = rate;
        E.n[p.cand]++;
        settle(E);
      }
    }
    g_events.push_back(p.e0);
    g_events.push_back(p.e1);
    p.e0 = p.e1 = nullptr;
  }
}

int find_entry(int dev, int site, long long key) {
  for (size_t i = 0; i < g_entries.size(); i++)
    if (g_entries[i].dev == dev && g_entries[i].site == site && g_entries[i].key == key) return (int)i;
  Entry E{};
  E.dev = dev;
  E.site = site;
  E.key = key;
  E.chosen = -1;
  g_entries.push_back(E);
  return (int)g_entries.size() - 1;
}

}  // namespace

int tune_class(long long n) {
  if (n < 1) n = 1;
  return (int)std::floor(2.0 * std::log2((double)n));
}

TuneTicket tune_begin(int site, long long key, double work, hipStream_t st) {
  TuneTicket t{0, -1};
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_force[site] >= 0) {
    t.cand = g_force[site];
    return t;
  }
  if (g_mode == 0) return t;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return t;
  drain(false);
  const int ei = find_entry(dev, site, key);
  Entry &E = g_entries[ei];
  if (E.chosen >= 0) {
    t.cand = E.chosen;
    return t;
  }
  int c = 0;
  for (int i = 1; i < kCands[site]; i++)
    if (E.issued[i] < E.issued[c]) c = i;
  // every candidate has its timed calls in flight (the host runs ahead of the
  // device): the default candidate, untimed, until the timings are read
  if (E.issued[c] >= kSamples) return t;
  t.cand = c;
  // time the call when a slot is free; otherwise run it untimed
  int slot = -1;
  for (size_t i = 0; i < g_pending.size(); i++)
    if (!g_pending[i].e0) {
      slot = (int)i;
      break;
    }
  if (slot < 0 && (int)g_pending.size() < kMaxPending) {
    g_pending.push_back(Pending{0, 0, 0, 0.0, nullptr, nullptr});
    slot = (int)g_pending.size() - 1;
  }
  if (slot < 0) return t;
  hipEvent_t e0 = take_event();
  if (!e0 || hipEventRecord(e0, st) != hipSuccess) {
    if (e0) g_events.push_back(e0);
    return t;
  }
  E.issued[c]++;
  // e1 is recorded by tune_end; until then drain() passes the slot by
  g_pending[slot] = Pending{ei, c, g_gen, work, e0, nullptr};
  t.slot = slot;
  return t;
}

void tune_end(const TuneTicket &t, hipStream_t st) {
  if (t.slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  Pending &p = g_pending[t.slot];
  hipEvent_t e1 = take_event();
  if (!e1 || hipEventRecord(e1, st) != hipSuccess) {
    if (e1) g_events.push_back(e1);
    g_events.push_back(p.e0);
    p.e0 = nullptr;
    return;
  }
  p.e1 = e1;
}

extern "C" int bqd2_set_autotune(int mode) {
  BQD2_CHECK(mode == 0 || mode == 1, BQD2_E_ARG, "bqd2_set_autotune: mode must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_mu);
  g_mode = mode;
  return BQD2_OK;
}

extern "C" int bqd2_autotune_force(int site, int cand) {
  BQD2_CHECK(site >= 0 && site < TUNE_NSITES, BQD2_E_ARG, "bqd2_autotune_force: no site %d", site);
  BQD2_CHECK(cand >= -1 && cand < kCands[site], BQD2_E_ARG, "bqd2_autotune_force: site %d has no candidate %d",
             site, cand);
  std::lock_guard<std::mutex> lk(g_mu);
  g_force[site] = cand;
  return BQD2_OK;
}

extern "C" int bqd2_autotune_info(int site, int wait, int max, long long *keys, int *chosen, float *best_rate,
                                  int *ncand) {
  BQD2_CHECK(site >= 0 && site < TUNE_NSITES, BQD2_E_ARG, "bqd2_autotune_info: no site %d", site);
  std::lock_guard<std::mutex> lk(g_mu);
  drain(wait != 0);
  if (ncand) *ncand = kCands[site];
  int dev = 0;
  BQD2_HIP(hipGetDevice(&dev));
  int cnt = 0;
  for (const Entry &E : g_entries) {
    if (E.dev != dev || E.site != site) continue;
    if (cnt < max) {
      if (keys) keys[cnt] = E.key;
      if (chosen) chosen[cnt] = E.chosen;
      if (best_rate)
        for (int c = 0; c < 4; c++) best_rate[4 * cnt + c] = (c < kCands[site] && E.n[c] > 0) ? E.best[c] : -1.0f;
    }
    cnt++;
  }
  return cnt;
}

extern "C" int bqd2_autotune_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  drain(true);
  // calls between tune_begin and tune_end keep their slots; their timings are
  // dropped when read (older generation)
  g_entries.clear();
  g_gen++;
  return BQD2_OK;
}
