// tune.h -- online autotuner of launch shapes (tune.hip).
//
// A site is one launch in the library whose tile or workgroup size can vary
// without changing a single bit of its output: every candidate computes each
// output element in the same lane-level order.  The tuner keys a site's calls
// by a size class, times each candidate on the real calls (HIP events around
// the launch, read back later without blocking) as a rate (the call's work per
// millisecond), and after kSamples rates of every candidate keeps the fastest
// for that class on that device.
#pragma once
#include <hip/hip_runtime.h>

enum TuneSite {
  TUNE_CGEMM = 0,  // trailing update of the complex float LU: 3M tiles 64 x 64 (0) or 128 x 64 (1)
  TUNE_GMV = 1,    // GMRES matrix-vector product: 4 (0), 2 (1), 8 (2) or 1 (3) rows per workgroup
  TUNE_NSITES = 2
};

struct TuneTicket {
  int cand;   // candidate to launch
  int slot;   // timing slot, -1 when the call is not timed
};

// size class: half octaves of n (n >= 1)
int tune_class(long long n);
// candidate for this call of (site, key) on the current device; starts timing
// on st when the class is still being explored.  work: the call's flops or
// bytes, which turns its timing into a rate comparable across the class
TuneTicket tune_begin(int site, long long key, double work, hipStream_t st);
// ends the timing started by tune_begin (no-op for untimed calls)
void tune_end(const TuneTicket &t, hipStream_t st);
