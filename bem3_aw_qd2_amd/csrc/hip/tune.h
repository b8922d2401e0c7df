// tune.h -- online autotuner of launch shapes (tune.hip).
//
// A site is one launch in the library whose tile or workgroup size can vary
// without changing a single bit of its output: every candidate computes each
// output element in the same lane-level order.  The tuner keys a site's calls
// by a size class, times each candidate on the real calls (HIP events around
// the launch, read back later without blocking), and after kSamples timings of
// every candidate keeps the fastest for that class on that device.
#pragma once
#include <hip/hip_runtime.h>

enum TuneSite {
  TUNE_CGEMM = 0,  // trailing update of the complex float LU: tile 128 (0) or 64 (1)
  TUNE_TRSM = 1,   // U12 = L11^-1 A12: L11 in LDS, 256 columns per workgroup (0); L11 read
                   // through the scalar cache, 64 (1) or 256 (2) columns per workgroup
  TUNE_GMV = 2,    // GMRES matrix-vector product: 4 (0), 2 (1), 8 (2) or 1 (3) rows per workgroup
  TUNE_NSITES = 3
};

struct TuneTicket {
  int cand;   // candidate to launch
  int slot;   // timing slot, -1 when the call is not timed
};

// size class: half octaves of n (n >= 1)
int tune_class(long long n);
// candidate for this call of (site, key) on the current device; starts timing
// on st when the class is still being explored
TuneTicket tune_begin(int site, long long key, hipStream_t st);
// ends the timing started by tune_begin (no-op for untimed calls)
void tune_end(const TuneTicket &t, hipStream_t st);
