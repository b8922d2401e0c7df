// common.h -- error plumbing shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include "../../../include/bqd2.h"

void bqd2_set_error(const char *fmt, ...);

#define BQD2_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      bqd2_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return BQD2_E_HIP;                                                            \
    }                                                                               \
  } while (0)

#define BQD2_CHECK(cond, code, ...)       \
  do {                                    \
    if (!(cond)) {                        \
      bqd2_set_error(__VA_ARGS__);        \
      return (code);                      \
    }                                     \
  } while (0)

static inline int bqd2_blocks(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// the process-wide progress callback (bqd2_set_progress): whether one is set,
// and a report from the calling thread
bool bqd2_progress_on();
void bqd2_progress(const char *stage, long long done, long long total);
