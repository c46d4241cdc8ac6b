// Error plumbing for the C-ABI: every entry point returns 0 on success and a
// non-zero status otherwise; mra_last_error() returns the message of the most
// recent failure on the calling host thread.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

namespace mra {
void set_last_error(const char* msg);
}

#define MRA_TRY_BEGIN {
#define MRA_TRY_END \
  }                 \
  return 0;
#define MRA_FAIL(msg)              \
  do {                             \
    ::mra::set_last_error(msg);    \
    return 1;                      \
  } while (0)
#define MRA_CHECK_LAUNCH(what)                                               \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess) {                                                  \
      char b_[256];                                                          \
      snprintf(b_, sizeof(b_), "%s: launch failed: %s", what, hipGetErrorString(e_)); \
      ::mra::set_last_error(b_);                                             \
      return 2;                                                              \
    }                                                                        \
  } while (0)
