// status.hpp -- error reporting shared by the translation units of libgng_pm.so.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gng_pm.h"

namespace gng {
// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string& msg);
}  // namespace gng

#define GNG_HIP(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return ::gng::fail(e_ == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,                  \
                         std::string(#call) + ": " + hipGetErrorString(e_));                     \
  } while (0)
