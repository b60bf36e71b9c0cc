// fot_hip.hpp -- HIP error checking shared by the HIP sources: a failing
// runtime call returns its fot_status (FOT_ERR_OUT_OF_MEMORY or
// FOT_ERR_HIP) with the call and HIP's message as the last error.
#ifndef FOT_HIP_HPP
#define FOT_HIP_HPP

#include <hip/hip_runtime.h>

#include "fot_internal.hpp"

#define FOT_HIP(call)                                                          \
  do {                                                                         \
    const hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess)                                                      \
      return ::fot::fail(e_ == hipErrorOutOfMemory ? FOT_ERR_OUT_OF_MEMORY     \
                                                   : FOT_ERR_HIP,              \
                         "%s: %s", #call, hipGetErrorString(e_));              \
  } while (0)

#endif
