// Shared host-side helpers of libchfsi: status propagation, thread-local
// error text, HIP checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/chfsi.h"

namespace chfsi {

void set_error(const char *fmt, ...);

#define CHFSI_HIP(...)                                                               \
  do {                                                                               \
    hipError_t err_ = (__VA_ARGS__);                                                      \
    if (err_ != hipSuccess) {                                                        \
      ::chfsi::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #__VA_ARGS__,          \
                         hipGetErrorString(err_));                                   \
      return CHFSI_ERR_HIP;                                                          \
    }                                                                                \
  } while (0)

#define CHFSI_TRY(...)                                                               \
  do {                                                                               \
    chfsi_status st_ = (__VA_ARGS__);                                                     \
    if (st_ != CHFSI_OK) return st_;                                                 \
  } while (0)

#define CHFSI_ARG(cond, ...)                                                         \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      ::chfsi::set_error(__VA_ARGS__);                                               \
      return CHFSI_ERR_ARG;                                                          \
    }                                                                                \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace chfsi
