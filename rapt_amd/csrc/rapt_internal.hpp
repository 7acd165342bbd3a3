// rapt_internal.hpp -- host-side context shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include <string>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"

struct rapt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  rapt::FieldC field{};
  bool field_set = false;
  std::string err;
};


static inline int set_err(rapt_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIP_OK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err((ctx), RAPT_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)
