// rapt_internal.hpp -- host-side context shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_kargs.hpp"

struct rapt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  rapt::FieldC field{};
  bool field_set = false;
  rapt::GridC grid{};         // uploaded by rapt_grid_set (device memory below)
  double* grid_mem = nullptr;    // axes and point-layout values
  double* grid_cells = nullptr;  // cell-layout values (nullptr: point layout)
  bool grid_set = false;
  rapt::ExprC expr{};            // uploaded by rapt_field_expr_set (device memory below)
  void* expr_mem = nullptr;
  bool expr_set = false;
  // the same program's instruction words on the host, how the tracer kernels
  // evaluate it (rapt_field_expr_mode), and its run-time compiled one-lane
  // tracer kernels once loaded (rapt_jit.hip)
  std::vector<int32_t> expr_code;
  int32_t expr_nb = 0, expr_ne = 0;
  int32_t expr_mode = RAPT_EXPR_COMPILED;
  hipFunction_t jit_fn[4] = {};  // Particle, GC, GC static TCB, Adaptive
  // device address of this device's allocation-free scratch (the work-queue
  // and step counters of rapt_trace_device_noalloc), from hipGetSymbolAddress
  unsigned long long* noalloc_scratch = nullptr;
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

// run-time compiled tracer kernels of a user-written field (rapt_jit.hip)
int rapt_jit_prepare(rapt_ctx* c, int kind, bool sp);
bool rapt_jit_launch(rapt_ctx* c, int kind, bool sp, const rapt::KArgs& A, int block,
                     int64_t lanes);

#define HIP_OK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err((ctx), RAPT_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)
