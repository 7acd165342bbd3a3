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

// kernels of a user-written field compiled at run time (rapt_jit.hip)
#define RAPT_JIT_SLOTS 19
#define RAPT_JIT_HALFBOUNCE 12  // field-line kernels among them
#define RAPT_JIT_LABELS 13
#define RAPT_JIT_BC 14
#define RAPT_JIT_BC_EQ 15
#define RAPT_JIT_FIELD_EVAL 16  // field evaluation, right-hand sides, diagnostics
#define RAPT_JIT_RHS_EVAL 17
#define RAPT_JIT_PARTICLE_GC 18

// Grid of a lane-group GuidingCenter launch: single layouts (L = 2, 4, 8)
// `groups` resident groups of L lanes; mixed layouts (lanes_per_particle 3,
// 5) `groups` octets, `quads` quads and `pairs` pairs, whose blocks are
// [0, oct_blocks) octets, [oct_blocks, wide_blocks) quads, the rest pairs.
struct GcnGrid {
  unsigned blocks;
  int32_t oct_blocks, wide_blocks;
};
static inline GcnGrid gcn_grid(int lanes_per_particle, int block, int64_t groups, int64_t quads,
                               int64_t pairs) {
  if (lanes_per_particle == 3 || lanes_per_particle == 5) {
    const int64_t ob = (8 * groups + block - 1) / block, qb = (4 * quads + block - 1) / block,
                  pb = (2 * pairs + block - 1) / block;
    return GcnGrid{(unsigned)(ob + qb + pb), (int32_t)ob, (int32_t)(ob + qb)};
  }
  const int L = lanes_per_particle == 2 ? 2 : (lanes_per_particle == 8 ? 8 : 4);
  return GcnGrid{(unsigned)((groups * L + block - 1) / block), 0, 0};
}

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
  hipFunction_t jit_fn[RAPT_JIT_SLOTS] = {};  // one lane: Particle, GC, GC static TCB,
                                              // Adaptive; then the GC lane-group kernels
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
int rapt_jit_prepare_gcn(rapt_ctx* c, int lanes_per_particle, bool sp);
bool rapt_jit_launch(rapt_ctx* c, int kind, bool sp, const rapt::KArgs& A, int block,
                     int64_t lanes);
bool rapt_jit_launch_gcn(rapt_ctx* c, int lanes_per_particle, bool sp, const rapt::KArgs& A,
                         int block, int64_t groups, int64_t quads, int64_t pairs);
int rapt_jit_run_slot(rapt_ctx* c, int slot, unsigned grid, unsigned block, void** args);

#define HIP_OK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err((ctx), RAPT_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)
