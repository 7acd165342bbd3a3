// rapt_kargs.hpp -- the tracer kernels' argument block and the kind/flag
// constants, shared by the library's translation units and by the tracer
// kernels compiled at run time for user-written fields (no host headers).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "rapt_fields.hpp"

namespace rapt {

constexpr int KIND_PARTICLE = 0;
constexpr int KIND_GC = 1;
constexpr int KIND_ADAPTIVE = 2;

constexpr int FLAG_STARTED = 1;   // advance() in progress (state slots valid)
constexpr int FLAG_GCMODE = 2;    // ADAPTIVE: current segment is a GuidingCenter
constexpr int FLAG_AINIT = 4;     // ADAPTIVE: initial mode not decided yet
constexpr int FLAG_FINAL = 8;     // advance() ended (any status but CAP): a relaunch skips it

struct CtlParams {  // scipy set_integrator knobs of one stepper
  double safe, facc1, facc2, beta, expo1;
};

struct KArgs {
  FieldC F;
  int32_t kind, eom, stepper, nmax;
  int32_t eq, check, width, pad;
  double rtol, atol, gc_rtol, gc_atol;
  CtlParams c853, c5;
  double max_step, first_step;
  double rkf_tol, rkf_hmax, rkf_hmin;
  double cycres, gcdt, epss, epst, delta;
  int64_t n;
  double* state;
  int32_t* flags;
  int32_t* status;
  int64_t* nrows;
  int64_t* naccpt;
  const int64_t* row_off;
  const int64_t* row_cap;
  double* rows;
  unsigned long long* work;  // [0] next work item, [1..5] counters; grouped GC
                             // kernels: [0], [6], [7] next item of queues 0, 1, 2
  const int64_t* order;      // work item -> particle index (nullptr: identity)
  int64_t split, split2;     // grouped GC kernels: work items [0, split) form queue 0,
                             // [split, split2) queue 1, [split2, n) queue 2
};

// The launch's kernel-argument block, re-addressed at a point of use (device
// code of the tracer kernels, which take KArgs by value as their first
// argument, so it starts at the kernarg segment pointer).  Arguments needed
// only when a particle starts, finishes or writes a row are read through
// this pointer, made opaque by an empty asm statement, so each load is
// issued where it is used.  Read through the by-value argument, every such
// load is hoisted to the kernel's entry and all of them stay in scalar
// registers for the whole launch, which spills them into VGPR lanes and
// reloads them with v_readlane inside the step loop.
using KArgsK = const __attribute__((address_space(4))) KArgs;
__device__ __forceinline__ KArgsK* cold() {
  KArgsK* p = (KArgsK*)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return p;
}

}  // namespace rapt
