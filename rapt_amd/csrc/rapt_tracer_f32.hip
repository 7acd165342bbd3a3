// rapt_tracer_f32.hip -- k_trace instantiated with the field and the
// finite-difference drift terms in float (RAPT_PREC_FP32_FIELD); state, stage
// sums, error norm and step control stay FP64.  A translation unit of its own
// so the two precisions compile in parallel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rapt_trace_kernel.hpp"

using namespace rapt;

template <int M, int K>
static void launch_f32(hipStream_t st, const KArgs& A, int block, int64_t lanes) {
  const int64_t g = (lanes + block - 1) / block;
  if constexpr (K == KIND_GC && M != F_VAR_EARTH_DIPOLE) {
    if (trace_use_sp<M>(A)) {
      hipLaunchKernelGGL((k_trace<M, K, RAPT_PREC_FP32_FIELD, true>), dim3((unsigned)g),
                         dim3(block), 0, st, A);
      return;
    }
  }
  hipLaunchKernelGGL((k_trace<M, K, RAPT_PREC_FP32_FIELD>), dim3((unsigned)g), dim3(block), 0,
                     st, A);
}

template <int K>
static void launch_f32_model(hipStream_t st, int model, const KArgs& A, int block,
                             int64_t lanes) {
  switch (model) {
    case 0: launch_f32<0, K>(st, A, block, lanes); break;
    case 1: launch_f32<1, K>(st, A, block, lanes); break;
    case 2: launch_f32<2, K>(st, A, block, lanes); break;
    case 3: launch_f32<3, K>(st, A, block, lanes); break;
    case 4: launch_f32<4, K>(st, A, block, lanes); break;
    case 5: launch_f32<5, K>(st, A, block, lanes); break;
    case 6: launch_f32<6, K>(st, A, block, lanes); break;
    case 7: launch_f32<7, K>(st, A, block, lanes); break;
    case 8: launch_f32<8, K>(st, A, block, lanes); break;
  }
}

void rapt_launch_trace_f32(hipStream_t st, int model, int kind, const KArgs& A, int block,
                           int64_t lanes) {
  switch (kind) {
    case KIND_PARTICLE: launch_f32_model<KIND_PARTICLE>(st, model, A, block, lanes); break;
    case KIND_GC: launch_f32_model<KIND_GC>(st, model, A, block, lanes); break;
    default: launch_f32_model<KIND_ADAPTIVE>(st, model, A, block, lanes); break;
  }
}
