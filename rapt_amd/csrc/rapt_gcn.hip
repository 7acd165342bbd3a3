// rapt_gcn.hip -- launches of the lane-group GuidingCenter kernels
// (rapt_gcn_kernel.hpp) for the built-in field models.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rapt_hip.h"
#include "rapt_gcn_kernel.hpp"
#include "rapt_internal.hpp"

using namespace rapt;

// the TCB + static-field instantiation when it applies (VarEarthDipole is
// never static)
template <int M>
static bool use_sp(const KArgs& A) {
  if constexpr (M == F_VAR_EARTH_DIPOLE) {
    return false;
  } else {
    return A.eom == EOM_TCB && A.F.is_static != 0;
  }
}

template <int M, int L>
static void launch_gcn_m(hipStream_t st, const KArgs& A, int block, int64_t groups) {
  const dim3 grid(gcn_grid(L, block, groups, 0, 0).blocks);
  if constexpr (M != F_VAR_EARTH_DIPOLE) {
    if (use_sp<M>(A)) {
      hipLaunchKernelGGL((k_trace_gcn<M, L, true>), grid, dim3(block), 0, st, A);
      return;
    }
  }
  hipLaunchKernelGGL((k_trace_gcn<M, L, false>), grid, dim3(block), 0, st, A);
}

template <int L>
static void launch_gcn_l(hipStream_t st, int model, const KArgs& A, int block, int64_t groups) {
  switch (model) {
    case 0: launch_gcn_m<0, L>(st, A, block, groups); break;
    case 1: launch_gcn_m<1, L>(st, A, block, groups); break;
    case 2: launch_gcn_m<2, L>(st, A, block, groups); break;
    case 3: launch_gcn_m<3, L>(st, A, block, groups); break;
    case 4: launch_gcn_m<4, L>(st, A, block, groups); break;
    case 5: launch_gcn_m<5, L>(st, A, block, groups); break;
    case 6: launch_gcn_m<6, L>(st, A, block, groups); break;
    case 7: launch_gcn_m<7, L>(st, A, block, groups); break;
    case 8: launch_gcn_m<8, L>(st, A, block, groups); break;
    case 9: launch_gcn_m<9, L>(st, A, block, groups); break;
  }
}

template <int M>
static void launch_mix_m(hipStream_t st, const KArgs& A, int block, int64_t octets,
                         int64_t quads, int64_t pairs) {
  const GcnGrid g = gcn_grid(5, block, octets, quads, pairs);
  const dim3 grid(g.blocks);
  if constexpr (M != F_VAR_EARTH_DIPOLE) {
    if (use_sp<M>(A)) {
      hipLaunchKernelGGL((k_trace_gcmix<M, true>), grid, dim3(block), 0, st, A, g.oct_blocks,
                         g.wide_blocks);
      return;
    }
  }
  hipLaunchKernelGGL((k_trace_gcmix<M, false>), grid, dim3(block), 0, st, A, g.oct_blocks,
                     g.wide_blocks);
}

// Single layouts (L = 2, 4, 8): `groups` resident groups of L lanes.  Mixed
// layouts (lanes_per_particle 3, 5): `groups` octets, `quads` quads and
// `pairs` pairs.
void rapt_launch_gcn(hipStream_t st, int model, int lanes_per_particle, const KArgs& A,
                     int block, int64_t groups, int64_t quads, int64_t pairs) {
  if (lanes_per_particle == 3 || lanes_per_particle == 5) {
    switch (model) {
      case 0: launch_mix_m<0>(st, A, block, groups, quads, pairs); break;
      case 1: launch_mix_m<1>(st, A, block, groups, quads, pairs); break;
      case 2: launch_mix_m<2>(st, A, block, groups, quads, pairs); break;
      case 3: launch_mix_m<3>(st, A, block, groups, quads, pairs); break;
      case 4: launch_mix_m<4>(st, A, block, groups, quads, pairs); break;
      case 5: launch_mix_m<5>(st, A, block, groups, quads, pairs); break;
      case 6: launch_mix_m<6>(st, A, block, groups, quads, pairs); break;
      case 7: launch_mix_m<7>(st, A, block, groups, quads, pairs); break;
      case 8: launch_mix_m<8>(st, A, block, groups, quads, pairs); break;
      case 9: launch_mix_m<9>(st, A, block, groups, quads, pairs); break;
    }
  } else if (lanes_per_particle == 2) {
    launch_gcn_l<2>(st, model, A, block, groups);
  } else if (lanes_per_particle == 8) {
    launch_gcn_l<8>(st, model, A, block, groups);
  } else {
    launch_gcn_l<4>(st, model, A, block, groups);
  }
}
