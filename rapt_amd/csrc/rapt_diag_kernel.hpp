// rapt_diag_kernel.hpp -- per-row trajectory diagnostics on the device (the
// kernel; rapt_diag.hip launches it for the built-in field models,
// rapt_jit.hip compiles it at run time for user-written fields).
//
// rapt_particle_gc: the guiding centre and first invariant of every row of
// Particle trajectories, Particle.guidingcenter() and Particle.mu()
// (rapt/Particle.py:463-482), which call utils.guidingcenter
// (rapt/utils.py:251-326) and utils.magnetic_moment (rapt/utils.py:189-216)
// once per row.  One lane per row: v = p / (gamma m) with gamma m =
// sqrt(m^2 + p.p / c^2); R solves R = r - rho(R), rho(q) = gamma m / (q |B(q)|^2)
// B(q) x v, by the reference's fixed-point iteration (first estimate r - rho(r),
// stop when |R - R_old| / |R| < tol, at most maxiter iterations); then
// v_par = v . B(R) / |B(R)|, speed |v|, and mu = gamma(speed)^2 m (v - v_par)
// (v + v_par) / (2 |B(R)|).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "rapt_hip.h"
#include "rapt_fields.hpp"

namespace rapt {

struct GcDiagArgs {
  FieldC F;
  int64_t n;
  const double* rows;    // [n][7] t, x, y, z, px, py, pz
  const double* mass;    // [n] or [1]
  const double* charge;  // [n] or [1]
  int64_t mq_stride;     // 1: per row, 0: one value for all rows
  double tol;
  int32_t maxiter;
  int32_t pad;
  double* out;           // [n][6] X, Y, Z, v_par, speed, mu
  int32_t* status;       // [n] iterations used, or RAPT_GC_NOCONV
};

template <int M>
__device__ __forceinline__ V3 gyrovector(const FieldC& F, double t, const V3& q, const V3& v,
                                         double gamma, double mass, double charge) {
  const V3 B = fieldB<M>(F, t, q.x, q.y, q.z);
  const double s = gamma * mass / (charge * dot3(B, B));
  const V3 c = cross3(B, v);
  return V3{s * c.x, s * c.y, s * c.z};
}

template <int M>
__global__ __launch_bounds__(256) void k_particle_gc(GcDiagArgs A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const FieldC& F = A.F;
  const double* row = A.rows + 7 * i;
  const double mass = A.mass[i * A.mq_stride], charge = A.charge[i * A.mq_stride];
  const double t = row[0];
  const V3 r{row[1], row[2], row[3]};
  const V3 p{row[4], row[5], row[6]};
  const double gm = __builtin_sqrt(mass * mass + dot3(p, p) / C_LIGHT_SQ);
  const V3 v{p.x / gm, p.y / gm, p.z / gm};
  const double vsq = dot3(v, v);
  const double gamma = 1.0 / __builtin_sqrt(1.0 - vsq / C_LIGHT_SQ);
  V3 g = gyrovector<M>(F, t, r, v, gamma, mass, charge);
  V3 gold{r.x - g.x, r.y - g.y, r.z - g.z};
  double* o = A.out + 6 * i;
  for (int32_t it = 1; it <= A.maxiter; ++it) {
    g = gyrovector<M>(F, t, gold, v, gamma, mass, charge);
    const V3 gc{r.x - g.x, r.y - g.y, r.z - g.z};
    const V3 dd{gc.x - gold.x, gc.y - gold.y, gc.z - gold.z};
    if (__builtin_sqrt(dot3(dd, dd)) / __builtin_sqrt(dot3(gc, gc)) < A.tol) {
      const V3 B = fieldB<M>(F, t, gc.x, gc.y, gc.z);
      const double Bn = __builtin_sqrt(dot3(B, B));
      const double vp = dot3(v, B) / Bn;
      const double spd = __builtin_sqrt(vsq);
      const double g2 = 1.0 / __builtin_sqrt(1.0 - (spd / C_LIGHT) * (spd / C_LIGHT));
      o[0] = gc.x;
      o[1] = gc.y;
      o[2] = gc.z;
      o[3] = vp;
      o[4] = spd;
      o[5] = g2 * g2 * mass * (spd - vp) * (spd + vp) / (2.0 * Bn);
      A.status[i] = it;
      return;
    }
    gold = gc;
  }
  for (int k = 0; k < 6; ++k) o[k] = __builtin_nan("");
  A.status[i] = RAPT_GC_NOCONV;
}

}  // namespace rapt
