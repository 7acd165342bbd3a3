// rapt_gcn_kernel.hpp -- GuidingCenter.advance with L = 2, 4 or 8 lanes per
// guiding centre: the kernels k_trace_gcn<M, L, SP> and k_trace_gcmix<M, SP>,
// shared by rapt_gcn.hip (the built-in field models) and the kernels that
// rapt_jit.hip compiles at run time for user-written fields.
//
// A group of L adjacent lanes carries one particle.  For L = 2 and 4 (inside
// one DPP quad) lane g of the group holds NC = 4 / L components of the state
// (x, y, z, ppar) and of every DOPRI5 stage -- components g*NC .. g*NC+NC-1 --
// so the stage combinations are split L ways; for L = 8 both quads of the
// group hold the four components as a quad does.  For a right-hand side the
// group gathers the stage vector with quad-permute DPP moves, then each lane
// evaluates the field at 8 / L points of the finite-difference stencil:
//
//   L = 8: lane 0 the centre, lane k in 1..3 +d on axis k, lane 7 - k -d on
//          axis k, lane 7 the centre again (row_half_mirror pairs lane i with
//          lane 7 - i, so each lane swaps its point with its pair's);
//   L = 4: lane 0 the centre (twice), lane k the +-d pair on axis k;
//   L = 2: lane 0 the centre, +-x and the centre again; lane 1 +-y and +-z.
//
// For a time-dependent field lane 0's spare evaluation (lane 7's for L = 8)
// is the centre at t - tstep.  Each lane forms its share of grad|B| and
// curl b, the shares are broadcast back, and every lane runs the shared
// guiding-centre algebra and keeps its own output components.  All
// reductions (error norm, HINIT norms) are summed in one fixed order from
// broadcast values, so the lanes of a group stay bit-identical, take the
// same branches, and the trajectories are bit-identical to the one-lane
// kernel's.  Driver semantics are those of k_trace<M, KIND_GC>
// (rapt_tracer.hip): rapt/GuidingCenter.py:442-458 with scipy's dopri5 call
// semantics.
//
// The layouts trade issue slots against latency: a wide group's step is
// short (L = 8 evaluates one stencil point per lane) but repeats the algebra
// in every lane; L = 2 issues about 30% fewer instructions per particle-step
// than L = 4, with longer per-lane chains.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_kargs.hpp"
#include "rapt_physics.hpp"
#include "rapt_steppers.hpp"

namespace rapt {


// quad_perm control that reads lane I of this lane's group of L lanes
// (2 bits per destination lane, lane 0 in bits 0-1)
template <int L, int I>
constexpr int grp_ctrl() {
  static_assert(L == 2 || L == 4, "two or four lanes per group");
  if constexpr (L == 4) {
    return I * 0x55;  // [I, I, I, I]
  } else {
    return I | (I << 2) | ((2 + I) << 4) | ((2 + I) << 6);  // [I, I, 2+I, 2+I]
  }
}

// value of lane I of this lane's group
template <int L, int I>
__device__ __forceinline__ double gb(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, grp_ctrl<L, I>(), 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), grp_ctrl<L, I>(), 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int L, int I>
__device__ __forceinline__ int gbi(int v) {
  return __builtin_amdgcn_mov_dpp(v, grp_ctrl<L, I>(), 0xF, 0xF, false);
}

// this lane's NC components of a four-component vector
template <int NC>
struct CV {
  double c[NC];
};
// components per lane of an L-lane group (a quad's layout for L = 8)
template <int L>
constexpr int ncomp() {
  return L >= 4 ? 1 : 4 / L;
}
template <int L>
using CVL = CV<ncomp<L>()>;

// value of lane 7 - i of this lane's half row of eight lanes (row_half_mirror)
__device__ __forceinline__ double gmir(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x141, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x141, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
// lanes 4-7 of each group of eight take the value of lanes 0-3 (row_shr:4
// written to DPP banks 1 and 3 only); lanes 0-3 keep their own
__device__ __forceinline__ int gupi(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xA, false);
}
__device__ __forceinline__ double gup(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = gupi((int)b), hi = gupi((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// lane 0's value in every lane of the group
template <int L>
__device__ __forceinline__ long long gbll0(long long v) {
  constexpr int LQ = L == 8 ? 4 : L;
  int lo = gbi<LQ, 0>((int)v), hi = gbi<LQ, 0>((int)(v >> 32));
  if constexpr (L == 8) {
    lo = gupi(lo);
    hi = gupi(hi);
  }
  return ((long long)hi << 32) | (long long)(unsigned)lo;
}

// the four components of the group's vector, in component order (for L = 8
// each quad gathers its own copy)
template <int L>
__device__ __forceinline__ void gather4(const CVL<L>& v, double out[4]) {
  if constexpr (L >= 4) {
    out[0] = gb<4, 0>(v.c[0]);
    out[1] = gb<4, 1>(v.c[0]);
    out[2] = gb<4, 2>(v.c[0]);
    out[3] = gb<4, 3>(v.c[0]);
  } else {
    out[0] = gb<2, 0>(v.c[0]);
    out[1] = gb<2, 0>(v.c[1]);
    out[2] = gb<2, 1>(v.c[0]);
    out[3] = gb<2, 1>(v.c[1]);
  }
}
// sum of squares of the four components, in component order
template <int L>
__device__ __forceinline__ double grp_sumsq(const CVL<L>& v) {
  double a[4];
  gather4<L>(v, a);
  return __builtin_fma(a[3], a[3], __builtin_fma(a[2], a[2], __builtin_fma(a[1], a[1], a[0] * a[0])));
}
// 1 when any component of the group's vector is not finite
template <int L>
__device__ __forceinline__ bool grp_nonfinite(const CVL<L>& v) {
  int f = 0;
#pragma unroll
  for (int i = 0; i < ncomp<L>(); ++i) f |= __builtin_isfinite(v.c[i]) ? 0 : 1;
  if constexpr (L >= 4) {
    return (gbi<4, 0>(f) | gbi<4, 1>(f) | gbi<4, 2>(f) | gbi<4, 3>(f)) != 0;
  } else {
    return (gbi<2, 0>(f) | gbi<2, 1>(f)) != 0;
  }
}

// |B| and b = B/|B| at one stencil point, as the one-lane stencil forms them
__device__ __forceinline__ void mag_unit(const V3& B, double& m, V3& u) {
  const double bb = dot3(B, B);
  const double ib = rsqrt_nr(bb);
  m = bb * ib;
  u = V3{B.x * ib, B.y * ib, B.z * ib};
}

// Stencil offsets of this lane's field evaluations, added to the centre as
// the one-lane stencil adds them (coordinate + (+-d) on the axis, + 0.0 off
// it).  The centre itself is coordinate + (-0.0), which is the coordinate
// exactly (-0.0 is the identity of IEEE addition), so no lane selects.
template <int L>
struct StencilOff {
  V3 o[8 / L];
};
template <int L>
__device__ __forceinline__ StencilOff<L> stencil_offsets(int q, double d) {
  const V3 ctr{-0.0, -0.0, -0.0};
  StencilOff<L> s;
  if constexpr (L == 8) {
    // lanes 0 and 7: the centre; lane k in 1..3: +d on axis k; lane 7 - k: -d
    if (q == 0 || q == 7) {
      s.o[0] = ctr;
    } else if (q < 4) {
      s.o[0] = V3{q == 1 ? d : 0.0, q == 2 ? d : 0.0, q == 3 ? d : 0.0};
    } else {
      s.o[0] = V3{q == 6 ? -d : 0.0, q == 5 ? -d : 0.0, q == 4 ? -d : 0.0};
    }
  } else if constexpr (L == 4) {
    // lane 0: the centre twice; lane k: +d and -d on axis k
    if (q == 0) {
      s.o[0] = ctr;
      s.o[1] = ctr;
    } else {
      s.o[0] = V3{q == 1 ? d : 0.0, q == 2 ? d : 0.0, q == 3 ? d : 0.0};
      s.o[1] = V3{q == 1 ? -d : 0.0, q == 2 ? -d : 0.0, q == 3 ? -d : 0.0};
    }
  } else {
    // lane 0: centre, +x, -x, centre; lane 1: +y, -y, +z, -z
    if (q == 0) {
      s.o[0] = ctr;
      s.o[1] = V3{d, 0.0, 0.0};
      s.o[2] = V3{-d, 0.0, 0.0};
      s.o[3] = ctr;
    } else {
      s.o[0] = V3{0.0, d, 0.0};
      s.o[1] = V3{0.0, -d, 0.0};
      s.o[2] = V3{0.0, 0.0, d};
      s.o[3] = V3{0.0, 0.0, -d};
    }
  }
  return s;
}

// guiding-centre right-hand side, four lanes: component q of the output
template <int M>
__device__ __forceinline__ CV<1> rhs_gc_quad(const FieldC& F, int eom, bool stat, double charge,
                                             double mass, double mu, double v, const GcK& K,
                                             bool eq, int q, const StencilOff<4>& O, double t,
                                             const CV<1>& yv, int64_t& nf) {
  ++nf;
  const double yq = yv.c[0];
  const double Y0 = gb<4, 0>(yq), Y1 = gb<4, 1>(yq), Y2 = gb<4, 2>(yq), Y3 = gb<4, 3>(yq);
  const double i2d = F.i2d;
  // lane 0: the centre (and, for a time-dependent field, the centre at
  // t - tstep); lane k: the +-d points on axis k
  const bool ctr = q == 0;
  const V3 Ba = fieldB<M>(F, t, Y0 + O.o[0].x, Y1 + O.o[0].y, Y2 + O.o[0].z);
  double tb = t;
  if (!stat) tb = ctr ? t - F.tstep : t;
  const V3 Bb = fieldB<M>(F, tb, Y0 + O.o[1].x, Y1 + O.o[1].y, Y2 + O.o[1].z);
  double ma, mb;
  V3 ua, ub_;
  mag_unit(Ba, ma, ua);
  mag_unit(Bb, mb, ub_);
  const V3 du{ua.x - ub_.x, ua.y - ub_.y, ua.z - ub_.z};
  const double g = (ma - mb) * i2d;
  const double e0 = (q == 0) ? Ba.x : g;
  const double e1 = (q == 0) ? Ba.y : ((q == 1) ? du.y : du.x);
  const double e2 = (q == 0) ? Ba.z : ((q == 3) ? du.y : du.z);
  const V3 B{gb<4, 0>(e0), gb<4, 0>(e1), gb<4, 0>(e2)};
  // |B| and b at the centre: lane 0's own mag_unit of Ba
  const double Bm = gb<4, 0>(ma);
  const V3 ub{gb<4, 0>(ua.x), gb<4, 0>(ua.y), gb<4, 0>(ua.z)};
  const V3 gB{gb<4, 1>(e0), gb<4, 2>(e0), gb<4, 3>(e0)};
  const V3 cb = curl_from_pairs(gb<4, 1>(e1), gb<4, 1>(e2), gb<4, 2>(e1), gb<4, 2>(e2),
                                gb<4, 3>(e1), gb<4, 3>(e2), i2d);
  V3 db{0.0, 0.0, 0.0};
  if (eom == EOM_TCB && !stat) {
    // lane 0's second point was B(t - tstep); B(t + tstep) is a third evaluation
    const V3 Bc = fieldB<M>(F, t + F.tstep, Y0, Y1, Y2);
    const double ic = rsqrt_nr(dot3(Bc, Bc));
    const double s = F.ihts;
    const double l0 = (Bc.x * ic - ub_.x) * s, l1 = (Bc.y * ic - ub_.y) * s,
                 l2 = (Bc.z * ic - ub_.z) * s;
    db = V3{gb<4, 0>(l0), gb<4, 0>(l1), gb<4, 0>(l2)};
  }
  const double Y[4] = {Y0, Y1, Y2, Y3};
  double o[4];
  const V3 Ec = fieldE<M>(F, t, Y0, Y1, Y2);
  gc_algebra_b(eom, has_E<M>(), Ec, stat, charge, mass, mu, v, K, eq, Y, B, Bm,
               ub, gB, cb, db, o);
  return CV<1>{{q == 0 ? o[0] : (q == 1 ? o[1] : (q == 2 ? o[2] : o[3]))}};
}

// guiding-centre right-hand side, two lanes: lane q returns components 2q, 2q+1
template <int M>
__device__ __forceinline__ CV<2> rhs_gc_pair(const FieldC& F, int eom, bool stat, double charge,
                                             double mass, double mu, double v, const GcK& K,
                                             bool eq, int q, const StencilOff<2>& O, double t,
                                             const CV<2>& yv, int64_t& nf) {
  ++nf;
  const double Y0 = gb<2, 0>(yv.c[0]), Y1 = gb<2, 0>(yv.c[1]);
  const double Y2 = gb<2, 1>(yv.c[0]), Y3 = gb<2, 1>(yv.c[1]);
  const double i2d = F.i2d;
  const bool l0 = q == 0;
  //   lane 0: centre, +x, -x, centre at t3;  lane 1: +y, -y, +z, -z
  double t3 = t;
  if (!stat) t3 = l0 ? t - F.tstep : t;
  const V3 B0 = fieldB<M>(F, t, Y0 + O.o[0].x, Y1 + O.o[0].y, Y2 + O.o[0].z);
  const V3 B1 = fieldB<M>(F, t, Y0 + O.o[1].x, Y1 + O.o[1].y, Y2 + O.o[1].z);
  const V3 B2 = fieldB<M>(F, t, Y0 + O.o[2].x, Y1 + O.o[2].y, Y2 + O.o[2].z);
  const V3 B3 = fieldB<M>(F, t3, Y0 + O.o[3].x, Y1 + O.o[3].y, Y2 + O.o[3].z);
  double m0, m1, m2, m3;
  V3 u0, u1, u2, u3;
  mag_unit(B0, m0, u0);
  mag_unit(B1, m1, u1);
  mag_unit(B2, m2, u2);
  mag_unit(B3, m3, u3);
  // lane 0 sends (B, d|B|/dx, dxy, dxz); lane 1 sends (d|B|/dy, d|B|/dz,
  // dyx, dyz, dzx, dzy), where dxy = b_y(+x) - b_y(-x) and so on
  const double e0 = l0 ? B0.x : (m0 - m1) * i2d;
  const double e1 = l0 ? B0.y : (m2 - m3) * i2d;
  const double e2 = l0 ? B0.z : u0.x - u1.x;
  const double e3 = l0 ? (m1 - m2) * i2d : u0.z - u1.z;
  const double e4 = l0 ? u1.y - u2.y : u2.x - u3.x;
  const double e5 = l0 ? u1.z - u2.z : u2.y - u3.y;
  const V3 B{gb<2, 0>(e0), gb<2, 0>(e1), gb<2, 0>(e2)};
  // |B| and b at the centre: lane 0's own mag_unit of B0
  const double Bm = gb<2, 0>(m0);
  const V3 ub{gb<2, 0>(u0.x), gb<2, 0>(u0.y), gb<2, 0>(u0.z)};
  const V3 gB{gb<2, 0>(e3), gb<2, 1>(e0), gb<2, 1>(e1)};
  const V3 cb = curl_from_pairs(gb<2, 0>(e4), gb<2, 0>(e5), gb<2, 1>(e2), gb<2, 1>(e3),
                                gb<2, 1>(e4), gb<2, 1>(e5), i2d);
  V3 db{0.0, 0.0, 0.0};
  if (eom == EOM_TCB && !stat) {
    // lane 0's fourth point was B(t - tstep); B(t + tstep) is a fifth evaluation
    const V3 Bc = fieldB<M>(F, t + F.tstep, Y0, Y1, Y2);
    const double ic = rsqrt_nr(dot3(Bc, Bc));
    const double s = F.ihts;
    const double l0x = (Bc.x * ic - u3.x) * s, l0y = (Bc.y * ic - u3.y) * s,
                 l0z = (Bc.z * ic - u3.z) * s;
    db = V3{gb<2, 0>(l0x), gb<2, 0>(l0y), gb<2, 0>(l0z)};
  }
  const double Y[4] = {Y0, Y1, Y2, Y3};
  double o[4];
  const V3 Ec = fieldE<M>(F, t, Y0, Y1, Y2);
  gc_algebra_b(eom, has_E<M>(), Ec, stat, charge, mass, mu, v, K, eq, Y, B, Bm,
               ub, gB, cb, db, o);
  return CV<2>{{l0 ? o[0] : o[2], l0 ? o[1] : o[3]}};
}

// guiding-centre right-hand side, eight lanes: component q & 3 of the output.
// Both quads hold the state, so each gathers it with quad moves.  Every lane
// evaluates one stencil point and swaps |B| and b with its mirror lane, which
// gives lanes 0-3 exactly the pairs a lane quad evaluates itself (lane k: +d
// and -d on axis k; lane 0: the centre and, for a time-dependent field, the
// centre at t - tstep); the lower quad runs the quad's broadcasts and
// algebra, and the upper quad takes its output.
template <int M>
__device__ __forceinline__ CV<1> rhs_gc_oct(const FieldC& F, int eom, bool stat, double charge,
                                            double mass, double mu, double v, const GcK& K,
                                            bool eq, int q, const StencilOff<8>& O, double t,
                                            const CV<1>& yv, int64_t& nf) {
  ++nf;
  const double yq = yv.c[0];
  const double Y0 = gb<4, 0>(yq), Y1 = gb<4, 1>(yq), Y2 = gb<4, 2>(yq), Y3 = gb<4, 3>(yq);
  const double i2d = F.i2d;
  double ta = t;
  if (!stat) ta = q == 7 ? t - F.tstep : t;
  const V3 Ba = fieldB<M>(F, ta, Y0 + O.o[0].x, Y1 + O.o[0].y, Y2 + O.o[0].z);
  double ma;
  V3 ua;
  mag_unit(Ba, ma, ua);
  const double mb = gmir(ma);
  const V3 ub_{gmir(ua.x), gmir(ua.y), gmir(ua.z)};
  const V3 du{ua.x - ub_.x, ua.y - ub_.y, ua.z - ub_.z};
  const double g = (ma - mb) * i2d;
  const double e0 = (q == 0) ? Ba.x : g;
  const double e1 = (q == 0) ? Ba.y : ((q == 1) ? du.y : du.x);
  const double e2 = (q == 0) ? Ba.z : ((q == 3) ? du.y : du.z);
  const V3 B{gb<4, 0>(e0), gb<4, 0>(e1), gb<4, 0>(e2)};
  const double Bm = gb<4, 0>(ma);
  const V3 ub{gb<4, 0>(ua.x), gb<4, 0>(ua.y), gb<4, 0>(ua.z)};
  const V3 gB{gb<4, 1>(e0), gb<4, 2>(e0), gb<4, 3>(e0)};
  const V3 cb = curl_from_pairs(gb<4, 1>(e1), gb<4, 1>(e2), gb<4, 2>(e1), gb<4, 2>(e2),
                                gb<4, 3>(e1), gb<4, 3>(e2), i2d);
  V3 db{0.0, 0.0, 0.0};
  if (eom == EOM_TCB && !stat) {
    // lane 0's partner point was B(t - tstep); B(t + tstep) is a second evaluation
    const V3 Bc = fieldB<M>(F, t + F.tstep, Y0, Y1, Y2);
    const double ic = rsqrt_nr(dot3(Bc, Bc));
    const double s = F.ihts;
    const double l0 = (Bc.x * ic - ub_.x) * s, l1 = (Bc.y * ic - ub_.y) * s,
                 l2 = (Bc.z * ic - ub_.z) * s;
    db = V3{gb<4, 0>(l0), gb<4, 0>(l1), gb<4, 0>(l2)};
  }
  const double Y[4] = {Y0, Y1, Y2, Y3};
  double o[4];
  const V3 Ec = fieldE<M>(F, t, Y0, Y1, Y2);
  gc_algebra_b(eom, has_E<M>(), Ec, stat, charge, mass, mu, v, K, eq, Y, B, Bm, ub, gB, cb, db,
               o);
  const int qc = q & 3;
  return CV<1>{{gup(qc == 0 ? o[0] : (qc == 1 ? o[1] : (qc == 2 ? o[2] : o[3])))}};
}

template <int M, int L>
__device__ __forceinline__ CVL<L> rhs_gcn(const FieldC& F, int eom, bool stat, double charge,
                                          double mass, double mu, double v, const GcK& K,
                                          bool eq, int q, const StencilOff<L>& O, double t,
                                          const CVL<L>& y, int64_t& nf) {
  if constexpr (L == 8) {
    return rhs_gc_oct<M>(F, eom, stat, charge, mass, mu, v, K, eq, q, O, t, y, nf);
  } else if constexpr (L == 4) {
    return rhs_gc_quad<M>(F, eom, stat, charge, mass, mu, v, K, eq, q, O, t, y, nf);
  } else {
    return rhs_gc_pair<M>(F, eom, stat, charge, mass, mu, v, K, eq, q, O, t, y, nf);
  }
}

// resident waves per SIMD the register allocation targets
#ifndef RAPT_GCN_WAVES
#define RAPT_GCN_WAVES 2
#endif

// The tracer loop of one lane group.  The work queue is split in three:
// queue 0 holds work items [0, A.split), queue 1 [A.split, A.split2) and
// queue 2 [A.split2, n) (counters A.work[0], A.work[6], A.work[7]; a single
// layout has everything in queue 0).  A group takes items from queue qown
// first and from the others, in queue order, once its own is drained; the
// queues hold descending costs, so that is the longest work left.  SP: the
// instantiation for
// the Tao-Chan-Brizard equations in a static field, with those two choices
// compiled in (no runtime branches on them in the right-hand side); the
// generic one reads them from A.eom and F.is_static.
template <int M, int L, bool SP>
__device__ __forceinline__ void gcn_body(const KArgs& A, const int qown) {
  constexpr int NC = ncomp<L>();
  using Y4 = CV<NC>;
  const FieldC& F = A.F;
  const int eom = SP ? (int)EOM_TCB : A.eom;
  const bool stat = SP || F.is_static != 0;
  const int64_t n = cold()->n;
  const int q = threadIdx.x & (L - 1);
  const int c0 = (L == 8 ? (q & 3) : q) * NC;  // first component this lane holds
  const bool owner = L < 8 || q < 4;           // the lane stores its components
  const StencilOff<L> off = stencil_offsets<L>(q, F.gstep);
  GcK gk{};
  bool active = false, exhausted = false;
  int64_t pi = 0;
  double x = 0;
  Y4 y{}, k1{};
  double h = 0, dt = 0, tend = 0, gm = 0, tcur = 0, mass = 0, charge = 0, mu = 0, v = 0;
  double xend = 0, hmax = 0, lfo = LFO0;  // lfo = log2(FACOLD)
  int nstep = 0, nacc_call = 0;
  bool last = false, reject = false, k1ok = false;
  bool drained[3] = {false, false, false};
  int64_t nrow = 0, rowcap = 0, rowbase = 0;
  int64_t c_acc = 0, c_rej = 0, c_stp = 0, c_fcn = 0, p_acc = 0;
  const int W = 5;
  const CtlParams& C = A.c5;
  const bool eq = A.eq != 0;

  auto write_row = [&](double trow) {
    double* r = cold()->rows + (rowbase + nrow) * W;
    if (q == 0) r[0] = trow;
    if (owner) {
#pragma unroll
      for (int i = 0; i < NC; ++i) r[1 + c0 + i] = y.c[i];
    }
    ++nrow;
  };
  auto begin_call = [&]() {
    xend = x + dt;
    hmax = (cold()->max_step == 0.0) ? xend - x : cold()->max_step;
    lfo = LFO0;
    nstep = 0;
    nacc_call = 0;
    last = false;
    reject = false;
  };
  auto finish = [&](int32_t code) {
    double* S = cold()->state;
    if (owner) {
#pragma unroll
      for (int i = 0; i < NC; ++i) S[(1 + c0 + i) * n + pi] = y.c[i];
    }
    if (q == 0) {
      S[0 * n + pi] = x;
      S[RAPT_S_MU * n + pi] = mu;
      S[RAPT_S_V * n + pi] = v;
      S[RAPT_S_DT * n + pi] = dt;
      S[RAPT_S_H * n + pi] = h;
      S[RAPT_S_TEND * n + pi] = tend;
      S[RAPT_S_GM * n + pi] = gm;
      S[RAPT_S_TCUR * n + pi] = tcur;
      cold()->flags[pi] = FLAG_STARTED | FLAG_GCMODE | (code != RAPT_ST_CAP ? FLAG_FINAL : 0);
      cold()->status[pi] = code;
      cold()->nrows[pi] = nrow;
      cold()->naccpt[pi] = p_acc;
    }
    active = false;
  };

  for (;;) {
    if (!active && !exhausted) {
      int64_t got = n;
      if (q == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // qown, then the other two in queue order
          const int qq = k == 0 ? qown : (k == 1 ? (qown == 0 ? 1 : 0) : (qown == 2 ? 1 : 2));
          if (got < n || drained[qq]) continue;
          const int64_t lo = qq == 0 ? 0 : (qq == 1 ? cold()->split : cold()->split2);
          const int64_t hi = qq == 0 ? cold()->split : (qq == 1 ? cold()->split2 : n);
          const int64_t t = (int64_t)atomicAdd(&cold()->work[qq == 0 ? 0 : (qq == 1 ? 6 : 7)], 1ULL);
          if (t < hi - lo) {
            got = lo + t;
          } else {
            drained[qq] = true;
          }
        }
        if (got < n && cold()->order) got = cold()->order[got];
      }
      pi = gbll0<L>(got);
      if (pi >= n) {
        exhausted = true;
      } else if (cold()->flags[pi] & FLAG_FINAL) {
        // finished in an earlier launch: no rows this time, status kept
        if (q == 0) {
          cold()->nrows[pi] = 0;
          cold()->naccpt[pi] = 0;
        }
      } else {
        active = true;
        const double* S = cold()->state;
        x = S[0 * n + pi];
#pragma unroll
        for (int i = 0; i < NC; ++i) y.c[i] = S[(1 + c0 + i) * n + pi];
        mass = S[RAPT_S_MASS * n + pi];
        charge = S[RAPT_S_CHARGE * n + pi];
        mu = S[RAPT_S_MU * n + pi];
        v = S[RAPT_S_V * n + pi];
        gk = gc_consts(eom, charge, mass, v, mu);
        dt = S[RAPT_S_DT * n + pi];
        h = S[RAPT_S_H * n + pi];
        tend = S[RAPT_S_TEND * n + pi];
        gm = S[RAPT_S_GM * n + pi];
        tcur = S[RAPT_S_TCUR * n + pi];
        const int32_t fl = cold()->flags[pi];
        nrow = 0;
        rowcap = cold()->row_cap[pi];
        rowbase = cold()->row_off[pi];
        p_acc = 0;
        k1ok = false;
        if (!(fl & FLAG_STARTED)) {
          tend = x + cold()->delta;
          h = cold()->first_step;
          tcur = x;
        }
        if (!(x < tend)) {
          finish(RAPT_ST_DONE);
        } else if (nrow >= rowcap) {
          finish(RAPT_ST_CAP);
        } else {
          begin_call();
        }
      }
    }
    if (__ballot(active || !exhausted) == 0) break;

    if (active && !k1ok) {
      k1 = rhs_gcn<M, L>(F, eom, stat, charge, mass, mu, v, gk, eq, q, off, x, y, c_fcn);
      k1ok = true;
    }
    if (active && h == 0.0) {
      // HINIT with group-summed norms (Hairer's dopri5, iord 5)
      Y4 sk, r1, r2;
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        sk.c[i] = A.atol + A.rtol * __builtin_fabs(y.c[i]);
        r1.c[i] = k1.c[i] / sk.c[i];
        r2.c[i] = y.c[i] / sk.c[i];
      }
      const double dnf = grp_sumsq<L>(r1), dny = grp_sumsq<L>(r2);
      double hh = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : __builtin_sqrt(dny / dnf) * 0.01;
      hh = fmin(hh, hmax);
      Y4 ye;
#pragma unroll
      for (int i = 0; i < NC; ++i) ye.c[i] = y.c[i] + hh * k1.c[i];
      const Y4 f1 =
          rhs_gcn<M, L>(F, eom, stat, charge, mass, mu, v, gk, eq, q, off, x + hh, ye, c_fcn);
#pragma unroll
      for (int i = 0; i < NC; ++i) r1.c[i] = (f1.c[i] - k1.c[i]) / sk.c[i];
      double der2 = grp_sumsq<L>(r1);
      der2 = __builtin_sqrt(der2) / hh;
      const double der12 = fmax(__builtin_fabs(der2), __builtin_sqrt(dnf));
      const double h1 = (der12 <= 1.0e-15) ? fmax(1.0e-6, __builtin_fabs(hh) * 1.0e-3)
                                           : pow(0.01 / der12, 1.0 / 5);
      h = fmin(100.0 * __builtin_fabs(hh), fmin(h1, hmax));
    }
    bool stepping = false;
    if (active) {
      if (nstep > A.nmax || 0.1 * __builtin_fabs(h) <= __builtin_fabs(x) * UROUND) {
        const int32_t code = (nstep > A.nmax) ? RAPT_ST_NSTEPS : RAPT_ST_HSMALL;
        write_row(x);
        tcur = x;
        finish(code);
      } else {
        if (x + 1.01 * h - xend > 0.0) {
          h = xend - x;
          last = true;
        }
        ++nstep;
        stepping = true;
      }
    }
    Y4 y1{}, k7{};
    double err = 0.0;
    double fob = 1.0;
    if (stepping) {
      fob = ctl_fob(C.beta, lfo);
      using namespace dp5;
      Y4 k2{}, k3{}, k4{}, k5{}, k6{};
#pragma unroll 1
      for (int s = 2; s <= 7; ++s) {
        Y4 ts;
        double xs;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
          switch (s) {
            case 2: ts.c[i] = y.c[i] + h * a21 * k1.c[i]; break;
            case 3: ts.c[i] = y.c[i] + h * (a31 * k1.c[i] + a32 * k2.c[i]); break;
            case 4:
              ts.c[i] = y.c[i] + h * (a41 * k1.c[i] + a42 * k2.c[i] + a43 * k3.c[i]);
              break;
            case 5:
              ts.c[i] = y.c[i] + h * (a51 * k1.c[i] + a52 * k2.c[i] + a53 * k3.c[i] +
                                      a54 * k4.c[i]);
              break;
            case 6:
              ts.c[i] = y.c[i] + h * (a61 * k1.c[i] + a62 * k2.c[i] + a63 * k3.c[i] +
                                      a64 * k4.c[i] + a65 * k5.c[i]);
              break;
            default:
              y1.c[i] = y.c[i] + h * (a71 * k1.c[i] + a73 * k3.c[i] + a74 * k4.c[i] +
                                      a75 * k5.c[i] + a76 * k6.c[i]);
              ts.c[i] = y1.c[i];
              break;
          }
        }
        switch (s) {
          case 2: xs = x + c2 * h; break;
          case 3: xs = x + c3 * h; break;
          case 4: xs = x + c4 * h; break;
          case 5: xs = x + c5 * h; break;
          default: xs = x + h; break;
        }
        const Y4 kk = rhs_gcn<M, L>(F, eom, stat, charge, mass, mu, v, gk, eq, q, off, xs, ts, c_fcn);
        if (s == 2) k2 = kk;
        else if (s == 3) k3 = kk;
        else if (s == 4) k4 = kk;
        else if (s == 5) k5 = kk;
        else if (s == 6) k6 = kk;
        else k7 = kk;
      }
      Y4 es;
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        const double e = (e1 * k1.c[i] + e3 * k3.c[i] + e4 * k4.c[i] + e5 * k5.c[i] +
                          e6 * k6.c[i] + e7 * k7.c[i]) * h;
        const double sk = A.atol + A.rtol * fmax(__builtin_fabs(y.c[i]), __builtin_fabs(y1.c[i]));
        es.c[i] = e / sk;
      }
      err = __builtin_sqrt(grp_sumsq<L>(es) * (1.0 / 4));
    }
    if (stepping) {
      bool call_done = false;
      if (q == 0) ++c_stp;
      const double l2e = log2(err);
      const double fac11 = exp2(C.expo1 * l2e);
      double fac = fac11 / fob;
      fac = fmax(C.facc2, fmin(C.facc1, fac / C.safe));
      double hnew = h / fac;
      if (err <= 1.0) {
        lfo = fmax(l2e, LFO0);
        ++nacc_call;
        if (q == 0) ++c_acc;
        ++p_acc;
        k1 = k7;
        y = y1;
        x = x + h;
        if (last) {
          h = hnew;
          call_done = true;
        } else {
          if (__builtin_fabs(hnew) > hmax) hnew = hmax;
          if (reject) hnew = fmin(__builtin_fabs(hnew), __builtin_fabs(h));
          reject = false;
          h = hnew;
        }
      } else {
        hnew = h / fmin(C.facc1, fac11 / C.safe);
        reject = true;
        if (nacc_call >= 1 && q == 0) ++c_rej;
        last = false;
        h = hnew;
      }
      if (call_done) {
        write_row(x);
        tcur = x;
        if (grp_nonfinite<L>(y)) {
          finish(RAPT_ST_NAN);
        } else {
          bool sw = false;
          if (cold()->check) {
            double yy[4];
            gather4<L>(y, yy);
            sw = !gc_isadiabatic<M>(F, cold()->epss, cold()->epst, x, yy, mass, charge, mu);
          }
          if (sw) {
            finish(RAPT_ST_SWITCH);
          } else if (x < tend) {
            if (nrow >= rowcap) {
              finish(RAPT_ST_CAP);
            } else {
              begin_call();
            }
          } else {
            finish(RAPT_ST_DONE);
          }
        }
      }
    }
  }
  if (q == 0) {
    atomicAdd(&cold()->work[1], (unsigned long long)c_acc);
    atomicAdd(&cold()->work[2], (unsigned long long)c_rej);
    atomicAdd(&cold()->work[3], (unsigned long long)c_stp);
    atomicAdd(&cold()->work[4], (unsigned long long)c_fcn);
  }
}

template <int M, int L, bool SP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RAPT_GCN_WAVES)))
void k_trace_gcn(KArgs A) {
  gcn_body<M, L, SP>(A, 0);
}

// Mixed layout: blocks [0, oct_blocks) trace with lane octets and take
// queue 0 first, blocks [oct_blocks, wide_blocks) with lane quads and take
// queue 1 first (the most expensive particles in cost order: their step
// chains are the launch's critical path, and a wide group's step is the
// shortest), the other blocks with lane pairs (fewer instructions per
// particle-step) and take queue 2 first.  Results do not depend on which
// layout traced a particle.
template <int M, bool SP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RAPT_GCN_WAVES)))
void k_trace_gcmix(KArgs A, int32_t oct_blocks, int32_t wide_blocks) {
  if ((int32_t)blockIdx.x < oct_blocks) {
    gcn_body<M, 8, SP>(A, 0);
  } else if ((int32_t)blockIdx.x < wide_blocks) {
    gcn_body<M, 4, SP>(A, 1);
  } else {
    gcn_body<M, 2, SP>(A, 2);
  }
}

}  // namespace rapt
