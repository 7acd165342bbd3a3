// rapt_gc4.hip -- GuidingCenter.advance with four lanes per guiding centre.
//
// A quad of lanes (DPP quad q = lane & 3) carries one particle: lane q holds
// component q of the state (x, y, z, ppar) and of every DOPRI5 stage, so the
// stage combinations are split four ways.  For a right-hand side the quad
// gathers the stage vector with quad-permute DPP moves, then each lane
// evaluates the field at two points of the finite-difference stencil (lane 0
// the centre, lanes 1-3 the +-x, +-y, +-z pairs), forms its share of
// grad|B| and curl b, and the shares are broadcast back.  Every lane then runs
// the same guiding-centre algebra and keeps its own output component.  All
// reductions (error norm, HINIT norms) are summed in one fixed order from
// broadcast values, so the four lanes stay bit-identical and take the same
// branches.  Driver semantics are those of k_trace<M, KIND_GC> (rapt_tracer.hip):
// rapt/GuidingCenter.py:442-458 with scipy's dopri5 call semantics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_internal.hpp"
#include "rapt_physics.hpp"
#include "rapt_steppers.hpp"

namespace rapt {

// value of lane I of this lane's quad (DPP quad_perm(I, I, I, I))
template <int I>
__device__ __forceinline__ double qb(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, I * 0x55, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), I * 0x55, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int I>
__device__ __forceinline__ int qbi(int v) {
  return __builtin_amdgcn_mov_dpp(v, I * 0x55, 0xF, 0xF, false);
}
__device__ __forceinline__ long long qbll0(long long v) {
  const int lo = qbi<0>((int)v), hi = qbi<0>((int)(v >> 32));
  return ((long long)hi << 32) | (long long)(unsigned)lo;
}
// sum of squares of the quad's four values, in component order
__device__ __forceinline__ double quad_sumsq(double v) {
  const double a = qb<0>(v), b = qb<1>(v), c = qb<2>(v), d = qb<3>(v);
  return __builtin_fma(d, d, __builtin_fma(c, c, __builtin_fma(b, b, a * a)));
}
__device__ __forceinline__ double sel4(int q, double a, double b, double c, double d) {
  return q == 0 ? a : (q == 1 ? b : (q == 2 ? c : d));
}

// guiding-centre right-hand side, component q of the output
template <int M>
__device__ __forceinline__ double rhs_gc_quad(const FieldC& F, int eom, double charge,
                                              double mass, double mu, double v, bool eq, int q,
                                              double t, double yq, int64_t& nf) {
  ++nf;
  const double Y0 = qb<0>(yq), Y1 = qb<1>(yq), Y2 = qb<2>(yq), Y3 = qb<3>(yq);
  const double d = F.gstep;
  const double i2d = 1.0 / (2.0 * d);
  // lane 0: the centre (and, for a time-dependent field, the centre at
  // t - tstep); lane k: the +-d points on axis k, formed as the one-lane
  // stencil forms them (coordinate + offset, +0.0 off the axis)
  const double ax = (q == 1) ? d : 0.0, ay = (q == 2) ? d : 0.0, az = (q == 3) ? d : 0.0;
  const double bx = (q == 1) ? -d : 0.0, by = (q == 2) ? -d : 0.0, bz = (q == 3) ? -d : 0.0;
  const bool ctr = q == 0;
  const V3 Ba = fieldB<M>(F, t, ctr ? Y0 : Y0 + ax, ctr ? Y1 : Y1 + ay, ctr ? Y2 : Y2 + az);
  double tb = t;
  if (!F.is_static) tb = ctr ? t - F.tstep : t;
  const V3 Bb = fieldB<M>(F, tb, ctr ? Y0 : Y0 + bx, ctr ? Y1 : Y1 + by, ctr ? Y2 : Y2 + bz);
  const double ba = dot3(Ba, Ba), bbb = dot3(Bb, Bb);
  const double ia = rsqrt_nr(ba), ib = rsqrt_nr(bbb);
  const double ma = ba * ia, mb = bbb * ib;
  const V3 ua{Ba.x * ia, Ba.y * ia, Ba.z * ia}, ub_{Bb.x * ib, Bb.y * ib, Bb.z * ib};
  const V3 du{ua.x - ub_.x, ua.y - ub_.y, ua.z - ub_.z};
  const double g = (ma - mb) * i2d;
  const double e0 = (q == 0) ? Ba.x : g;
  const double e1 = (q == 0) ? Ba.y : ((q == 1) ? du.y : du.x);
  const double e2 = (q == 0) ? Ba.z : ((q == 3) ? du.y : du.z);
  const V3 B{qb<0>(e0), qb<0>(e1), qb<0>(e2)};
  const V3 gB{qb<1>(e0), qb<2>(e0), qb<3>(e0)};
  const V3 cb = curl_from_pairs(qb<1>(e1), qb<1>(e2), qb<2>(e1), qb<2>(e2), qb<3>(e1),
                                qb<3>(e2), i2d);
  V3 db{0.0, 0.0, 0.0};
  if (eom == EOM_TCB && !F.is_static) {
    // lane 0's second point was B(t - tstep); B(t + tstep) is a third evaluation
    const V3 Bc = fieldB<M>(F, t + F.tstep, Y0, Y1, Y2);
    const double ic = rsqrt_nr(dot3(Bc, Bc));
    const double s = 0.5 / F.tstep;
    const double l0 = (Bc.x * ic - ub_.x) * s, l1 = (Bc.y * ic - ub_.y) * s,
                 l2 = (Bc.z * ic - ub_.z) * s;
    db = V3{qb<0>(l0), qb<0>(l1), qb<0>(l2)};
  }
  const double Y[4] = {Y0, Y1, Y2, Y3};
  double o[4];
  gc_algebra(eom, has_E<M>(), fieldE<M>(F), F.is_static != 0, charge, mass, mu, v, eq, Y, B, gB,
             cb, db, o);
  return sel4(q, o[0], o[1], o[2], o[3]);
}

// resident waves per SIMD the register allocation targets
#ifndef RAPT_GC4_WAVES
#define RAPT_GC4_WAVES 2
#endif

template <int M>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RAPT_GC4_WAVES)))
void k_trace_gc4(KArgs A) {
  const FieldC& F = A.F;
  const int64_t n = A.n;
  double* S = A.state;
  const int q = threadIdx.x & 3;
  bool active = false, exhausted = false;
  int64_t pi = 0;
  double x = 0, y = 0, k1 = 0;
  double h = 0, dt = 0, tend = 0, gm = 0, tcur = 0, mass = 0, charge = 0, mu = 0, v = 0;
  double xend = 0, hmax = 0, facold = 1e-4;
  int nstep = 0, nacc_call = 0;
  bool last = false, reject = false, k1ok = false;
  int64_t nrow = 0, rowcap = 0, rowbase = 0;
  int64_t c_acc = 0, c_rej = 0, c_stp = 0, c_fcn = 0, p_acc = 0;
  const int W = 5;
  const CtlParams& C = A.c5;
  const bool eq = A.eq != 0;

  auto write_row = [&](double trow) {
    double* r = A.rows + (rowbase + nrow) * W;
    if (q == 0) r[0] = trow;
    r[1 + q] = y;
    ++nrow;
  };
  auto begin_call = [&]() {
    xend = x + dt;
    hmax = (A.max_step == 0.0) ? xend - x : A.max_step;
    facold = 1e-4;
    nstep = 0;
    nacc_call = 0;
    last = false;
    reject = false;
  };
  auto finish = [&](int32_t code) {
    S[(1 + q) * n + pi] = y;
    if (q == 0) {
      S[0 * n + pi] = x;
      S[RAPT_S_MU * n + pi] = mu;
      S[RAPT_S_V * n + pi] = v;
      S[RAPT_S_DT * n + pi] = dt;
      S[RAPT_S_H * n + pi] = h;
      S[RAPT_S_TEND * n + pi] = tend;
      S[RAPT_S_GM * n + pi] = gm;
      S[RAPT_S_TCUR * n + pi] = tcur;
      A.flags[pi] = FLAG_STARTED | FLAG_GCMODE | (code != RAPT_ST_CAP ? FLAG_FINAL : 0);
      A.status[pi] = code;
      A.nrows[pi] = nrow;
      A.naccpt[pi] = p_acc;
    }
    active = false;
  };

  for (;;) {
    if (!active && !exhausted) {
      int64_t got = 0;
      if (q == 0) {
        got = (int64_t)atomicAdd(&A.work[0], 1ULL);
        if (got < n && A.order) got = A.order[got];
      }
      pi = qbll0(got);
      if (pi >= n) {
        exhausted = true;
      } else if (A.flags[pi] & FLAG_FINAL) {
        // finished in an earlier launch: no rows this time, status kept
        if (q == 0) {
          A.nrows[pi] = 0;
          A.naccpt[pi] = 0;
        }
      } else {
        active = true;
        x = S[0 * n + pi];
        y = S[(1 + q) * n + pi];
        mass = S[RAPT_S_MASS * n + pi];
        charge = S[RAPT_S_CHARGE * n + pi];
        mu = S[RAPT_S_MU * n + pi];
        v = S[RAPT_S_V * n + pi];
        dt = S[RAPT_S_DT * n + pi];
        h = S[RAPT_S_H * n + pi];
        tend = S[RAPT_S_TEND * n + pi];
        gm = S[RAPT_S_GM * n + pi];
        tcur = S[RAPT_S_TCUR * n + pi];
        const int32_t fl = A.flags[pi];
        nrow = 0;
        rowcap = A.row_cap[pi];
        rowbase = A.row_off[pi];
        p_acc = 0;
        k1ok = false;
        if (!(fl & FLAG_STARTED)) {
          tend = x + A.delta;
          h = A.first_step;
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
      k1 = rhs_gc_quad<M>(F, A.eom, charge, mass, mu, v, eq, q, x, y, c_fcn);
      k1ok = true;
    }
    if (active && h == 0.0) {
      // HINIT with quad-summed norms (Hairer's dopri5, iord 5)
      const double sk = A.atol + A.rtol * __builtin_fabs(y);
      const double dnf = quad_sumsq(k1 / sk), dny = quad_sumsq(y / sk);
      double hh = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : __builtin_sqrt(dny / dnf) * 0.01;
      hh = fmin(hh, hmax);
      const double f1 = rhs_gc_quad<M>(F, A.eom, charge, mass, mu, v, eq, q, x + hh,
                                       y + hh * k1, c_fcn);
      double der2 = quad_sumsq((f1 - k1) / sk);
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
    double y1 = 0.0, k7 = 0.0, err = 0.0;
    if (stepping) {
      using namespace dp5;
      double k2 = 0, k3 = 0, k4 = 0, k5 = 0, k6 = 0;
#pragma unroll 1
      for (int s = 2; s <= 7; ++s) {
        double ts, xs;
        switch (s) {
          case 2: ts = y + h * a21 * k1; xs = x + c2 * h; break;
          case 3: ts = y + h * (a31 * k1 + a32 * k2); xs = x + c3 * h; break;
          case 4: ts = y + h * (a41 * k1 + a42 * k2 + a43 * k3); xs = x + c4 * h; break;
          case 5: ts = y + h * (a51 * k1 + a52 * k2 + a53 * k3 + a54 * k4); xs = x + c5 * h; break;
          case 6:
            ts = y + h * (a61 * k1 + a62 * k2 + a63 * k3 + a64 * k4 + a65 * k5);
            xs = x + h;
            break;
          default:
            y1 = y + h * (a71 * k1 + a73 * k3 + a74 * k4 + a75 * k5 + a76 * k6);
            ts = y1;
            xs = x + h;
            break;
        }
        const double kk = rhs_gc_quad<M>(F, A.eom, charge, mass, mu, v, eq, q, xs, ts, c_fcn);
        if (s == 2) k2 = kk;
        else if (s == 3) k3 = kk;
        else if (s == 4) k4 = kk;
        else if (s == 5) k5 = kk;
        else if (s == 6) k6 = kk;
        else k7 = kk;
      }
      const double e = (e1 * k1 + e3 * k3 + e4 * k4 + e5 * k5 + e6 * k6 + e7 * k7) * h;
      const double sk = A.atol + A.rtol * fmax(__builtin_fabs(y), __builtin_fabs(y1));
      err = __builtin_sqrt(quad_sumsq(e / sk) * (1.0 / 4));
    }
    if (stepping) {
      bool call_done = false;
      if (q == 0) ++c_stp;
      const double fac11 = pow(err, C.expo1);
      double fac = fac11 / pow(facold, C.beta);
      fac = fmax(C.facc2, fmin(C.facc1, fac / C.safe));
      double hnew = h / fac;
      if (err <= 1.0) {
        facold = fmax(err, 1.0e-4);
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
        const int fin = __builtin_isfinite(y) ? 0 : 1;
        const bool finite = (qbi<0>(fin) | qbi<1>(fin) | qbi<2>(fin) | qbi<3>(fin)) == 0;
        if (!finite) {
          finish(RAPT_ST_NAN);
        } else {
          bool sw = false;
          if (A.check) {
            const double yy[4] = {qb<0>(y), qb<1>(y), qb<2>(y), qb<3>(y)};
            sw = !gc_isadiabatic<M>(F, A, x, yy, mass, charge, mu);
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
    atomicAdd(&A.work[1], (unsigned long long)c_acc);
    atomicAdd(&A.work[2], (unsigned long long)c_rej);
    atomicAdd(&A.work[3], (unsigned long long)c_stp);
    atomicAdd(&A.work[4], (unsigned long long)c_fcn);
  }
}

}  // namespace rapt

using namespace rapt;

template <int M>
static void launch_gc4_m(hipStream_t st, const KArgs& A, int block, int64_t groups) {
  const int64_t lanes = groups * 4;
  hipLaunchKernelGGL(k_trace_gc4<M>, dim3((unsigned)((lanes + block - 1) / block)), dim3(block),
                     0, st, A);
}

void rapt_launch_gc4(hipStream_t st, int model, const KArgs& A, int block, int64_t groups) {
  switch (model) {
    case 0: launch_gc4_m<0>(st, A, block, groups); break;
    case 1: launch_gc4_m<1>(st, A, block, groups); break;
    case 2: launch_gc4_m<2>(st, A, block, groups); break;
    case 3: launch_gc4_m<3>(st, A, block, groups); break;
    case 4: launch_gc4_m<4>(st, A, block, groups); break;
    case 5: launch_gc4_m<5>(st, A, block, groups); break;
    case 6: launch_gc4_m<6>(st, A, block, groups); break;
    case 7: launch_gc4_m<7>(st, A, block, groups); break;
    case 8: launch_gc4_m<8>(st, A, block, groups); break;
  }
}
