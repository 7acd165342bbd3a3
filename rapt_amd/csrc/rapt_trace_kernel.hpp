// rapt_trace_kernel.hpp -- the persistent batched tracer kernel k_trace<M, KIND, P>
// (field model, tracer kind, field precision), shared by the translation
// units that instantiate it (rapt_tracer.hip: FP64, rapt_tracer_f32.hip:
// FP32 field).
//
// One lane carries one particle: its FP64 state, its integrator state (step
// size, PI-controller memory, call bookkeeping) and, in ADAPTIVE mode, its
// Particle/GuidingCenter segment mode.  The kernel loop is step-synchronous:
// every iteration performs one full Runge-Kutta step attempt for every active
// lane, so the expensive stage evaluations run converged; accept/reject,
// output rows and integrate()-call boundaries are per-lane bookkeeping.  A lane
// whose particle finishes takes the next one from a global counter (work
// refill), so a launch with fewer resident lanes than particles balances
// particles of very different step counts.
//
// Driver semantics follow the reference exactly:
//   Particle.advance  rapt/Particle.py:270-309  (dop853, beta 0.1, dt from the
//                     cyclotron period, row time r.t+dt, tcur one dt ahead)
//   GuidingCenter.advance rapt/GuidingCenter.py:442-458 (dopri5, row [r.t, y])
//   Adaptive.advance  rapt/Adaptive.py:187-222 with Particle.init /
//                     GuidingCenter.init (rapt/Particle.py:149-164,
//                     rapt/GuidingCenter.py:168-186)
// and scipy.integrate.ode's call semantics (one integrate() per output row,
// step size carried between calls, at most nsteps+1 steps per call).
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

// ---- right-hand-side functors --------------------------------------------------
// P = RAPT_PREC_FP64 or RAPT_PREC_FP32_FIELD (field and drift terms in float)
template <int M, int P>
struct LorentzRHS {
  const FieldC& F;
  double charge, mass, gm;
  bool eq;
  int64_t nf;
  __device__ __forceinline__ void operator()(double t, const double* y, double* o) {
    if constexpr (P == RAPT_PREC_FP32_FIELD) {
      rhs_lorentz_f32<M>(F, charge, mass, gm, eq, t, y, o);
    } else {
      rhs_lorentz<M>(F, charge, mass, gm, eq, t, y, o);
    }
    ++nf;
  }
};

// SP: the Tao-Chan-Brizard equations in a static field, compiled in
template <int M, int P, bool SP>
struct GcRHS {
  const FieldC& F;
  int eom;
  double charge, mass, mu, v;
  bool eq;
  int64_t nf;
  GcK K;  // gc_consts(eom, charge, mass, v, mu), formed once per step
  __device__ __forceinline__ void operator()(double t, const double* y, double* o) {
    const int e = SP ? (int)EOM_TCB : eom;
    const bool stat = SP || F.is_static != 0;
    if constexpr (P == RAPT_PREC_FP32_FIELD) {
      rhs_gc_f32<M>(F, e, stat, charge, mass, mu, v, K, eq, t, y, o);
    } else {
      rhs_gc<M>(F, e, stat, charge, mass, mu, v, K, eq, t, y, o);
    }
    ++nf;
  }
};

// ---- the persistent tracer kernel ---------------------------------------------------
// SP: for KIND_GC, the instantiation for the Tao-Chan-Brizard equations in a
// static field (those choices compiled in); the generic one reads A.eom and
// F.is_static.
// Register budget: two waves per SIMD (256 registers per lane) for the
// analytic fields, where it costs 12-92 bytes of spills per lane in the GC
// and Particle kernels and hides the dependent-chain latency better (one-lane
// GC +21% FP64, +31% FP32-field on 1,048,576 particles); one wave for
// gridded fields, where two spill 180 bytes per lane in the Particle kernel
// and ran 15% slower (65,536 protons in a 257 x 257 x 193 dipole grid), and
// for interpreted user-written fields, whose interpreter calls carry a
// register file (compiled ones are inline code like the analytic models).
template <int M>
constexpr int trace_waves() {
  return M == F_GRID || (M == F_EXPR && !EXPR_COMPILED) ? 1 : 2;
}
template <int M, int KIND, int P, bool SP = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(trace_waves<M>())))
void k_trace(KArgs A) {
  const int eom = SP ? (int)EOM_TCB : A.eom;
  const FieldC& F = A.F;
  const int64_t n = cold()->n;
  bool active = false, exhausted = false;
  int64_t pi = 0;
  // per-lane particle + integrator state
  double x = 0, y[6] = {0, 0, 0, 0, 0, 0}, k1[6] = {0, 0, 0, 0, 0, 0};
  double h = 0, dt = 0, tend = 0, gm = 0, tcur = 0, mass = 0, charge = 0, mu = 0, v = 0;
  double xend = 0, hmax = 0, lfo = LFO0;  // lfo = log2(FACOLD)
  int nstep = 0, nacc_call = 0, stepper = A.stepper;
  bool last = false, reject = false, k1ok = false, gcmode = (KIND == KIND_GC);
  int64_t nrow = 0, rowcap = 0, rowbase = 0;
  int64_t c_acc = 0, c_rej = 0, c_stp = 0, c_fcn = 0, c_sw = 0, p_acc = 0;
  int32_t st = RAPT_ST_PENDING;

  auto write_row = [&](double trow) {
    const int W = cold()->width;
    double* r = cold()->rows + (rowbase + nrow) * W;
    if constexpr (KIND == KIND_ADAPTIVE) {
      r[0] = gcmode ? 1.0 : 0.0;
      r[1] = trow;
      if (gcmode) {
        r[2] = y[0]; r[3] = y[1]; r[4] = y[2]; r[5] = y[3]; r[6] = mu; r[7] = 0.0;
      } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) r[2 + i] = y[i];
      }
    } else {
      r[0] = trow;
      if constexpr (KIND == KIND_GC) {
#pragma unroll
        for (int i = 0; i < 4; ++i) r[1 + i] = y[i];
      } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) r[1 + i] = y[i];
      }
    }
    ++nrow;
  };
  // start an integrate() call towards x + dt (rapt/Particle.py:305, rapt/GuidingCenter.py:453)
  auto begin_call = [&]() {
    xend = x + dt;
    hmax = (cold()->max_step == 0.0) ? xend - x : cold()->max_step;
    lfo = LFO0;
    nstep = 0;
    nacc_call = 0;
    last = false;
    reject = false;
    if (stepper == RAPT_STEP_RKF45) h = A.rkf_hmax;
  };
  // start a new advance() of the current segment at its current state
  auto begin_advance = [&](double tdelta_end) {
    tend = tdelta_end;
    h = cold()->first_step;
    tcur = x;
    if (!gcmode) {
      const V3 p{y[3], y[4], y[5]};
      gm = __builtin_sqrt(mass * mass + dot3(p, p) / C_LIGHT_SQ);
      const V3 vel{p.x / gm, p.y / gm, p.z / gm};
      const double gamma = 1.0 / __builtin_sqrt(1.0 - dot3(vel, vel) / C_LIGHT_SQ);
      const double B = magB<M>(F, x, y[0], y[1], y[2]);
      dt = TWO_PI * gamma * mass / B / __builtin_fabs(charge) / cold()->cycres;
      stepper = (KIND == KIND_ADAPTIVE) ? RAPT_STEP_DOP853 : A.stepper;
    } else if constexpr (KIND == KIND_ADAPTIVE) {
      dt = cold()->gcdt;
      stepper = RAPT_STEP_DOPRI5;
    }
    k1ok = false;
  };
  auto finish = [&](int32_t code) {
    st = code;
    double* S = cold()->state;
    S[0 * n + pi] = x;
#pragma unroll
    for (int i = 0; i < 6; ++i) S[(1 + i) * n + pi] = y[i];
    S[RAPT_S_MU * n + pi] = mu;
    S[RAPT_S_V * n + pi] = v;
    S[RAPT_S_DT * n + pi] = dt;
    S[RAPT_S_H * n + pi] = h;
    S[RAPT_S_TEND * n + pi] = tend;
    S[RAPT_S_GM * n + pi] = gm;
    S[RAPT_S_TCUR * n + pi] = tcur;
    int32_t fl = FLAG_STARTED | (gcmode ? FLAG_GCMODE : 0) | (st != RAPT_ST_CAP ? FLAG_FINAL : 0);
    cold()->flags[pi] = fl;
    cold()->status[pi] = st;
    cold()->nrows[pi] = nrow;
    cold()->naccpt[pi] = p_acc;
    active = false;
  };

  for (;;) {
    // ---- refill: take the next particle ---------------------------------------
    if (!active && !exhausted) {
      pi = (int64_t)atomicAdd(&cold()->work[0], 1ULL);
      if (pi < n && cold()->order) pi = cold()->order[pi];
      if (pi >= n) {
        exhausted = true;
      } else if (cold()->flags[pi] & FLAG_FINAL) {
        // finished in an earlier launch: no rows this time, status kept
        cold()->nrows[pi] = 0;
        cold()->naccpt[pi] = 0;
      } else {
        active = true;
        const double* S = cold()->state;
        x = S[0 * n + pi];
#pragma unroll
        for (int i = 0; i < 6; ++i) y[i] = S[(1 + i) * n + pi];
        mass = S[RAPT_S_MASS * n + pi];
        charge = S[RAPT_S_CHARGE * n + pi];
        mu = S[RAPT_S_MU * n + pi];
        v = S[RAPT_S_V * n + pi];
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
        if constexpr (KIND == KIND_ADAPTIVE) {
          gcmode = (fl & FLAG_GCMODE) != 0;
          stepper = gcmode ? RAPT_STEP_DOPRI5 : RAPT_STEP_DOP853;
          if (fl & FLAG_AINIT) {
            // Adaptive.__init__ (rapt/Adaptive.py:96-104): particle state in
            gcmode = false;
            if (particle_isadiabatic<M>(F, cold()->epss, cold()->epst, x, y, mass, charge)) {
              if (particle_to_gc<M>(F, x, y, mass, charge, mu, v)) {
                gcmode = true;
              } else {
                finish(RAPT_ST_GCITER);
                continue;
              }
            }
            if (rowcap < 1) {
              finish(RAPT_ST_CAP);
              continue;
            }
            write_row(x);
            begin_advance(x + cold()->delta);
          } else if (!(fl & FLAG_STARTED)) {
            begin_advance(x + cold()->delta);
          }
        } else {
          if (!(fl & FLAG_STARTED)) begin_advance(x + cold()->delta);
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

    // ---- f0 = f(x, y) at a call start without a carried FSAL stage ------------
    if (active && !k1ok) {
      if (!gcmode) {
        LorentzRHS<M, P> f{F, charge, mass, gm, A.eq != 0, 0};
        if (stepper != RAPT_STEP_RKF45) f(x, y, k1);
        gm = f.gm;
        c_fcn += f.nf;
      } else {
        GcRHS<M, P, SP> f{F, eom, charge, mass, mu, v, A.eq != 0, 0, gc_consts(eom, charge, mass, v, mu)};
        if (stepper != RAPT_STEP_RKF45) f(x, y, k1);
        c_fcn += f.nf;
      }
      k1ok = true;
    }
    // ---- HINIT on the first call of an advance (first_step == 0) -------------
    if (active && h == 0.0) {
      if (!gcmode) {
        LorentzRHS<M, P> f{F, charge, mass, gm, A.eq != 0, 0};
        h = hinit<6, P == RAPT_PREC_FP32_FIELD>(f, x, y, k1, stepper == RAPT_STEP_DOP853 ? 8 : 5, hmax, A.atol, A.rtol);
        gm = f.gm;
        c_fcn += f.nf;
      } else {
        GcRHS<M, P, SP> f{F, eom, charge, mass, mu, v, A.eq != 0, 0, gc_consts(eom, charge, mass, v, mu)};
        const double at = (KIND == KIND_ADAPTIVE) ? A.gc_atol : A.atol;
        const double rt = (KIND == KIND_ADAPTIVE) ? A.gc_rtol : A.rtol;
        h = hinit<4, P == RAPT_PREC_FP32_FIELD>(f, x, y, k1, stepper == RAPT_STEP_DOP853 ? 8 : 5, hmax, at, rt);
        c_fcn += f.nf;
      }
    }
    // ---- step start checks (Hairer's core loop head) --------------------------
    bool stepping = false;
    if (active) {
      if (stepper == RAPT_STEP_RKF45) {
        if (x + h > xend) h = xend - x;
        stepping = true;
        ++nstep;
      } else if (nstep > A.nmax || 0.1 * __builtin_fabs(h) <= __builtin_fabs(x) * UROUND) {
        const int32_t code = (nstep > A.nmax) ? RAPT_ST_NSTEPS : RAPT_ST_HSMALL;
        write_row((KIND != KIND_GC && !gcmode) ? xend : x);
        if (!gcmode) tcur = x + dt; else tcur = x;
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
    // ---- one full step attempt --------------------------------------------------
    double y1[6], k7[6], err = 0.0, fob = 1.0;
    if (stepping) {
      if (stepper != RAPT_STEP_RKF45)
        fob = ctl_fob(((stepper == RAPT_STEP_DOP853) ? A.c853 : A.c5).beta, lfo);
      if (!gcmode) {
        LorentzRHS<M, P> f{F, charge, mass, gm, A.eq != 0, 0};
        if (KIND == KIND_ADAPTIVE || stepper == RAPT_STEP_DOP853) {
          err = dop853_attempt<6, P == RAPT_PREC_FP32_FIELD>(f, x, h, y, k1, y1, A.atol, A.rtol);
        } else {
          err = rkf45_attempt<6>(f, x, h, y, y1);
        }
        gm = f.gm;
        c_fcn += f.nf;
      } else {
        GcRHS<M, P, SP> f{F, eom, charge, mass, mu, v, A.eq != 0, 0, gc_consts(eom, charge, mass, v, mu)};
        const double at = (KIND == KIND_ADAPTIVE) ? A.gc_atol : A.atol;
        const double rt = (KIND == KIND_ADAPTIVE) ? A.gc_rtol : A.rtol;
        if (KIND == KIND_ADAPTIVE || stepper == RAPT_STEP_DOPRI5) {
          err = dopri5_attempt<4, P == RAPT_PREC_FP32_FIELD>(f, x, h, y, k1, y1, k7, at, rt);
        } else {
          err = rkf45_attempt<4>(f, x, h, y, y1);
        }
        c_fcn += f.nf;
      }
    }
    // DOP853: the 12th evaluation f(x+h, y1) only after acceptance
    if (KIND != KIND_GC && stepping && !gcmode && stepper == RAPT_STEP_DOP853 && err <= 1.0) {
      LorentzRHS<M, P> f{F, charge, mass, gm, A.eq != 0, 0};
      f(x + h, y1, k7);
      gm = f.gm;
      c_fcn += f.nf;
    }
    // ---- step-size control and call bookkeeping ---------------------------------
    if (stepping) {
      const int N = gcmode ? 4 : 6;
      bool call_done = false;
      int32_t fail = 0;
      ++c_stp;
      if (stepper == RAPT_STEP_RKF45) {
        // rapt/rkf.py:122-139
        if (err <= A.rkf_tol) {
          x = x + h;
          for (int i = 0; i < 6; ++i)
            if (i < N) y[i] = y1[i];
          ++c_acc;
          ++p_acc;
        } else {
          ++c_rej;
        }
        h = h * fmin(fmax(0.84 * pow(A.rkf_tol / err, 0.25), 0.1), 4.0);
        if (h > A.rkf_hmax) {
          h = A.rkf_hmax;
        } else if (h < A.rkf_hmin || !(h > 0.0)) {
          // with hmin = 0, a non-finite error shrinks h to zero, where rkf
          // would repeat a zero step for ever
          fail = RAPT_ST_HMIN;
        }
        if (!fail && !(x < xend)) call_done = true;
      } else {
        const CtlParams& C = (stepper == RAPT_STEP_DOP853) ? A.c853 : A.c5;
        const double l2e = log2(err);
        const double fac11 = exp2(C.expo1 * l2e);
        double fac = fac11 / fob;
        fac = fmax(C.facc2, fmin(C.facc1, fac / C.safe));
        double hnew = h / fac;
        if (err <= 1.0) {
          lfo = fmax(l2e, LFO0);
          ++nacc_call;
          ++c_acc;
          ++p_acc;
          for (int i = 0; i < 6; ++i)
            if (i < N) {
              k1[i] = k7[i];
              y[i] = y1[i];
            }
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
          if (nacc_call >= 1) ++c_rej;
          last = false;
          h = hnew;
        }
      }
      if (fail) {
        write_row(gcmode ? x : xend);
        tcur = gcmode ? x : x + dt;
        finish(fail);
      } else if (call_done) {
        // the reference appends one trajectory row per integrate() call
        write_row(gcmode ? x : xend);
        tcur = gcmode ? x : x + dt;
        bool finite = true;
        for (int i = 0; i < 6; ++i)
          if (i < N) finite = finite && __builtin_isfinite(y[i]);
        bool sw = false;
        if (!finite) {
          finish(RAPT_ST_NAN);
        } else {
          if (KIND == KIND_ADAPTIVE || cold()->check) {
            if (!gcmode) {
              sw = particle_isadiabatic<M>(F, cold()->epss, cold()->epst, xend, y, mass, charge);
            } else {
              sw = !gc_isadiabatic<M>(F, cold()->epss, cold()->epst, x, y, mass, charge, mu);
            }
          }
          if (sw && KIND != KIND_ADAPTIVE) {
            finish(RAPT_ST_SWITCH);
          } else if (sw) {
            // Adaptive.advance (rapt/Adaptive.py:205-222): convert and restart
            ++c_sw;
            const double trow = gcmode ? x : xend;
            bool ok = true;
            if (!gcmode) {
              ok = particle_to_gc<M>(F, trow, y, mass, charge, mu, v);
              if (ok) gcmode = true;
            } else {
              gc_to_particle<M>(F, trow, y, mass, charge, mu);
              gcmode = false;
            }
            x = trow;
            if (!ok) {
              finish(RAPT_ST_GCITER);
            } else if (nrow >= rowcap) {
              finish(RAPT_ST_CAP);
            } else {
              write_row(trow);
              // t = current.tcur = segment start; advance(delta - t)
              begin_advance(trow + (cold()->delta - trow));
              if (!(x < tend)) {
                finish(RAPT_ST_DONE);
              } else if (nrow >= rowcap) {
                finish(RAPT_ST_CAP);
              } else {
                begin_call();
              }
            }
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
  // lane counters -> global (one atomic per lane and counter)
  atomicAdd(&cold()->work[1], (unsigned long long)c_acc);
  atomicAdd(&cold()->work[2], (unsigned long long)c_rej);
  atomicAdd(&cold()->work[3], (unsigned long long)c_stp);
  atomicAdd(&cold()->work[4], (unsigned long long)c_fcn);
  atomicAdd(&cold()->work[5], (unsigned long long)c_sw);
}

// ---- batched field evaluation (the _Field methods) -------------------------------
template <int M>
__global__ __launch_bounds__(256) void k_field_eval(FieldC F, int what, int64_t n,
                                                    const double* __restrict__ tpos,
                                                    double* __restrict__ out, int w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = tpos[4 * i], x = tpos[4 * i + 1], y = tpos[4 * i + 2], z = tpos[4 * i + 3];
  double* o = out + (int64_t)w * i;
  switch (what) {
    case 0: {
      V3 B = fieldB<M>(F, t, x, y, z);
      o[0] = B.x; o[1] = B.y; o[2] = B.z;
      break;
    }
    case 1: {
      V3 E = fieldE<M>(F, t, x, y, z);
      o[0] = E.x; o[1] = E.y; o[2] = E.z;
      break;
    }
    case 2: {
      V3 b = unitb<M>(F, t, x, y, z);
      o[0] = b.x; o[1] = b.y; o[2] = b.z;
      break;
    }
    case 3: o[0] = magB<M>(F, t, x, y, z); break;
    case 4:
    case 7: {
      V3 g, c;
      grad_curl<M>(F, t, x, y, z, g, c);
      const V3& r = (what == 4) ? g : c;
      o[0] = r.x; o[1] = r.y; o[2] = r.z;
      break;
    }
    case 5: {
      const double d = F.gstep, i2d = F.i2d;
      for (int k = 0; k < 3; ++k) {
        V3 Bp = fieldB<M>(F, t, x + (k == 0 ? d : 0.0), y + (k == 1 ? d : 0.0), z + (k == 2 ? d : 0.0));
        V3 Bm = fieldB<M>(F, t, x - (k == 0 ? d : 0.0), y - (k == 1 ? d : 0.0), z - (k == 2 ? d : 0.0));
        o[0 * 3 + k] = (Bp.x - Bm.x) * i2d;
        o[1 * 3 + k] = (Bp.y - Bm.y) * i2d;
        o[2 * 3 + k] = (Bp.z - Bm.z) * i2d;
      }
      break;
    }
    case 6: o[0] = curvature<M>(F, t, x, y, z); break;
    case 8: {
      if (F.is_static) {
        o[0] = 0.0;
      } else {
        const double d = F.tstep;
        o[0] = (magB<M>(F, t + d, x, y, z) - magB<M>(F, t - d, x, y, z)) / d / 2.0;
      }
      break;
    }
    case 9: {
      V3 r{0.0, 0.0, 0.0};
      if (!F.is_static) r = dbdt<M>(F, t, x, y, z);
      o[0] = r.x; o[1] = r.y; o[2] = r.z;
      break;
    }
    case 10: o[0] = lengthscale<M>(F, t, x, y, z); break;
    case 11: o[0] = F.is_static ? __builtin_nan("") : timescale<M>(F, t, x, y, z); break;
    default: break;
  }
}

// ---- batched right-hand sides (the tracers' equations of motion) ---------------
// mode KIND_PARTICLE: ty [n][7] = (t, x, y, z, px, py, pz), k [n][3] = (mass,
// charge, gamma m frozen for a static field), out [n][6]; KIND_GC: ty [n][5]
// = (t, x, y, z, ppar), k [n][4] = (mass, charge, mu, v), out [n][4].  The
// functions the tracer kernels integrate, one evaluation per row.
template <int M>
__global__ __launch_bounds__(256) void k_rhs_eval(FieldC F, int mode, int eom, int eq, int64_t n,
                                                  const double* __restrict__ ty,
                                                  const double* __restrict__ k,
                                                  double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mode == KIND_PARTICLE) {
    const double* r = ty + 7 * i;
    double gm = k[3 * i + 2];
    rhs_lorentz<M>(F, k[3 * i + 1], k[3 * i], gm, eq != 0, r[0], r + 1, out + 6 * i);
  } else {
    const double* r = ty + 5 * i;
    const double mass = k[4 * i], charge = k[4 * i + 1], mu = k[4 * i + 2], v = k[4 * i + 3];
    const GcK K = gc_consts(eom, charge, mass, v, mu);
    rhs_gc<M>(F, eom, F.is_static != 0, charge, mass, mu, v, K, eq != 0, r[0], r + 1,
              out + 4 * i);
  }
}

// the one-lane GC kernel's TCB + static-field instantiation applies
// (VarEarthDipole is never static)
template <int M>
static inline bool trace_use_sp(const KArgs& A) {
  return M != F_VAR_EARTH_DIPOLE && A.eom == EOM_TCB && A.F.is_static != 0;
}

}  // namespace rapt
