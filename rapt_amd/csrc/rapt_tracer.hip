// rapt_tracer.hip -- persistent batched tracer kernels and the C-ABI
// (include/rapt_hip.h) for gfx950.
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
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <math.h>
#include <new>
#include <string>
#include <vector>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_internal.hpp"
#include "rapt_physics.hpp"
#include "rapt_steppers.hpp"

namespace rapt {

// ---- right-hand-side functors --------------------------------------------------
template <int M>
struct LorentzRHS {
  const FieldC& F;
  double charge, mass, gm;
  bool eq;
  int64_t nf;
  __device__ __forceinline__ void operator()(double t, const double* y, double* o) {
    rhs_lorentz<M>(F, charge, mass, gm, eq, t, y, o);
    ++nf;
  }
};

template <int M>
struct GcRHS {
  const FieldC& F;
  int eom;
  double charge, mass, mu, v;
  bool eq;
  int64_t nf;
  __device__ __forceinline__ void operator()(double t, const double* y, double* o) {
    rhs_gc<M>(F, eom, charge, mass, mu, v, eq, t, y, o);
    ++nf;
  }
};

// ---- the persistent tracer kernel ---------------------------------------------------
template <int M, int KIND>
__global__ __launch_bounds__(256) void k_trace(KArgs A) {
  const FieldC& F = A.F;
  const int64_t n = A.n;
  double* S = A.state;
  bool active = false, exhausted = false;
  int64_t pi = 0;
  // per-lane particle + integrator state
  double x = 0, y[6] = {0, 0, 0, 0, 0, 0}, k1[6] = {0, 0, 0, 0, 0, 0};
  double h = 0, dt = 0, tend = 0, gm = 0, tcur = 0, mass = 0, charge = 0, mu = 0, v = 0;
  double xend = 0, hmax = 0, facold = 1e-4;
  int nstep = 0, nacc_call = 0, stepper = A.stepper;
  bool last = false, reject = false, k1ok = false, gcmode = (KIND == KIND_GC);
  int64_t nrow = 0, rowcap = 0, rowbase = 0;
  int64_t c_acc = 0, c_rej = 0, c_stp = 0, c_fcn = 0, c_sw = 0, p_acc = 0;
  int32_t st = RAPT_ST_PENDING;
  const int W = A.width;

  auto write_row = [&](double trow) {
    double* r = A.rows + (rowbase + nrow) * W;
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
    hmax = (A.max_step == 0.0) ? xend - x : A.max_step;
    facold = 1e-4;
    nstep = 0;
    nacc_call = 0;
    last = false;
    reject = false;
    if (stepper == RAPT_STEP_RKF45) h = A.rkf_hmax;
  };
  // start a new advance() of the current segment at its current state
  auto begin_advance = [&](double tdelta_end) {
    tend = tdelta_end;
    h = A.first_step;
    tcur = x;
    if (!gcmode) {
      const V3 p{y[3], y[4], y[5]};
      gm = __builtin_sqrt(mass * mass + dot3(p, p) / C_LIGHT_SQ);
      const V3 vel{p.x / gm, p.y / gm, p.z / gm};
      const double gamma = 1.0 / __builtin_sqrt(1.0 - dot3(vel, vel) / C_LIGHT_SQ);
      const double B = magB<M>(F, x, y[0], y[1], y[2]);
      dt = TWO_PI * gamma * mass / B / __builtin_fabs(charge) / A.cycres;
      stepper = (KIND == KIND_ADAPTIVE) ? RAPT_STEP_DOP853 : A.stepper;
    } else if constexpr (KIND == KIND_ADAPTIVE) {
      dt = A.gcdt;
      stepper = RAPT_STEP_DOPRI5;
    }
    k1ok = false;
  };
  auto finish = [&](int32_t code) {
    st = code;
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
    int32_t fl = FLAG_STARTED | (gcmode ? FLAG_GCMODE : 0);
    A.flags[pi] = fl;
    A.status[pi] = st;
    A.nrows[pi] = nrow;
    A.naccpt[pi] = p_acc;
    active = false;
  };

  for (;;) {
    // ---- refill: take the next particle ---------------------------------------
    if (!active && !exhausted) {
      pi = (int64_t)atomicAdd(&A.work[0], 1ULL);
      if (pi < n && A.order) pi = A.order[pi];
      if (pi >= n) {
        exhausted = true;
      } else {
        active = true;
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
        const int32_t fl = A.flags[pi];
        nrow = 0;
        rowcap = A.row_cap[pi];
        rowbase = A.row_off[pi];
        p_acc = 0;
        k1ok = false;
        if constexpr (KIND == KIND_ADAPTIVE) {
          gcmode = (fl & FLAG_GCMODE) != 0;
          stepper = gcmode ? RAPT_STEP_DOPRI5 : RAPT_STEP_DOP853;
          if (fl & FLAG_AINIT) {
            // Adaptive.__init__ (rapt/Adaptive.py:96-104): particle state in
            gcmode = false;
            if (particle_isadiabatic<M>(F, A, x, y, mass, charge)) {
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
            begin_advance(x + A.delta);
          } else if (!(fl & FLAG_STARTED)) {
            begin_advance(x + A.delta);
          }
        } else {
          if (!(fl & FLAG_STARTED)) begin_advance(x + A.delta);
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
    if (__ballot(active) == 0) break;

    // ---- f0 = f(x, y) at a call start without a carried FSAL stage ------------
    if (active && !k1ok) {
      if (!gcmode) {
        LorentzRHS<M> f{F, charge, mass, gm, A.eq != 0, 0};
        if (stepper != RAPT_STEP_RKF45) f(x, y, k1);
        gm = f.gm;
        c_fcn += f.nf;
      } else {
        GcRHS<M> f{F, A.eom, charge, mass, mu, v, A.eq != 0, 0};
        if (stepper != RAPT_STEP_RKF45) f(x, y, k1);
        c_fcn += f.nf;
      }
      k1ok = true;
    }
    // ---- HINIT on the first call of an advance (first_step == 0) -------------
    if (active && h == 0.0) {
      if (!gcmode) {
        LorentzRHS<M> f{F, charge, mass, gm, A.eq != 0, 0};
        h = hinit<6>(f, x, y, k1, stepper == RAPT_STEP_DOP853 ? 8 : 5, hmax, A.atol, A.rtol);
        gm = f.gm;
        c_fcn += f.nf;
      } else {
        GcRHS<M> f{F, A.eom, charge, mass, mu, v, A.eq != 0, 0};
        const double at = (KIND == KIND_ADAPTIVE) ? A.gc_atol : A.atol;
        const double rt = (KIND == KIND_ADAPTIVE) ? A.gc_rtol : A.rtol;
        h = hinit<4>(f, x, y, k1, stepper == RAPT_STEP_DOP853 ? 8 : 5, hmax, at, rt);
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
    double y1[6], k7[6], err = 0.0;
    if (stepping) {
      if (!gcmode) {
        LorentzRHS<M> f{F, charge, mass, gm, A.eq != 0, 0};
        if (KIND == KIND_ADAPTIVE || stepper == RAPT_STEP_DOP853) {
          err = dop853_attempt<6>(f, x, h, y, k1, y1, A.atol, A.rtol);
        } else {
          err = rkf45_attempt<6>(f, x, h, y, y1);
        }
        gm = f.gm;
        c_fcn += f.nf;
      } else {
        GcRHS<M> f{F, A.eom, charge, mass, mu, v, A.eq != 0, 0};
        const double at = (KIND == KIND_ADAPTIVE) ? A.gc_atol : A.atol;
        const double rt = (KIND == KIND_ADAPTIVE) ? A.gc_rtol : A.rtol;
        if (KIND == KIND_ADAPTIVE || stepper == RAPT_STEP_DOPRI5) {
          err = dopri5_attempt<4>(f, x, h, y, k1, y1, k7, at, rt);
        } else {
          err = rkf45_attempt<4>(f, x, h, y, y1);
        }
        c_fcn += f.nf;
      }
    }
    // DOP853: the 12th evaluation f(x+h, y1) only after acceptance
    if (KIND != KIND_GC && stepping && !gcmode && stepper == RAPT_STEP_DOP853 && err <= 1.0) {
      LorentzRHS<M> f{F, charge, mass, gm, A.eq != 0, 0};
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
        } else if (h < A.rkf_hmin) {
          fail = RAPT_ST_HMIN;
        }
        if (!fail && !(x < xend)) call_done = true;
      } else {
        const CtlParams& C = (stepper == RAPT_STEP_DOP853) ? A.c853 : A.c5;
        const double fac11 = pow(err, C.expo1);
        double fac = fac11 / pow(facold, C.beta);
        fac = fmax(C.facc2, fmin(C.facc1, fac / C.safe));
        double hnew = h / fac;
        if (err <= 1.0) {
          facold = fmax(err, 1.0e-4);
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
          if (KIND == KIND_ADAPTIVE || A.check) {
            if (!gcmode) {
              sw = particle_isadiabatic<M>(F, A, xend, y, mass, charge);
            } else {
              sw = !gc_isadiabatic<M>(F, A, x, y, mass, charge, mu);
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
              begin_advance(trow + (A.delta - trow));
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
  atomicAdd(&A.work[1], (unsigned long long)c_acc);
  atomicAdd(&A.work[2], (unsigned long long)c_rej);
  atomicAdd(&A.work[3], (unsigned long long)c_stp);
  atomicAdd(&A.work[4], (unsigned long long)c_fcn);
  atomicAdd(&A.work[5], (unsigned long long)c_sw);
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
      V3 E = fieldE<M>(F);
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
      const double d = F.gstep, i2d = 1.0 / (2.0 * d);
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

}  // namespace rapt

// ============================================================================
// host side
// ============================================================================
using namespace rapt;

struct rapt_batch {
  rapt_ctx* ctx = nullptr;
  int64_t n = 0;
  double* d_state = nullptr;
  double* d_state0 = nullptr;  // state as uploaded (rapt_batch_reset)
  int32_t* d_flags0 = nullptr;
  int32_t* d_flags = nullptr;
  int32_t* d_status = nullptr;
  int64_t* d_nrows = nullptr;
  int64_t* d_naccpt = nullptr;
  int64_t* d_rowoff = nullptr;
  int64_t* d_rowcap = nullptr;
  int64_t* d_order = nullptr;  // work order: particle indices by output step
  double* d_rows = nullptr;
  int32_t width = 0;
  int64_t total_rows = 0;
  unsigned long long* d_work = nullptr;
};


static int field_width(int what) {
  switch (what) {
    case 0: case 1: case 2: case 4: case 7: case 9: return 3;
    case 5: return 9;
    case 3: case 6: case 8: case 10: case 11: return 1;
    default: return -1;
  }
}

template <int M>
static void launch_field_eval(rapt_ctx* c, int what, int64_t n, const double* dt, double* dout,
                              int w) {
  const int B = 256;
  const int64_t g = (n + B - 1) / B;
  hipLaunchKernelGGL(k_field_eval<M>, dim3((unsigned)g), dim3(B), 0, c->stream, c->field, what,
                     n, dt, dout, w);
}

static CtlParams ctl(double safety, double dfactor, double ifactor, double beta, bool dop853) {
  CtlParams p;
  p.safe = safety == 0.0 ? 0.9 : safety;
  const double fac1 = dfactor == 0.0 ? (dop853 ? 0.333 : 0.2) : dfactor;
  const double fac2 = ifactor == 0.0 ? (dop853 ? 6.0 : 10.0) : ifactor;
  p.facc1 = 1.0 / fac1;
  p.facc2 = 1.0 / fac2;
  if (dop853) {
    p.beta = beta <= 0.0 ? 0.0 : beta;
    p.expo1 = 1.0 / 8.0 - p.beta * 0.2;
  } else {
    p.beta = beta == 0.0 ? 0.04 : (beta < 0.0 ? 0.0 : beta);
    p.expo1 = 0.2 - p.beta * 0.75;
  }
  return p;
}

void rapt_launch_gc4(hipStream_t st, int model, const KArgs& A, int block, int64_t groups);

template <int M, int K>
static void launch_trace(rapt_ctx* c, const KArgs& A, int block, int64_t lanes) {
  const int64_t g = (lanes + block - 1) / block;
  hipLaunchKernelGGL((k_trace<M, K>), dim3((unsigned)g), dim3(block), 0, c->stream, A);
}

template <int K>
static void launch_model(rapt_ctx* c, const KArgs& A, int block, int64_t lanes) {
  switch (c->field.model) {
    case 0: launch_trace<0, K>(c, A, block, lanes); break;
    case 1: launch_trace<1, K>(c, A, block, lanes); break;
    case 2: launch_trace<2, K>(c, A, block, lanes); break;
    case 3: launch_trace<3, K>(c, A, block, lanes); break;
    case 4: launch_trace<4, K>(c, A, block, lanes); break;
    case 5: launch_trace<5, K>(c, A, block, lanes); break;
    case 6: launch_trace<6, K>(c, A, block, lanes); break;
    case 7: launch_trace<7, K>(c, A, block, lanes); break;
    case 8: launch_trace<8, K>(c, A, block, lanes); break;
  }
}

extern "C" {

int rapt_ctx_create(int device, rapt_ctx** out) {
  if (!out) return RAPT_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RAPT_E_HIP;
  if (device < 0 || device >= ndev) return RAPT_E_ARG;
  rapt_ctx* c = new (std::nothrow) rapt_ctx();
  if (!c) return RAPT_E_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return RAPT_E_HIP;
  }
  *out = c;
  return RAPT_OK;
}

int rapt_ctx_destroy(rapt_ctx* c) {
  if (!c) return RAPT_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RAPT_OK;
}

const char* rapt_last_error(rapt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rapt_field_set(rapt_ctx* c, const rapt_field_desc* f) {
  if (!c || !f) return RAPT_E_ARG;
  if (f->model < 0 || f->model >= RAPT_FIELD_NMODELS)
    return set_err(c, RAPT_E_ARG, "unknown field model %d", f->model);
  if (!(f->gstep > 0.0)) return set_err(c, RAPT_E_ARG, "gradientstepsize must be > 0");
  if (!f->is_static && !(f->tstep > 0.0))
    return set_err(c, RAPT_E_ARG, "timederivstepsize must be > 0");
  static_assert(sizeof(FieldC) == sizeof(rapt_field_desc), "descriptor layout");
  memcpy(&c->field, f, sizeof(FieldC));
  c->field_set = true;
  return RAPT_OK;
}

int rapt_field_eval(rapt_ctx* c, int what, int64_t n, const double* tpos, double* out) {
  if (!c || (!tpos && n) || (!out && n) || n < 0) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  const int w = field_width(what);
  if (w < 0) return set_err(c, RAPT_E_ARG, "bad field quantity %d", what);
  if (n == 0) return RAPT_OK;
  HIP_OK(c, hipSetDevice(c->device));
  double *dt = nullptr, *dout = nullptr;
  HIP_OK(c, hipMalloc(&dt, sizeof(double) * 4 * n));
  if (hipMalloc(&dout, sizeof(double) * w * n) != hipSuccess) {
    (void)hipFree(dt);
    return set_err(c, RAPT_E_OOM, "hipMalloc failed");
  }
  int rc = RAPT_OK;
  if (hipMemcpyAsync(dt, tpos, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream) !=
      hipSuccess) {
    rc = set_err(c, RAPT_E_HIP, "H2D copy failed");
  } else {
    switch (c->field.model) {
      case 0: launch_field_eval<0>(c, what, n, dt, dout, w); break;
      case 1: launch_field_eval<1>(c, what, n, dt, dout, w); break;
      case 2: launch_field_eval<2>(c, what, n, dt, dout, w); break;
      case 3: launch_field_eval<3>(c, what, n, dt, dout, w); break;
      case 4: launch_field_eval<4>(c, what, n, dt, dout, w); break;
      case 5: launch_field_eval<5>(c, what, n, dt, dout, w); break;
      case 6: launch_field_eval<6>(c, what, n, dt, dout, w); break;
      case 7: launch_field_eval<7>(c, what, n, dt, dout, w); break;
      case 8: launch_field_eval<8>(c, what, n, dt, dout, w); break;
    }
    if (hipGetLastError() != hipSuccess) {
      rc = set_err(c, RAPT_E_HIP, "field kernel launch failed");
    } else if (hipMemcpyAsync(out, dout, sizeof(double) * w * n, hipMemcpyDeviceToHost,
                              c->stream) != hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess) {
      rc = set_err(c, RAPT_E_HIP, "field eval failed");
    }
  }
  (void)hipFree(dt);
  (void)hipFree(dout);
  return rc;
}

int rapt_batch_create(rapt_ctx* c, int64_t n, rapt_batch** out) {
  if (!c || !out || n <= 0) return c ? set_err(c, RAPT_E_ARG, "bad batch size") : RAPT_E_ARG;
  HIP_OK(c, hipSetDevice(c->device));
  rapt_batch* b = new (std::nothrow) rapt_batch();
  if (!b) return RAPT_E_OOM;
  b->ctx = c;
  b->n = n;
  bool ok = hipMalloc(&b->d_state, sizeof(double) * RAPT_NSLOT * n) == hipSuccess &&
            hipMalloc(&b->d_state0, sizeof(double) * RAPT_NSLOT * n) == hipSuccess &&
            hipMalloc(&b->d_flags0, sizeof(int32_t) * n) == hipSuccess &&
            hipMalloc(&b->d_flags, sizeof(int32_t) * n) == hipSuccess &&
            hipMalloc(&b->d_status, sizeof(int32_t) * n) == hipSuccess &&
            hipMalloc(&b->d_nrows, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_naccpt, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_rowoff, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_rowcap, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_order, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_work, sizeof(unsigned long long) * 8) == hipSuccess;
  if (!ok) {
    rapt_batch_destroy(b);
    return set_err(c, RAPT_E_OOM, "device allocation of a %lld-particle batch failed",
                   (long long)n);
  }
  *out = b;
  return RAPT_OK;
}

int rapt_batch_destroy(rapt_batch* b) {
  if (!b) return RAPT_E_ARG;
  if (b->ctx) (void)hipSetDevice(b->ctx->device);
  (void)hipFree(b->d_state);
  (void)hipFree(b->d_state0);
  (void)hipFree(b->d_flags0);
  (void)hipFree(b->d_flags);
  (void)hipFree(b->d_status);
  (void)hipFree(b->d_nrows);
  (void)hipFree(b->d_naccpt);
  (void)hipFree(b->d_rowoff);
  (void)hipFree(b->d_rowcap);
  (void)hipFree(b->d_order);
  (void)hipFree(b->d_rows);
  (void)hipFree(b->d_work);
  delete b;
  return RAPT_OK;
}

int rapt_batch_upload(rapt_batch* b, const double* state, const int32_t* flags) {
  if (!b || !state || !flags) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipMemcpyAsync(b->d_state, state, sizeof(double) * RAPT_NSLOT * b->n,
                           hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_flags, flags, sizeof(int32_t) * b->n, hipMemcpyHostToDevice,
                           c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_state0, b->d_state, sizeof(double) * RAPT_NSLOT * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_flags0, b->d_flags, sizeof(int32_t) * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  // Work order: shortest output step first.  For guiding centres the output
  // step is a fixed fraction of the bounce period and the step count scales
  // with delta / bounce period, so this starts the longest traces first and
  // puts traces of similar length into the same wavefront.  Entries without
  // an output step (0: computed in the kernel) keep their input order after
  // the others.
  std::vector<int64_t> order((size_t)b->n);
  const double* dts = state + (size_t)RAPT_S_DT * b->n;
  for (int64_t i = 0; i < b->n; ++i) order[(size_t)i] = i;
  std::stable_sort(order.begin(), order.end(), [dts](int64_t a, int64_t q) {
    const double ka = dts[a] > 0.0 ? dts[a] : HUGE_VAL, kq = dts[q] > 0.0 ? dts[q] : HUGE_VAL;
    return ka < kq;
  });
  HIP_OK(c, hipMemcpyAsync(b->d_order, order.data(), sizeof(int64_t) * b->n,
                           hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

int rapt_batch_reset(rapt_batch* b) {
  if (!b) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipMemcpyAsync(b->d_state, b->d_state0, sizeof(double) * RAPT_NSLOT * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_flags, b->d_flags0, sizeof(int32_t) * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  return RAPT_OK;
}

int rapt_batch_set_rows(rapt_batch* b, int32_t width, const int64_t* row_cap) {
  if (!b || !row_cap || width <= 0 || width > 16) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * b->n);
  if (!off) return set_err(c, RAPT_E_OOM, "host allocation failed");
  int64_t tot = 0;
  for (int64_t i = 0; i < b->n; ++i) {
    if (row_cap[i] < 0) {
      free(off);
      return set_err(c, RAPT_E_ARG, "negative row capacity");
    }
    off[i] = tot;
    tot += row_cap[i];
  }
  if (tot * width != b->total_rows * b->width || !b->d_rows) {
    (void)hipFree(b->d_rows);
    b->d_rows = nullptr;
    if (tot > 0 && hipMalloc(&b->d_rows, sizeof(double) * width * tot) != hipSuccess) {
      free(off);
      b->total_rows = 0;
      return set_err(c, RAPT_E_OOM, "trajectory buffer of %lld rows failed", (long long)tot);
    }
  }
  b->total_rows = tot;
  b->width = width;
  hipError_t e1 = hipMemcpyAsync(b->d_rowoff, off, sizeof(int64_t) * b->n,
                                 hipMemcpyHostToDevice, c->stream);
  hipError_t e2 = hipMemcpyAsync(b->d_rowcap, row_cap, sizeof(int64_t) * b->n,
                                 hipMemcpyHostToDevice, c->stream);
  hipError_t e3 = hipStreamSynchronize(c->stream);
  free(off);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return set_err(c, RAPT_E_HIP, "row table upload failed");
  return RAPT_OK;
}

int rapt_batch_run(rapt_batch* b, const rapt_trace_desc* d, rapt_counters* cnt) {
  if (!b || !d) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (d->mode < 0 || d->mode > 2) return set_err(c, RAPT_E_ARG, "bad mode %d", d->mode);
  const int width = d->mode == RAPT_MODE_PARTICLE ? 7 : (d->mode == RAPT_MODE_GC ? 5 : 8);
  if (b->width != width)
    return set_err(c, RAPT_E_STATE, "row width %d does not match mode %d", b->width, d->mode);
  if (d->eom < 0 || d->eom > 2) return set_err(c, RAPT_E_ARG, "bad eom %d", d->eom);
  if (d->stepper < 0 || d->stepper > 2) return set_err(c, RAPT_E_ARG, "bad stepper");
  if (d->mode == RAPT_MODE_PARTICLE && d->stepper == RAPT_STEP_DOPRI5)
    return set_err(c, RAPT_E_ARG, "PARTICLE mode steps with DOP853 or RKF45");
  if (d->mode == RAPT_MODE_GC && d->stepper == RAPT_STEP_DOP853)
    return set_err(c, RAPT_E_ARG, "GC mode steps with DOPRI5 or RKF45");
  if (d->lanes_per_particle != 0 && d->lanes_per_particle != 1 && d->lanes_per_particle != 4)
    return set_err(c, RAPT_E_ARG, "lanes_per_particle must be 0, 1 or 4");
  if (d->work_order != 0 && d->work_order != 1)
    return set_err(c, RAPT_E_ARG, "work_order must be 0 or 1");
  if (d->precision != RAPT_PREC_FP64)
    return set_err(c, RAPT_E_ARG, "precision %d is not built", d->precision);
  if (d->mode == RAPT_MODE_ADAPTIVE && !(d->gctimestep > 0.0))
    return set_err(c, RAPT_E_ARG, "ADAPTIVE needs params['GCtimestep'] > 0");
  if (d->mode == RAPT_MODE_PARTICLE && !(d->cyclotronresolution > 0.0))
    return set_err(c, RAPT_E_ARG, "cyclotronresolution must be > 0");
  if (d->stepper == RAPT_STEP_RKF45 && !(d->rkf_hmax > 0.0 && d->rkf_hmin > 0.0))
    return set_err(c, RAPT_E_ARG, "RKF45 needs rkf_hmax, rkf_hmin > 0");
  const int block = d->block > 0 ? d->block : 256;
  if (block % 64 != 0 || block > 256) return set_err(c, RAPT_E_ARG, "block must be 64..256, x64");
  HIP_OK(c, hipSetDevice(c->device));
  KArgs A;
  memset(&A, 0, sizeof A);
  A.F = c->field;
  A.kind = d->mode;
  A.eom = d->eom;
  A.stepper = d->stepper;
  A.nmax = d->nsteps > 0 ? d->nsteps : 500;
  A.eq = d->enforce_equatorial;
  A.check = d->check_adiabaticity;
  A.width = width;
  A.rtol = d->rtol;
  A.atol = d->atol;
  A.gc_rtol = d->gc_rtol > 0 ? d->gc_rtol : d->rtol;
  A.gc_atol = d->gc_atol > 0 ? d->gc_atol : d->atol;
  if (d->mode == RAPT_MODE_ADAPTIVE) {
    // Particle segments: scipy dop853 as set by Particle.advance; GC segments: dopri5 defaults
    A.c853 = ctl(d->safety, d->dfactor, d->ifactor, d->beta, true);
    A.c5 = ctl(0.9, 0.2, 10.0, 0.0, false);
  } else {
    A.c853 = ctl(d->safety, d->dfactor, d->ifactor, d->beta, true);
    A.c5 = ctl(d->safety, d->dfactor, d->ifactor, d->beta, false);
  }
  A.max_step = d->max_step;
  A.first_step = d->first_step;
  A.rkf_tol = d->rkf_tol;
  A.rkf_hmax = d->rkf_hmax;
  A.rkf_hmin = d->rkf_hmin;
  A.cycres = d->cyclotronresolution;
  A.gcdt = d->gctimestep;
  A.epss = d->epss;
  A.epst = d->epst;
  A.delta = d->delta;
  A.n = b->n;
  A.state = b->d_state;
  A.flags = b->d_flags;
  A.status = b->d_status;
  A.nrows = b->d_nrows;
  A.naccpt = b->d_naccpt;
  A.row_off = b->d_rowoff;
  A.row_cap = b->d_rowcap;
  A.rows = b->d_rows;
  A.work = b->d_work;
  A.order = d->work_order == 0 ? b->d_order : nullptr;
  int64_t lanes = d->lanes > 0 ? d->lanes : b->n;
  if (lanes > b->n) lanes = b->n;
  if (lanes < 1) lanes = 1;
  const bool quad = d->mode == RAPT_MODE_GC && d->stepper == RAPT_STEP_DOPRI5 &&
                    d->lanes_per_particle != 1;
  HIP_OK(c, hipMemsetAsync(b->d_work, 0, sizeof(unsigned long long) * 8, c->stream));
  HIP_OK(c, hipEventRecord(c->ev0, c->stream));
  if (quad) {
    int64_t groups = d->lanes > 0 ? (d->lanes + 3) / 4 : b->n;
    if (groups > b->n) groups = b->n;
    rapt_launch_gc4(c->stream, c->field.model, A, block, groups);
  } else {
    switch (d->mode) {
      case RAPT_MODE_PARTICLE: launch_model<KIND_PARTICLE>(c, A, block, lanes); break;
      case RAPT_MODE_GC: launch_model<KIND_GC>(c, A, block, lanes); break;
      default: launch_model<KIND_ADAPTIVE>(c, A, block, lanes); break;
    }
  }
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipEventRecord(c->ev1, c->stream));
  unsigned long long w[8];
  HIP_OK(c, hipMemcpyAsync(w, b->d_work, sizeof w, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (cnt) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    cnt->naccpt = (int64_t)w[1];
    cnt->nrejct = (int64_t)w[2];
    cnt->nstep = (int64_t)w[3];
    cnt->nfcn = (int64_t)w[4];
    cnt->nswitch = (int64_t)w[5];
    cnt->nrows = 0;
    cnt->kernel_ms = ms;
  }
  return RAPT_OK;
}

int rapt_batch_download(rapt_batch* b, double* state, int32_t* flags, int32_t* status,
                        int64_t* nrows, double* rows, int64_t* naccpt) {
  if (!b) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  const int64_t n = b->n;
  if (state)
    HIP_OK(c, hipMemcpyAsync(state, b->d_state, sizeof(double) * RAPT_NSLOT * n,
                             hipMemcpyDeviceToHost, c->stream));
  if (flags)
    HIP_OK(c, hipMemcpyAsync(flags, b->d_flags, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (status)
    HIP_OK(c, hipMemcpyAsync(status, b->d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (nrows)
    HIP_OK(c, hipMemcpyAsync(nrows, b->d_nrows, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (naccpt)
    HIP_OK(c, hipMemcpyAsync(naccpt, b->d_naccpt, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (rows && b->total_rows > 0)
    HIP_OK(c, hipMemcpyAsync(rows, b->d_rows, sizeof(double) * b->width * b->total_rows,
                             hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

int rapt_trace(rapt_ctx* c, const rapt_trace_desc* d, int64_t n, double* state, int32_t* flags,
               const int64_t* row_cap, double* rows, int64_t* nrows, int32_t* status,
               rapt_counters* cnt) {
  if (!c || !d || !state || !flags || !row_cap || !nrows || !status || n <= 0)
    return RAPT_E_ARG;
  rapt_batch* b = nullptr;
  int rc = rapt_batch_create(c, n, &b);
  if (rc) return rc;
  const int width = d->mode == RAPT_MODE_PARTICLE ? 7 : (d->mode == RAPT_MODE_GC ? 5 : 8);
  if (!(rc = rapt_batch_upload(b, state, flags)) &&
      !(rc = rapt_batch_set_rows(b, width, row_cap)) && !(rc = rapt_batch_run(b, d, cnt)))
    rc = rapt_batch_download(b, state, flags, status, nrows, rows, nullptr);
  rapt_batch_destroy(b);
  return rc;
}

}  // extern "C"
