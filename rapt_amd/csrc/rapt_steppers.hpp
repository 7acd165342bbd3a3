// rapt_steppers.hpp -- embedded Runge-Kutta pairs for one lane (FP64).
//
// DOPRI5 (Dormand-Prince 5(4)) and DOP853 (Dormand-Prince 8(5,3)) with the
// step-size control of Hairer & Wanner's codes (Solving ODEs I, 2nd ed.,
// 1993, II.4-II.5) as scipy.integrate.ode drives them
// (scipy/integrate/_ode.py:1121-1250): error norm with
// sc_i = atol + rtol*max(|y_i|, |y1_i|), the PI controller with FACOLD reset
// per integrate() call, HINIT on the first call only, and the step size
// carried from one call to the next.  RKF45 follows rapt/rkf.py:13-143.
//
// Each *_attempt performs one full step attempt (all stages) so that every
// lane of a wavefront evaluates its right-hand side at the same program point.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace rapt {

constexpr double UROUND = 2.3e-16;

// The controller's step-size factor err^expo1 / facold^beta in base-2
// logarithms.  The drivers keep lfo = log2(facold) instead of facold:
// FACOLD = max(err, 1e-4) of the last accepted step becomes
// lfo = max(log2(err), LFO0), with lfo = LFO0 = log2(1e-4) at the start of
// each integrate() call.  One log2 per step attempt and two exp2 replace two
// pow calls (on gfx950 an FP64 pow is ~220 vector instructions, log2 ~80,
// exp2 ~32); facold^beta depends on the previous step only, so a driver
// forms it before the stages, off the chain that ends the step.  The factors
// match pow to a few ulp (exp2's argument carries the rounding of
// |expo1 log2(err)| <= ~10); err = 0, inf and NaN give pow's 0, inf and NaN,
// and beta = 0 gives facold^beta = 1 exactly.
constexpr double LFO0 = -13.287712379549449;  // log2(1e-4)
__device__ __forceinline__ double ctl_fob(double beta, double lfo) {
  return beta == 0.0 ? 1.0 : exp2(beta * lfo);
}

// Magnitude that scales component i in the error norms.  G = false: |a_i|,
// scipy's componentwise sc_i = atol + rtol max(|y_i|, |y1_i|).  G = true
// (FP32-field mode): the largest |a_j| of the component's 3-vector (x, y, z |
// momentum | ppar), so a component passing through zero, like z at the
// equator, keeps its vector's scale: with float forces the rounding noise of
// a step is ~1e-7 of |v| h, which a componentwise scale rtol |z| ~ rtol |v| h
// near z = 0 can never accept.
template <int N, bool G>
__device__ __forceinline__ double err_mag(const double* a, int i) {
  if constexpr (!G) {
    return __builtin_fabs(a[i]);
  } else {
    const int g0 = (i / 3) * 3;
    double m = __builtin_fabs(a[g0]);
    if (g0 + 1 < N) m = fmax(m, __builtin_fabs(a[g0 + 1]));
    if (g0 + 2 < N) m = fmax(m, __builtin_fabs(a[g0 + 2]));
    return m;
  }
}

// ---- DOPRI5 -------------------------------------------------------------------
namespace dp5 {
constexpr double c2 = 0.2, c3 = 0.3, c4 = 0.8, c5 = 8.0 / 9.0;
constexpr double a21 = 0.2, a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                 a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
constexpr double a71 = 35.0 / 384.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0,
                 a75 = -2187.0 / 6784.0, a76 = 11.0 / 84.0;
constexpr double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0,
                 e5 = -17253.0 / 339200.0, e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
}  // namespace dp5

// One DOPRI5 step attempt from (x, y) with k1 = f(x, y).  Writes the 5th-order
// solution to y1 and f(x+h, y1) to k7; returns the scaled error norm.
template <int N, bool G = false, class RHS>
__device__ __forceinline__ double dopri5_attempt(RHS& f, double x, double h, const double* y,
                                                 const double* k1, double* y1, double* k7,
                                                 double atol, double rtol) {
  using namespace dp5;
  double k2[N], k3[N], k4[N], k5[N], k6[N], t[N], kk[N];
  // one right-hand-side call site for all six stages keeps the hot loop small
  // enough for the instruction cache; the stage index is wave-uniform
#pragma unroll 1
  for (int s = 2; s <= 7; ++s) {
    double xs;
    switch (s) {
      case 2:
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = y[i] + h * a21 * k1[i];
        xs = x + c2 * h;
        break;
      case 3:
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
        xs = x + c3 * h;
        break;
      case 4:
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
        xs = x + c4 * h;
        break;
      case 5:
#pragma unroll
        for (int i = 0; i < N; ++i)
          t[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
        xs = x + c5 * h;
        break;
      case 6:
#pragma unroll
        for (int i = 0; i < N; ++i)
          t[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] +
                             a65 * k5[i]);
        xs = x + h;
        break;
      default:
#pragma unroll
        for (int i = 0; i < N; ++i) {
          y1[i] = y[i] + h * (a71 * k1[i] + a73 * k3[i] + a74 * k4[i] + a75 * k5[i] +
                              a76 * k6[i]);
          t[i] = y1[i];
        }
        xs = x + h;
        break;
    }
    f(xs, t, kk);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (s == 2) k2[i] = kk[i];
      else if (s == 3) k3[i] = kk[i];
      else if (s == 4) k4[i] = kk[i];
      else if (s == 5) k5[i] = kk[i];
      else if (s == 6) k6[i] = kk[i];
      else k7[i] = kk[i];
    }
  }
  double err = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double e = (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] +
                      e7 * k7[i]) * h;
    const double sk = atol + rtol * fmax(err_mag<N, G>(y, i), err_mag<N, G>(y1, i));
    const double q = e / sk;
    err = __builtin_fma(q, q, err);
  }
  return __builtin_sqrt(err * (1.0 / N));
}

// ---- DOP853 -------------------------------------------------------------------
namespace dp8 {
constexpr double c2 = 0.526001519587677318785587544488e-01,
                 c3 = 0.789002279381515978178381316732e-01,
                 c4 = 0.118350341907227396726757197510e+00,
                 c5 = 0.281649658092772603273242802490e+00,
                 c6 = 0.333333333333333333333333333333e+00, c7 = 0.25e+00,
                 c8 = 0.307692307692307692307692307692e+00,
                 c9 = 0.651282051282051282051282051282e+00, c10 = 0.6e+00,
                 c11 = 0.857142857142857142857142857142e+00;
constexpr double a21 = 5.26001519587677318785587544488e-2,
                 a31 = 1.97250569845378994544595329183e-2,
                 a32 = 5.91751709536136983633785987549e-2,
                 a41 = 2.95875854768068491816892993775e-2,
                 a43 = 8.87627564304205475450678981324e-2,
                 a51 = 2.41365134159266685502369798665e-1,
                 a53 = -8.84549479328286085344864962717e-1,
                 a54 = 9.24834003261792003115737966543e-1,
                 a61 = 3.7037037037037037037037037037e-2,
                 a64 = 1.70828608729473871279604482173e-1,
                 a65 = 1.25467687566822425016691814123e-1, a71 = 3.7109375e-2,
                 a74 = 1.70252211019544039314978060272e-1,
                 a75 = 6.02165389804559606850219397283e-2, a76 = -1.7578125e-2,
                 a81 = 3.70920001185047927108779319836e-2,
                 a84 = 1.70383925712239993810214054705e-1,
                 a85 = 1.07262030446373284651809199168e-1,
                 a86 = -1.53194377486244017527936158236e-2,
                 a87 = 8.27378916381402288758473766002e-3,
                 a91 = 6.24110958716075717114429577812e-1,
                 a94 = -3.36089262944694129406857109825e0,
                 a95 = -8.68219346841726006818189891453e-1,
                 a96 = 2.75920996994467083049415600797e1,
                 a97 = 2.01540675504778934086186788979e1,
                 a98 = -4.34898841810699588477366255144e1,
                 a101 = 4.77662536438264365890433908527e-1,
                 a104 = -2.48811461997166764192642586468e0,
                 a105 = -5.90290826836842996371446475743e-1,
                 a106 = 2.12300514481811942347288949897e1,
                 a107 = 1.52792336328824235832596922938e1,
                 a108 = -3.32882109689848629194453265587e1,
                 a109 = -2.03312017085086261358222928593e-2,
                 a111 = -9.3714243008598732571704021658e-1,
                 a114 = 5.18637242884406370830023853209e0,
                 a115 = 1.09143734899672957818500254654e0,
                 a116 = -8.14978701074692612513997267357e0,
                 a117 = -1.85200656599969598641566180701e1,
                 a118 = 2.27394870993505042818970056734e1,
                 a119 = 2.49360555267965238987089396762e0,
                 a1110 = -3.0467644718982195003823669022e0,
                 a121 = 2.27331014751653820792359768449e0,
                 a124 = -1.05344954667372501984066689879e1,
                 a125 = -2.00087205822486249909675718444e0,
                 a126 = -1.79589318631187989172765950534e1,
                 a127 = 2.79488845294199600508499808837e1,
                 a128 = -2.85899827713502369474065508674e0,
                 a129 = -8.87285693353062954433549289258e0,
                 a1210 = 1.23605671757943030647266201528e1,
                 a1211 = 6.43392746015763530355970484046e-1;
constexpr double b1 = 5.42937341165687622380535766363e-2, b6 = 4.45031289275240888144113950566e0,
                 b7 = 1.89151789931450038304281599044e0, b8 = -5.8012039600105847814672114227e0,
                 b9 = 3.1116436695781989440891606237e-1,
                 b10 = -1.52160949662516078556178806805e-1,
                 b11 = 2.01365400804030348374776537501e-1,
                 b12 = 4.47106157277725905176885569043e-2;
constexpr double bhh1 = 0.244094488188976377952755905512e+00,
                 bhh2 = 0.733846688281611857341361741547e+00,
                 bhh3 = 0.220588235294117647058823529412e-01;
constexpr double er1 = 0.1312004499419488073250102996e-01,
                 er6 = -0.1225156446376204440720569753e+01,
                 er7 = -0.4957589496572501915214079952e+00,
                 er8 = 0.1664377182454986536961530415e+01,
                 er9 = -0.3503288487499736816886487290e+00,
                 er10 = 0.3341791187130174790297318841e+00,
                 er11 = 0.8192320648511571246570742613e-01,
                 er12 = -0.2235530786388629525884427845e-01;
}  // namespace dp8

// c0 k0 + c1 k1 + ... accumulated left to right in fused multiply-adds
__device__ __forceinline__ double fsum_acc(double s) { return s; }
template <class... R>
__device__ __forceinline__ double fsum_acc(double s, double c, double k, R... r) {
  return fsum_acc(__builtin_fma(c, k, s), r...);
}
template <class... R>
__device__ __forceinline__ double fsum(double c, double k, R... r) {
  return fsum_acc(c * k, r...);
}

// One DOP853 step attempt (11 stage evaluations).  Writes the 8th-order
// solution to y1; returns the combined 5th/3rd-order error norm.  The caller
// evaluates f(x+h, y1) after acceptance (the 12th evaluation).  The stage
// sums y + h (a_1 k_1 + a_2 k_2 + ...) are formed in fused multiply-adds in
// the order written (the reference's Fortran rounds each product and sum:
// the results differ by rounding only, and the sums take ~40% fewer
// instructions, about a fifth of a Lorentz-force step).
template <int N, bool G = false, class RHS>
__device__ __forceinline__ double dop853_attempt(RHS& f, double x, double h, const double* y,
                                                 const double* k1, double* y1, double atol,
                                                 double rtol) {
  using namespace dp8;
  double k2[N], k3[N], k4[N], k5[N], k6[N], k7[N], k8[N], k9[N], k10[N], t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = __builtin_fma(h, a21 * k1[i], y[i]);
  f(x + c2 * h, t, k2);
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = __builtin_fma(h, fsum(a31, k1[i], a32, k2[i]), y[i]);
  f(x + c3 * h, t, k3);
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = __builtin_fma(h, fsum(a41, k1[i], a43, k3[i]), y[i]);
  f(x + c4 * h, t, k4);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h, fsum(a51, k1[i], a53, k3[i], a54, k4[i]), y[i]);
  f(x + c5 * h, t, k5);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h, fsum(a61, k1[i], a64, k4[i], a65, k5[i]), y[i]);
  f(x + c6 * h, t, k6);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h, fsum(a71, k1[i], a74, k4[i], a75, k5[i], a76, k6[i]), y[i]);
  f(x + c7 * h, t, k7);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h, fsum(a81, k1[i], a84, k4[i], a85, k5[i], a86, k6[i], a87, k7[i]),
                         y[i]);
  f(x + c8 * h, t, k8);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(
        h, fsum(a91, k1[i], a94, k4[i], a95, k5[i], a96, k6[i], a97, k7[i], a98, k8[i]), y[i]);
  f(x + c9 * h, t, k9);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h,
                         fsum(a101, k1[i], a104, k4[i], a105, k5[i], a106, k6[i], a107, k7[i],
                              a108, k8[i], a109, k9[i]),
                         y[i]);
  f(x + c10 * h, t, k10);
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h,
                         fsum(a111, k1[i], a114, k4[i], a115, k5[i], a116, k6[i], a117, k7[i],
                              a118, k8[i], a119, k9[i], a1110, k10[i]),
                         y[i]);
  f(x + c11 * h, t, k2);  // stage 11 overwrites k2, as in Hairer's code
  const double xph = x + h;
#pragma unroll
  for (int i = 0; i < N; ++i)
    t[i] = __builtin_fma(h,
                         fsum(a121, k1[i], a124, k4[i], a125, k5[i], a126, k6[i], a127, k7[i],
                              a128, k8[i], a129, k9[i], a1210, k10[i], a1211, k2[i]),
                         y[i]);
  f(xph, t, k3);  // stage 12 overwrites k3
  double err = 0.0, err2 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double bs = fsum(b1, k1[i], b6, k6[i], b7, k7[i], b8, k8[i], b9, k9[i], b10, k10[i],
                           b11, k2[i], b12, k3[i]);
    y1[i] = __builtin_fma(h, bs, y[i]);
    const double sk = atol + rtol * fmax(err_mag<N, G>(y, i), err_mag<N, G>(y1, i));
    const double isk = 1.0 / sk;
    const double e3 = fsum_acc(bs, -bhh1, k1[i], -bhh2, k9[i], -bhh3, k3[i]) * isk;
    err2 = __builtin_fma(e3, e3, err2);
    const double e5 = fsum(er1, k1[i], er6, k6[i], er7, k7[i], er8, k8[i], er9, k9[i],
                           er10, k10[i], er11, k2[i], er12, k3[i]) * isk;
    err = __builtin_fma(e5, e5, err);
  }
  double deno = err + 0.01 * err2;
  if (deno <= 0.0) deno = 1.0;
  return __builtin_fabs(h) * err * __builtin_sqrt(1.0 / (N * deno));
}

// ---- HINIT (initial step guess) ----------------------------------------------
template <int N, bool G = false, class RHS>
__device__ __forceinline__ double hinit(RHS& f, double x, const double* y, const double* f0,
                                        int iord, double hmax, double atol, double rtol) {
  double dnf = 0.0, dny = 0.0;
  double sk[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    sk[i] = atol + rtol * err_mag<N, G>(y, i);
    const double a = f0[i] / sk[i], b = y[i] / sk[i];
    dnf = __builtin_fma(a, a, dnf);
    dny = __builtin_fma(b, b, dny);
  }
  double h = (dnf <= 1.0e-10 || dny <= 1.0e-10) ? 1.0e-6 : __builtin_sqrt(dny / dnf) * 0.01;
  h = fmin(h, hmax);
  double y1[N], f1[N];
#pragma unroll
  for (int i = 0; i < N; ++i) y1[i] = y[i] + h * f0[i];
  f(x + h, y1, f1);
  double der2 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double q = (f1[i] - f0[i]) / sk[i];
    der2 = __builtin_fma(q, q, der2);
  }
  der2 = __builtin_sqrt(der2) / h;
  const double der12 = fmax(__builtin_fabs(der2), __builtin_sqrt(dnf));
  const double h1 = (der12 <= 1.0e-15) ? fmax(1.0e-6, __builtin_fabs(h) * 1.0e-3)
                                       : pow(0.01 / der12, 1.0 / iord);
  return fmin(100.0 * __builtin_fabs(h), fmin(h1, hmax));
}

// ---- RKF45 (rapt/rkf.py:13-143) ---------------------------------------------------
namespace rkf {
constexpr double a2 = 2.500000000000000e-01, a3 = 3.750000000000000e-01,
                 a4 = 9.230769230769231e-01, a5 = 1.000000000000000e+00,
                 a6 = 5.000000000000000e-01;
constexpr double b21 = 2.500000000000000e-01, b31 = 9.375000000000000e-02,
                 b32 = 2.812500000000000e-01, b41 = 8.793809740555303e-01,
                 b42 = -3.277196176604461e+00, b43 = 3.320892125625853e+00,
                 b51 = 2.032407407407407e+00, b52 = -8.000000000000000e+00,
                 b53 = 7.173489278752436e+00, b54 = -2.058966861598441e-01,
                 b61 = -2.962962962962963e-01, b62 = 2.000000000000000e+00,
                 b63 = -1.381676413255361e+00, b64 = 4.529727095516569e-01,
                 b65 = -2.750000000000000e-01;
constexpr double r1 = 2.777777777777778e-03, r3 = -2.994152046783626e-02,
                 r4 = -2.919989367357789e-02, r5 = 2.000000000000000e-02,
                 r6 = 3.636363636363636e-02;
constexpr double c1 = 1.157407407407407e-01, c3 = 5.489278752436647e-01,
                 c4 = 5.353313840155945e-01, c5 = -2.000000000000000e-01;
}  // namespace rkf

// One RKF45 step attempt; y1 receives the 4th-order update, returns the
// reference's error measure max_i |r . k| / h.
template <int N, class RHS>
__device__ __forceinline__ double rkf45_attempt(RHS& f, double t, double h, const double* x,
                                                double* y1) {
  using namespace rkf;
  double k1[N], k2[N], k3[N], k4[N], k5[N], k6[N], tmp[N];
  f(t, x, k1);
#pragma unroll
  for (int i = 0; i < N; ++i) { k1[i] *= h; tmp[i] = x[i] + b21 * k1[i]; }
  f(t + a2 * h, tmp, k2);
#pragma unroll
  for (int i = 0; i < N; ++i) { k2[i] *= h; tmp[i] = x[i] + b31 * k1[i] + b32 * k2[i]; }
  f(t + a3 * h, tmp, k3);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    k3[i] *= h;
    tmp[i] = x[i] + b41 * k1[i] + b42 * k2[i] + b43 * k3[i];
  }
  f(t + a4 * h, tmp, k4);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    k4[i] *= h;
    tmp[i] = x[i] + b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
  }
  f(t + a5 * h, tmp, k5);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    k5[i] *= h;
    tmp[i] = x[i] + b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
  }
  f(t + a6 * h, tmp, k6);
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    k6[i] *= h;
    r = fmax(r, __builtin_fabs(r1 * k1[i] + r3 * k3[i] + r4 * k4[i] + r5 * k5[i] + r6 * k6[i]) / h);
    y1[i] = x[i] + c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
  }
  return r;
}

}  // namespace rapt
