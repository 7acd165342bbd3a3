// rapt_bounce_kernel.hpp -- field-line kernels (half-bounce path, eye,
// surrogate labels, BounceCenter), shared by rapt_bounce.hip (the built-in
// field models) and the kernels that rapt_jit.hip compiles at run time for
// user-written fields.
//
// Half-bounce path length S_b on the device, the output
// step of GuidingCenter.advance (bounceperiod()/bounceresolution,
// rapt/GuidingCenter.py:443-446, 593-606).
//
// Per lane (one particle):
//  1. trace the field line both ways as rapt/fieldline.py:37-105 does: chunks
//     of length ds = 1/(curvature * fieldlineresolution), each integrated by
//     RKF45 (rapt/rkf.py) with tol 1e-3 forward and 1e-4 backward, hmax = ds,
//     hmin = 1e-6, stopping after the chunk whose end has |B| > Bm; a chunk
//     that fails the hmin test ends the half line without its points;
//  2. keep the points inside B <= Bm plus one beyond each mirror point
//     (rapt/flutils.py:282-293); three points use the reference's equatorial
//     formula pi sqrt(2 Bm / B'') (:295-305);
//  3. otherwise interpolate B(s) by the quadratic spline scipy's
//     interp1d(kind='quadratic') builds (not-a-knot knots at the data
//     midpoints without the first and last, tridiagonal collocation), find
//     the mirror points with Brent's method in the end intervals, and
//     integrate ds / sqrt(1 - B/Bm) between them.  The integral uses the
//     substitution s = (a+b)/2 - (b-a)/2 cos(theta), which removes the
//     inverse-square-root end-point singularities, and 32 panels of 16-point
//     Gauss-Legendre; the reference calls QUADPACK with epsrel 1e-4.
// Points live in a scratch buffer laid out [point][lane] so that lanes at
// the same step touch consecutive addresses.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_kargs.hpp"
#include "rapt_steppers.hpp"

namespace rapt {

static constexpr double GL16_X[8] = {0.09501250983763745, 0.2816035507792589,
                                     0.45801677765722737, 0.6178762444026438,
                                     0.755404408355003,   0.8656312023878318,
                                     0.9445750230732326,  0.9894009349916499};
static constexpr double GL16_W[8] = {0.18945061045506859, 0.1826034150449236,
                                     0.16915651939500262, 0.14959598881657676,
                                     0.12462897125553403, 0.09515851168249259,
                                     0.062253523938647706, 0.027152459411754037};

enum { BST_OK = 1, BST_OVERFLOW = -20, BST_NOBRACKET = -21, BST_ROOT = -22, BST_NAN = -23 };

struct BounceArgs {
  FieldC F;
  int64_t n;
  const double* tpos;
  const double* Bm;
  double res;
  int32_t cap;   // points per half line
  int32_t what;  // k_halfbounce output: 0 S_b, 1 the second invariant I
  int64_t lanes; // scratch stride
  double* S;     // [2*cap][lanes] arc length
  double* Bv;    // [2*cap][lanes] |B|
  double* Cc;    // [2*cap][lanes] spline coefficients
  double* Wk;    // [2*cap][lanes] elimination workspace
  double* Zs;    // [2*cap][lanes] z of the trace points (labels only, else nullptr)
  double* Rs;    // [2*cap][lanes] cylindrical radius of the trace points (labels only)
  double* out;
  int32_t* status;
};

template <int M>
struct LineRHS {
  const FieldC& F;
  double t, sign;
  __device__ __forceinline__ void operator()(double, const double* Y, double* o) {
    const V3 b = unitb<M>(F, t, Y[1], Y[2], Y[3]);
    o[0] = sign;
    o[1] = sign * b.x;
    o[2] = sign * b.y;
    o[3] = sign * b.z;
  }
};

// one chunk of Fieldline.trace: rkf(deriv, 0, ds, Y, tol, hmax=ds, hmin=1e-6).
// Accepted points are written at idx(base + dir*(np+1+j)); *np advances only
// when the chunk completes.  Returns 1 ok, 0 hmin failure, -1 overflow.
template <int M>
__device__ int line_chunk(const BounceArgs& A, int64_t lane, double t, double sign, double tol,
                          double ds, double* Y, int32_t* np, int32_t base, int32_t dir) {
  LineRHS<M> f{A.F, t, sign};
  double tl = 0.0, h = ds, Y1[4];
  int32_t m = *np;
  while (tl < ds) {
    if (tl + h > ds) h = ds - tl;
    const double r = rkf45_attempt<4>(f, tl, h, Y, Y1);
    if (r <= tol) {
      tl = tl + h;
      Y[0] = Y1[0]; Y[1] = Y1[1]; Y[2] = Y1[2]; Y[3] = Y1[3];
      ++m;
      if (m >= A.cap) return -1;
      const int64_t k = base + dir * m;
      A.S[k * A.lanes + lane] = Y[0];
      A.Bv[k * A.lanes + lane] = magB<M>(A.F, t, Y[1], Y[2], Y[3]);
      if (A.Zs) {
        A.Zs[k * A.lanes + lane] = Y[3];
        A.Rs[k * A.lanes + lane] = __builtin_sqrt(Y[1] * Y[1] + Y[2] * Y[2]);
      }
    }
    h = h * fmin(fmax(0.84 * pow(tol / r, 0.25), 0.1), 4.0);
    if (h > ds) {
      h = ds;
    } else if (h < 1e-6) {
      return 0;
    }
  }
  *np = m;
  return 1;
}

// quadratic spline on x[lo..lo+nt-1] (array index k -> value at k*lanes+lane)
struct Spline {
  const double* X;
  const double* C;
  int64_t lanes, lane, lo;
  int32_t nt;
  __device__ __forceinline__ double x(int32_t j) const { return X[(lo + j) * lanes + lane]; }
  // knot i of the not-a-knot vector for k = 2
  __device__ __forceinline__ double knot(int32_t i) const {
    if (i <= 2) return x(0);
    if (i >= nt) return x(nt - 1);
    return (x(i - 2) + x(i - 1)) / 2;
  }
  // de Boor's BSPLVB for k = 2 on [t_l, t_{l+1})
  __device__ __forceinline__ void basis(int32_t l, double s, double b[3]) const {
    double left[3], right[3];
    b[0] = 1.0;
    for (int j = 1; j <= 2; ++j) {
      left[j] = s - knot(l + 1 - j);
      right[j] = knot(l + j) - s;
      double saved = 0.0;
      for (int r = 0; r < j; ++r) {
        const double temp = b[r] / (right[r + 1] + left[j - r]);
        b[r] = saved + right[r + 1] * temp;
        saved = left[j - r] * temp;
      }
      b[j] = saved;
    }
  }
  __device__ double eval(double s) const {
    int32_t lo_i = 2, hi_i = nt - 1;  // find l in [2, nt-1] with t_l <= s < t_{l+1}
    while (hi_i - lo_i > 1) {
      const int32_t mid = (lo_i + hi_i) / 2;
      if (knot(mid) <= s) lo_i = mid; else hi_i = mid;
    }
    int32_t l = (knot(hi_i) <= s) ? hi_i : lo_i;
    double b[3];
    basis(l, s, b);
    return b[0] * C[(lo + l - 2) * lanes + lane] + b[1] * C[(lo + l - 1) * lanes + lane] +
           b[2] * C[(lo + l) * lanes + lane];
  }
};

// Brent's method (R. P. Brent, Algorithms for Minimization without
// Derivatives, 1973, ch. 4) with scipy.optimize.brentq's default tolerances
template <class Fn>
__device__ bool brent(Fn& f, double a, double b, double& root) {
  const double xtol = 2e-12, rtol = 8.881784197001252e-16;
  double fa = f(a), fb = f(b);
  if (fa == 0.0) { root = a; return true; }
  if (fb == 0.0) { root = b; return true; }
  if ((fa > 0.0) == (fb > 0.0)) return false;
  double c = a, fc = fa, d = b - a, e = d;
  for (int it = 0; it < 100; ++it) {
    if ((fb > 0.0) == (fc > 0.0)) { c = a; fc = fa; d = b - a; e = d; }
    if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
    const double tol1 = xtol + rtol * fabs(b);
    const double m = 0.5 * (c - b);
    if (fabs(m) <= tol1 || fb == 0.0) { root = b; return true; }
    if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
      double p, q, r;
      const double s = fb / fa;
      if (a == c) {
        p = 2.0 * m * s;
        q = 1.0 - s;
      } else {
        q = fa / fc;
        r = fb / fc;
        p = s * (2.0 * m * q * (q - r) - (b - a) * (r - 1.0));
        q = (q - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (p > 0.0) q = -q; else p = -p;
      if (2.0 * p < fmin(3.0 * m * q - fabs(tol1 * q), fabs(e * q))) {
        e = d;
        d = p / q;
      } else {
        d = m;
        e = m;
      }
    } else {
      d = m;
      e = m;
    }
    a = b;
    fa = fb;
    b += (fabs(d) > tol1) ? d : (m > 0.0 ? tol1 : -tol1);
    fb = f(b);
  }
  root = b;
  return true;
}

// Integrals over the field line through (x0, y0, z0) between the mirror
// points of B = Bm: sb = int ds / sqrt(1 - B/Bm) (flutils.halfbouncepath,
// rapt/flutils.py:254-316), eye = int sqrt(1 - B/Bm) ds (flutils.eye,
// rapt/flutils.py:65-151; returns 0 where the reference does: Bm below the
// minimum of B on the line, or an equatorial particle to 1e-12), and lat =
// latitude (rad) of the mirror point along +b (when A.Zs is set).  The eye
// integral uses the same quadratic spline and Gauss-Legendre quadrature as
// sb for every pitch angle; the reference switches to Simpson's rule on the
// trace points below 70 degrees, a rule whose discretisation error depends on
// where the trace points fall and so adds jitter to finite differences of I.
template <int M>
__device__ __forceinline__ int32_t line_integrals(const BounceArgs& A, int64_t lane, double t, double x0,
                                  double y0, double z0, double Bm, double* sb_out,
                                  double* eye_out, double* lat_out) {
  const FieldC& F = A.F;
  const int32_t base = A.cap;  // array index of the start point
  const double ds = 1.0 / curvature<M>(F, t, x0, y0, z0) / A.res;
  // A chunk length that is not finite and positive (a NaN start point or
  // mirror field, a zero or infinite curvature such as |B| = 0 or a straight
  // line) would let a chunk end without adding a point, and the trace below
  // would never reach its mirror field.
  if (!(ds > 0.0 && ds < __builtin_inf()) || !(Bm > 0.0 && Bm < __builtin_inf())) {
    if (sb_out) *sb_out = __builtin_nan("");
    if (eye_out) *eye_out = __builtin_nan("");
    if (lat_out) *lat_out = __builtin_nan("");
    return BST_NAN;
  }
  A.S[(int64_t)base * A.lanes + lane] = 0.0;
  A.Bv[(int64_t)base * A.lanes + lane] = magB<M>(F, t, x0, y0, z0);
  if (A.Zs) {
    A.Zs[(int64_t)base * A.lanes + lane] = z0;
    A.Rs[(int64_t)base * A.lanes + lane] = __builtin_sqrt(x0 * x0 + y0 * y0);
  }
  int32_t st = BST_OK;
  int32_t nf = 0, nb = 0;
  for (int side = 0; side < 2 && st == BST_OK; ++side) {
    const double sign = side == 0 ? 1.0 : -1.0;
    const double tol = side == 0 ? 1e-3 : 1e-4;
    const int32_t dir = side == 0 ? 1 : -1;
    int32_t* np = side == 0 ? &nf : &nb;
    double Y[4] = {0.0, x0, y0, z0};
    for (;;) {
      const int32_t before = *np;
      const int rc = line_chunk<M>(A, lane, t, sign, tol, ds, Y, np, base, dir);
      if (rc < 0) { st = BST_OVERFLOW; break; }
      if (rc == 0) break;
      if (*np == before) { st = BST_NAN; break; }  // no progress: cannot reach Bm
      if (magB<M>(F, t, Y[1], Y[2], Y[3]) > Bm) break;
    }
  }
  double sb = __builtin_nan(""), eye = __builtin_nan(""), lat = __builtin_nan("");
  if (st == BST_OK) {
    // curve = array indices [base-nb, base+nf]; keep B <= Bm plus one each side
    const int32_t lo = base - nb, hi = base + nf;
    int32_t first = -1, lastin = -1;
    double bmin = A.Bv[(int64_t)lo * A.lanes + lane];
    for (int32_t k = lo; k <= hi; ++k) {
      const double bk = A.Bv[(int64_t)k * A.lanes + lane];
      bmin = fmin(bmin, bk);
      if (bk <= Bm) {
        if (first < 0) first = k;
        lastin = k;
      }
    }
    int32_t i1, i2;
    const int32_t ncur = hi - lo + 1;
    if (first < 0) {
      i1 = lo + (ncur - 3) / 2;
      i2 = lo + (ncur + 1) / 2;
    } else {
      i1 = first - 1 < lo ? lo : first - 1;
      i2 = lastin + 1 > hi ? hi : lastin + 1;
    }
    const int32_t nt = i2 - i1 + 1;
    auto X = [&](int32_t j) { return A.S[(int64_t)(i1 + j) * A.lanes + lane]; };
    auto Yv = [&](int32_t j) { return A.Bv[(int64_t)(i1 + j) * A.lanes + lane]; };
    const bool eye_zero = bmin > Bm || __builtin_fabs(bmin - Bm) / Bm < 1e-12;
    if (nt == 3) {
      // equatorial: B(s) ~ B(s0) + B''/2 (s - s0)^2 from the Lagrange parabola
      const double s12 = X(0) - X(1), s23 = X(1) - X(2), s13 = X(0) - X(2);
      const double B2s = 2 * (Yv(0) * s23 - Yv(1) * s13 + Yv(2) * s12) / (s12 * s13 * s23);
      sb = 3.141592653589793 * __builtin_sqrt(2 * Bm / B2s);
      const double a2 = fmax(2.0 * (Bm - Yv(1)) / B2s, 0.0);
      eye = eye_zero ? 0.0 : 0.5 * 3.141592653589793 * a2 * __builtin_sqrt(B2s / (2.0 * Bm));
      if (A.Zs) {
        const double r1 = A.Rs[(int64_t)(i1 + 1) * A.lanes + lane];
        lat = atan(__builtin_sqrt(a2) / r1);
      }
    } else if (nt < 3) {
      st = BST_NOBRACKET;
    } else {
      Spline sp{A.S, A.Cc, A.lanes, lane, i1, nt};
      // tridiagonal collocation system (row j: N_{j-1}, N_j, N_{j+1} at x_j)
      double cprev = 0.0, dprev = Yv(0);
      A.Wk[(int64_t)i1 * A.lanes + lane] = 0.0;
      A.Cc[(int64_t)i1 * A.lanes + lane] = dprev;
      for (int32_t j = 1; j < nt; ++j) {
        double sub, dia, sup;
        if (j == nt - 1) {
          sub = 0.0; dia = 1.0; sup = 0.0;
        } else {
          double b[3];
          sp.basis(j + 1, X(j), b);
          sub = b[0]; dia = b[1]; sup = b[2];
        }
        const double m = dia - sub * cprev;
        cprev = sup / m;
        dprev = (Yv(j) - sub * dprev) / m;
        A.Wk[(int64_t)(i1 + j) * A.lanes + lane] = cprev;
        A.Cc[(int64_t)(i1 + j) * A.lanes + lane] = dprev;
      }
      for (int32_t j = nt - 2; j >= 0; --j) {
        const int64_t kj = (int64_t)(i1 + j) * A.lanes + lane;
        const int64_t kn = (int64_t)(i1 + j + 1) * A.lanes + lane;
        A.Cc[kj] = A.Cc[kj] - A.Wk[kj] * A.Cc[kn];
      }
      auto g = [&](double s) { return sp.eval(s) - Bm; };
      double sm1, sm2;
      if (!brent(g, X(0), X(1), sm1) || !brent(g, X(nt - 2), X(nt - 1), sm2)) {
        st = BST_ROOT;
      } else {
        // s = mid - half cos(theta) removes the end-point singularities
        const double mid = 0.5 * (sm1 + sm2), half = 0.5 * (sm2 - sm1);
        const int P = 32;
        const double pw = 3.141592653589793 / P;
        double acc = 0.0, acc_i = 0.0;
        for (int p = 0; p < P; ++p) {
          const double c0 = (p + 0.5) * pw;
          for (int q = 0; q < 16; ++q) {
            const double xg = (q < 8) ? -GL16_X[7 - q] : GL16_X[q - 8];
            const double wg = (q < 8) ? GL16_W[7 - q] : GL16_W[q - 8];
            const double th = c0 + 0.5 * pw * xg;
            const double s = mid - half * cos(th);
            const double val = 1.0 - sp.eval(s) / Bm;
            const double sn = sin(th);
            acc += wg * sn / __builtin_sqrt(val);
            if (eye_out) acc_i += wg * sn * __builtin_sqrt(fmax(val, 0.0));
          }
        }
        sb = acc * 0.5 * pw * half;
        eye = eye_zero ? 0.0 : acc_i * 0.5 * pw * half;
        if (!__builtin_isfinite(sb)) st = BST_NAN;
        if (A.Zs) {
          // quadratic Lagrange interpolation of (z, rho) at the +b mirror point
          double zz = 0.0, rr = 0.0;
          for (int a = 0; a < 3; ++a) {
            const int32_t ja = nt - 3 + a;
            double w = 1.0;
            for (int b2 = 0; b2 < 3; ++b2)
              if (b2 != a) w *= (sm2 - X(nt - 3 + b2)) / (X(ja) - X(nt - 3 + b2));
            zz += w * A.Zs[(int64_t)(i1 + ja) * A.lanes + lane];
            rr += w * A.Rs[(int64_t)(i1 + ja) * A.lanes + lane];
          }
          lat = atan2(zz, rr);
        }
      }
    }
  }
  if (sb_out) *sb_out = sb;
  if (eye_out) *eye_out = eye;
  if (lat_out) *lat_out = lat;
  return st;
}

template <int M>
__global__ __launch_bounds__(256) void k_halfbounce(BounceArgs A) {
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= A.lanes) return;
  for (int64_t i = lane; i < A.n; i += A.lanes) {
    double sb, ic = 0.0;
    const int32_t st = line_integrals<M>(A, lane, A.tpos[4 * i], A.tpos[4 * i + 1],
                                         A.tpos[4 * i + 2], A.tpos[4 * i + 3], A.Bm[i], &sb,
                                         A.what ? &ic : nullptr, nullptr);
    A.out[i] = A.what ? ic : sb;
    A.status[i] = st;
  }
}

// Surrogate labels from one bounce-averaged drift evaluation per query
// (BounceCenter.advance's deriv, rapt/BounceCenter.py:235-244, at t = 0):
// query (L, equatorial pitch angle alpha [deg], log10 K [eV]) starts at the
// equatorial point (L Re, 0, 0); v from K (utils.speedfromKE), Bm =
// B_eq / sin^2 alpha, T_b = 2 S_b / v (flutils.bounceperiod), grad I by
// flutils.gradI (rapt/flutils.py:153-229: central differences of eye across
// the field, step eyegradientstep, one-sided where an I is 0), v_d =
// gamma m v^2 / (q S_b B^2) grad I x B, omega_d = v_d . phi_hat / rho.
// out[3 i + 0..2] = (T_b [s], omega_d [rad/s], mirror latitude [deg]).
struct LabelArgs {
  double mass, charge, dstep;
  double Re, eV;  // Earth radius [m] and joules per eV, as the caller's constants
};

template <int M>
__global__ __launch_bounds__(256) void k_bounce_labels(BounceArgs A, LabelArgs L) {
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= A.lanes) return;
  const FieldC& F = A.F;
  const double Re = L.Re;
  const double mc2 = L.mass * C_LIGHT * C_LIGHT;
  for (int64_t i = lane; i < A.n; i += A.lanes) {
    const double Lsh = A.tpos[3 * i], alpha = A.tpos[3 * i + 1] * (3.141592653589793 / 180.0);
    const double K = pow(10.0, A.tpos[3 * i + 2]) * L.eV;
    // speedfromKE (rapt/utils.py:218-249): non-relativistic below K/mc2 = 1e-6
    double v;
    if (K / mc2 < 1e-6) {
      v = __builtin_sqrt(2.0 * K / L.mass);
    } else {
      const double g = 1.0 + K / mc2;
      v = C_LIGHT * __builtin_sqrt(1.0 - 1.0 / (g * g));
    }
    const double gamma = 1.0 / __builtin_sqrt(1.0 - (v / C_LIGHT) * (v / C_LIGHT));
    const double x0 = Lsh * Re, y0 = 0.0, z0 = 0.0, t = 0.0;
    const V3 B = fieldB<M>(F, t, x0, y0, z0);
    const double bsq = dot3(B, B);
    const double sa = sin(alpha);
    const double Bm = __builtin_sqrt(bsq) / (sa * sa);
    double sb, lat, ic;
    int32_t st = line_integrals<M>(A, lane, t, x0, y0, z0, Bm, &sb, &ic, &lat);
    // gradI directions (rapt/flutils.py:196-207)
    const double mx = fmax(__builtin_fabs(x0), __builtin_fabs(y0));
    V3 r{x0 / mx, y0 / mx, 0.0};
    const double ib = 1.0 / __builtin_sqrt(bsq);
    const V3 b{B.x * ib, B.y * ib, B.z * ib};
    const double rb = dot3(r, b);
    r = V3{r.x - rb * b.x, r.y - rb * b.y, r.z - rb * b.z};
    const double rn = 1.0 / __builtin_sqrt(dot3(r, r));
    const V3 v1{r.x * rn, r.y * rn, r.z * rn};
    const V3 v2 = cross3(b, v1);
    double dI[2];
    for (int k = 0; k < 2 && st == BST_OK; ++k) {
      const V3 u = k == 0 ? v1 : v2;
      double Ip, Imn;
      st = line_integrals<M>(A, lane, t, x0 + L.dstep * u.x, y0 + L.dstep * u.y,
                             z0 + L.dstep * u.z, Bm, nullptr, &Ip, nullptr);
      if (st != BST_OK) break;
      st = line_integrals<M>(A, lane, t, x0 - L.dstep * u.x, y0 - L.dstep * u.y,
                             z0 - L.dstep * u.z, Bm, nullptr, &Imn, nullptr);
      if (st != BST_OK) break;
      if (Ip == 0.0) dI[k] = (ic - Imn) / L.dstep;
      else if (Imn == 0.0) dI[k] = (Ip - ic) / L.dstep;
      else dI[k] = (Ip - Imn) / (2.0 * L.dstep);
    }
    double Tb = __builtin_nan(""), wd = __builtin_nan(""), latd = __builtin_nan("");
    if (st == BST_OK) {
      const V3 gI{dI[0] * v1.x + dI[1] * v2.x, dI[0] * v1.y + dI[1] * v2.y,
                  dI[0] * v1.z + dI[1] * v2.z};
      const double coef = gamma * L.mass * v * v / (L.charge * sb * bsq);
      const V3 c = cross3(gI, B);
      const double rho = __builtin_sqrt(x0 * x0 + y0 * y0);
      // phi_hat = (-y, x, 0) / rho
      wd = coef * (c.y * x0 - c.x * y0) / (rho * rho);
      Tb = 2.0 * sb / v;
      latd = lat * (180.0 / 3.141592653589793);
    }
    A.out[3 * i] = Tb;
    A.out[3 * i + 1] = wd;
    A.out[3 * i + 2] = latd;
    A.status[i] = st;
  }
}

// ---- BounceCenter.advance (rapt/BounceCenter.py:206-251) -----------------------
// The drift right-hand side of the reference's deriv (:235-244, the
// non-equatorial branch): gamma m v^2 / (q S_b |B|^2) grad I x B, with S_b and
// grad I from the field-line integrals above.  A failed line integral stores
// its status and makes the derivative NaN.
template <int M>
struct DriftRHS {
  const BounceArgs& A;
  int64_t lane;
  double Bm, coef, dstep;
  int32_t st;
  int64_t nf;
  __device__ __forceinline__ void operator()(double t, const double* Y, double* out) {
    ++nf;
    const FieldC& F = A.F;
    const V3 B = fieldB<M>(F, t, Y[0], Y[1], Y[2]);
    const double bsq = dot3(B, B);
    double sb, ic;
    int32_t s = line_integrals<M>(A, lane, t, Y[0], Y[1], Y[2], Bm, &sb, &ic, nullptr);
    // gradI directions (rapt/flutils.py:196-207)
    const double mx = fmax(__builtin_fabs(Y[0]), __builtin_fabs(Y[1]));
    V3 r{Y[0] / mx, Y[1] / mx, 0.0};
    const double ib = 1.0 / __builtin_sqrt(bsq);
    const V3 b{B.x * ib, B.y * ib, B.z * ib};
    const double rb = dot3(r, b);
    r = V3{r.x - rb * b.x, r.y - rb * b.y, r.z - rb * b.z};
    const double rn = 1.0 / __builtin_sqrt(dot3(r, r));
    const V3 v1{r.x * rn, r.y * rn, r.z * rn};
    const V3 v2 = cross3(b, v1);
    double dI[2] = {0.0, 0.0};
    for (int k = 0; k < 2 && s == BST_OK; ++k) {
      const V3 u = k == 0 ? v1 : v2;
      double Ip = 0.0, Imn = 0.0;
      s = line_integrals<M>(A, lane, t, Y[0] + dstep * u.x, Y[1] + dstep * u.y,
                            Y[2] + dstep * u.z, Bm, nullptr, &Ip, nullptr);
      if (s != BST_OK) break;
      s = line_integrals<M>(A, lane, t, Y[0] - dstep * u.x, Y[1] - dstep * u.y,
                            Y[2] - dstep * u.z, Bm, nullptr, &Imn, nullptr);
      if (s != BST_OK) break;
      if (Ip == 0.0) dI[k] = (ic - Imn) / dstep;
      else if (Imn == 0.0) dI[k] = (Ip - ic) / dstep;
      else dI[k] = (Ip - Imn) / (2.0 * dstep);
    }
    if (s != BST_OK) {
      if (st == BST_OK) st = s;
      out[0] = out[1] = out[2] = __builtin_nan("");
      return;
    }
    const V3 gI{dI[0] * v1.x + dI[1] * v2.x, dI[0] * v1.y + dI[1] * v2.y,
                dI[0] * v1.z + dI[1] * v2.z};
    const double c = coef / (sb * bsq);
    const V3 x = cross3(gI, B);
    out[0] = c * x.x;
    out[1] = c * x.y;
    out[2] = c * x.z;
  }
};

struct BCArgs {
  int64_t n;
  double* tpos;          // [n][4] (t, x, y, z): start in, end out
  const double* Bm;      // [n] mirror field
  const double* coef;    // [n] gamma m v^2 / q
  const double* dt;      // [n] output step
  const double* tcur;    // [n] first row time (np.arange start)
  const int64_t* nrow;   // [n] rows to write
  const int64_t* row_off;
  double* rows;          // [.][4]
  int32_t* status;
  int64_t* naccpt;
  double rtol, atol, dstep;
  int32_t nmax, pad;
  CtlParams c5;
};

// The equatorial branch of deriv (rapt/BounceCenter.py:237-238): gradient
// plus inertial drift, _v_grad + _v_inert (:154-204), for bounce centres
// with no parallel motion.  _v_grad = mu (B x grad|B|) / (q gamma |B|^2);
// _v_inert = m |v_g| / (q |B|^2 2d) B x (v_g(+) - v_g(-)), the gradient drift
// differenced at +-d along its own direction, where the reference shifts the
// whole 4-vector (t, x, y, z) by d (t, v_g/|v_g|) -- the time shift included.
// The reference calls both with Y[:4], the 3-vector (x, y, z), which the
// fields' `t, x, y, z = tpos` rejects; here they take (t, x, y, z).
template <int M>
struct EqDriftRHS {
  const FieldC& F;
  double mu, mass, charge, gamma;
  int32_t st;
  int64_t nf;
  __device__ __forceinline__ V3 vgrad(double t, double x, double y, double z) const {
    const V3 B = fieldB<M>(F, t, x, y, z);
    const double bsq = dot3(B, B);
    const double d = F.gstep;
    // _Field.gradB (rapt/fields.py:111-131): central differences of |B|
    const double gx = (magB<M>(F, t, x + d, y, z) - magB<M>(F, t, x - d, y, z)) / (2.0 * d);
    const double gy = (magB<M>(F, t, x, y + d, z) - magB<M>(F, t, x, y - d, z)) / (2.0 * d);
    const double gz = (magB<M>(F, t, x, y, z + d) - magB<M>(F, t, x, y, z - d)) / (2.0 * d);
    const V3 c = cross3(B, V3{gx, gy, gz});
    const double den = charge * gamma * bsq;
    return V3{mu * c.x / den, mu * c.y / den, mu * c.z / den};
  }
  __device__ __forceinline__ void operator()(double t, const double* Y, double* out) {
    ++nf;
    const V3 vg = vgrad(t, Y[0], Y[1], Y[2]);
    const double vm = __builtin_sqrt(dot3(vg, vg));
    const double d = F.gstep;
    const double ux = vg.x / vm, uy = vg.y / vm, uz = vg.z / vm;
    const V3 v1 = vgrad(t - d * t, Y[0] - d * ux, Y[1] - d * uy, Y[2] - d * uz);
    const V3 v2 = vgrad(t + d * t, Y[0] + d * ux, Y[1] + d * uy, Y[2] + d * uz);
    const V3 B = fieldB<M>(F, t, Y[0], Y[1], Y[2]);
    const double s = mass * vm / (charge * dot3(B, B) * 2.0 * d);
    const V3 c = cross3(B, V3{v2.x - v1.x, v2.y - v1.y, v2.z - v1.z});
    out[0] = vg.x + s * c.x;
    out[1] = vg.y + s * c.y;
    out[2] = vg.z + s * c.z;
  }
};

// One lane per bounce centre: r = ode(deriv).set_integrator("dopri5", ...),
// then for each t in arange(tcur, tcur + delta, dt): integrate(r.t + dt) and
// append [t, r.y] (the row's time is the arange value, one step behind the
// state, as the reference writes it).  scipy's call semantics as in k_trace.
template <class RHS>
__device__ __forceinline__ void bc_run(RHS& f, const BCArgs& C, int64_t i) {
  const CtlParams& P = C.c5;
  {
    double x = C.tpos[4 * i];
    double y[3] = {C.tpos[4 * i + 1], C.tpos[4 * i + 2], C.tpos[4 * i + 3]};
    double k1[3], y1[3], k7[3];
    f(x, y, k1);
    double h = 0.0;
    int32_t code = 1;
    int64_t acc = 0;
    const double dt = C.dt[i];
    for (int64_t r = 0; r < C.nrow[i] && code == 1; ++r) {
      const double xend = x + dt;
      const double hmax = xend - x;
      if (h == 0.0) h = hinit<3>(f, x, y, k1, 5, hmax, C.atol, C.rtol);
      double facold = 1.0e-4;
      int nstep = 0;
      bool last = false, reject = false;
      for (;;) {
        if (f.st != BST_OK) { code = f.st; break; }
        if (nstep > C.nmax) { code = RAPT_ST_NSTEPS; break; }
        if (0.1 * __builtin_fabs(h) <= __builtin_fabs(x) * UROUND) { code = RAPT_ST_HSMALL; break; }
        if (x + 1.01 * h - xend > 0.0) {
          h = xend - x;
          last = true;
        }
        ++nstep;
        const double err = dopri5_attempt<3>(f, x, h, y, k1, y1, k7, C.atol, C.rtol);
        if (f.st != BST_OK) { code = f.st; break; }
        const double fac11 = pow(err, P.expo1);
        double fac = fac11 / pow(facold, P.beta);
        fac = fmax(P.facc2, fmin(P.facc1, fac / P.safe));
        double hnew = h / fac;
        if (err <= 1.0) {
          facold = fmax(err, 1.0e-4);
          ++acc;
          for (int k = 0; k < 3; ++k) {
            k1[k] = k7[k];
            y[k] = y1[k];
          }
          x = x + h;
          if (last) {
            h = hnew;
            break;
          }
          if (__builtin_fabs(hnew) > hmax) hnew = hmax;
          if (reject) hnew = fmin(__builtin_fabs(hnew), __builtin_fabs(h));
          reject = false;
          h = hnew;
        } else {
          hnew = h / fmin(P.facc1, fac11 / P.safe);
          reject = true;
          last = false;
          h = hnew;
        }
      }
      double* row = C.rows + (C.row_off[i] + r) * 4;
      row[0] = C.tcur[i] + (double)r * dt;
      row[1] = y[0];
      row[2] = y[1];
      row[3] = y[2];
    }
    C.tpos[4 * i] = x;
    C.tpos[4 * i + 1] = y[0];
    C.tpos[4 * i + 2] = y[1];
    C.tpos[4 * i + 3] = y[2];
    C.status[i] = code;
    C.naccpt[i] = acc;
  }
}

template <int M>
__global__ __launch_bounds__(256) void k_bounce_center(BounceArgs A, BCArgs C) {
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= A.lanes) return;
  for (int64_t i = lane; i < C.n; i += A.lanes) {
    DriftRHS<M> f{A, lane, C.Bm[i], C.coef[i], C.dstep, BST_OK, 0};
    bc_run(f, C, i);
  }
}

// Per-centre constants of the equatorial drift: mu, mass, charge, gamma
struct EqArgs {
  const double* k;  // [n][4]
};

template <int M>
__global__ __launch_bounds__(64) void k_bounce_center_eq(FieldC F, BCArgs C, EqArgs E) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C.n) return;
  EqDriftRHS<M> f{F, E.k[4 * i], E.k[4 * i + 1], E.k[4 * i + 2], E.k[4 * i + 3], BST_OK, 0};
  bc_run(f, C, i);
}

}  // namespace rapt
