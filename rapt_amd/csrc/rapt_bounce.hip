// rapt_bounce.hip -- half-bounce path length S_b on the device, the output
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
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_internal.hpp"
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
  int32_t pad;
  int64_t lanes; // scratch stride
  double* S;     // [2*cap][lanes] arc length
  double* Bv;    // [2*cap][lanes] |B|
  double* Cc;    // [2*cap][lanes] spline coefficients
  double* Wk;    // [2*cap][lanes] elimination workspace
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

template <int M>
__global__ __launch_bounds__(256) void k_halfbounce(BounceArgs A) {
  const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= A.lanes) return;
  const FieldC& F = A.F;
  const int32_t base = A.cap;  // array index of the start point
  for (int64_t i = lane; i < A.n; i += A.lanes) {
    const double t = A.tpos[4 * i];
    const double x0 = A.tpos[4 * i + 1], y0 = A.tpos[4 * i + 2], z0 = A.tpos[4 * i + 3];
    const double Bm = A.Bm[i];
    const double ds = 1.0 / curvature<M>(F, t, x0, y0, z0) / A.res;
    A.S[(int64_t)base * A.lanes + lane] = 0.0;
    A.Bv[(int64_t)base * A.lanes + lane] = magB<M>(F, t, x0, y0, z0);
    int32_t st = BST_OK;
    int32_t nf = 0, nb = 0;
    for (int side = 0; side < 2 && st == BST_OK; ++side) {
      const double sign = side == 0 ? 1.0 : -1.0;
      const double tol = side == 0 ? 1e-3 : 1e-4;
      const int32_t dir = side == 0 ? 1 : -1;
      int32_t* np = side == 0 ? &nf : &nb;
      double Y[4] = {0.0, x0, y0, z0};
      for (;;) {
        const int rc = line_chunk<M>(A, lane, t, sign, tol, ds, Y, np, base, dir);
        if (rc < 0) { st = BST_OVERFLOW; break; }
        if (rc == 0) break;
        if (magB<M>(F, t, Y[1], Y[2], Y[3]) > Bm) break;
      }
    }
    double sb = __builtin_nan("");
    if (st == BST_OK) {
      // curve = array indices [base-nb, base+nf]; keep B <= Bm plus one each side
      const int32_t lo = base - nb, hi = base + nf;
      int32_t first = -1, lastin = -1;
      for (int32_t k = lo; k <= hi; ++k) {
        if (A.Bv[(int64_t)k * A.lanes + lane] <= Bm) {
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
      if (nt == 3) {
        const double s12 = X(0) - X(1), s23 = X(1) - X(2), s13 = X(0) - X(2);
        const double B2s = 2 * (Yv(0) * s23 - Yv(1) * s13 + Yv(2) * s12) / (s12 * s13 * s23);
        sb = 3.141592653589793 * __builtin_sqrt(2 * Bm / B2s);
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
          const double mid = 0.5 * (sm1 + sm2), half = 0.5 * (sm2 - sm1);
          const int P = 32;
          const double pw = 3.141592653589793 / P;
          double acc = 0.0;
          for (int p = 0; p < P; ++p) {
            const double c0 = (p + 0.5) * pw;
            for (int q = 0; q < 16; ++q) {
              const double xg = (q < 8) ? -GL16_X[7 - q] : GL16_X[q - 8];
              const double wg = (q < 8) ? GL16_W[7 - q] : GL16_W[q - 8];
              const double th = c0 + 0.5 * pw * xg;
              const double s = mid - half * cos(th);
              const double val = 1.0 - sp.eval(s) / Bm;
              acc += wg * sin(th) / __builtin_sqrt(val);
            }
          }
          sb = acc * 0.5 * pw * half;
          if (!__builtin_isfinite(sb)) st = BST_NAN;
        }
      }
    }
    A.out[i] = sb;
    A.status[i] = st;
  }
}

}  // namespace rapt

using namespace rapt;

template <int M>
static void launch_hb(rapt_ctx* c, const BounceArgs& A) {
  const int B = 256;
  hipLaunchKernelGGL(k_halfbounce<M>, dim3((unsigned)((A.lanes + B - 1) / B)), dim3(B), 0,
                     c->stream, A);
}

extern "C" int rapt_halfbounce(rapt_ctx* c, int64_t n, const double* tpos, const double* Bm,
                               double fieldlineresolution, double* sb, int32_t* status) {
  if (!c || n < 0 || (n && (!tpos || !Bm || !sb || !status))) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!(fieldlineresolution > 0.0))
    return set_err(c, RAPT_E_ARG, "fieldlineresolution must be > 0");
  if (n == 0) return RAPT_OK;
  HIP_OK(c, hipSetDevice(c->device));
  BounceArgs A;
  memset(&A, 0, sizeof A);
  A.F = c->field;
  A.n = n;
  A.res = fieldlineresolution;
  A.cap = 2048;
  A.lanes = n < 16384 ? n : 16384;
  const size_t npts = (size_t)2 * A.cap * A.lanes;
  double *d_tpos = nullptr, *d_bm = nullptr, *d_out = nullptr, *scratch = nullptr;
  int32_t* d_st = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&d_tpos, sizeof(double) * 4 * n) != hipSuccess ||
      hipMalloc(&d_bm, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&d_out, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&d_st, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&scratch, sizeof(double) * 4 * npts) != hipSuccess) {
    rc = set_err(c, RAPT_E_OOM, "bounce scratch allocation failed");
  } else {
    A.tpos = d_tpos;
    A.Bm = d_bm;
    A.out = d_out;
    A.status = d_st;
    A.S = scratch;
    A.Bv = scratch + npts;
    A.Cc = scratch + 2 * npts;
    A.Wk = scratch + 3 * npts;
    if (hipMemcpyAsync(d_tpos, tpos, sizeof(double) * 4 * n, hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        hipMemcpyAsync(d_bm, Bm, sizeof(double) * n, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess) {
      rc = set_err(c, RAPT_E_HIP, "bounce input copy failed");
    } else {
      switch (c->field.model) {
        case 1: launch_hb<1>(c, A); break;
        case 2: launch_hb<2>(c, A); break;
        case 5: launch_hb<5>(c, A); break;
        case 8: launch_hb<8>(c, A); break;
        default:
          rc = set_err(c, RAPT_E_ARG, "bounce period needs a dipole-type field (model %d)",
                       c->field.model);
      }
      if (rc == RAPT_OK) {
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(sb, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipMemcpyAsync(status, d_st, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
          rc = set_err(c, RAPT_E_HIP, "bounce kernel failed");
      }
    }
  }
  (void)hipFree(d_tpos);
  (void)hipFree(d_bm);
  (void)hipFree(d_out);
  (void)hipFree(d_st);
  (void)hipFree(scratch);
  return rc;
}
