// rapt_fields.hpp -- device field models, finite-difference derivatives and
// right-hand sides of the tracer (gfx950, FP64 state).
//
// Field formulas follow rapt/fields.py; the finite-difference stencils follow
// _Field.gradB / curlb / jacobianB / dbdt (rapt/fields.py:111-245) with the
// reference's step sizes.  Differences from the reference's arithmetic are
// confined to rounding: divisions by a common factor become one reciprocal
// and products, and |r|^5 is r2*r2*sqrt(r2).  The guiding-centre RHS
// evaluates B once per stencil point and derives both |B| (for grad|B|) and
// b = B/|B| (for curl b) from it, where the reference evaluates the field
// twice at each of the six points.
#pragma once
#ifndef __HIPCC_RTC__  // (run-time compiled kernels get the runtime from the compiler)
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

// No FMA contraction anywhere after this header (the library is also built
// with -ffp-contract=off): finite differences must cancel exactly where the
// reference's do (a uniform field has zero curl), which fused a*b-c*d breaks,
// and the one-lane and four-lane guiding-centre kernels must round every
// expression the same way to give identical trajectories.  Products are
// fused only where written as fma (dot3, mirroring numpy's BLAS dot, and the
// sums of squares of the error norms).
#pragma clang fp contract(off)

namespace rapt {

constexpr double C_LIGHT = 299792458.0;
constexpr double C_LIGHT_SQ = 89875517873681764.0;  // c**2 as a rounded integer
constexpr double TWO_PI = 6.283185307179586;

enum FieldModel : int {
  F_ZERO = 0, F_EARTH_DIPOLE = 1, F_DOUBLE_DIPOLE = 2, F_UNIFORM_BZ = 3,
  F_UNIFORM_CROSSED_EB = 4, F_VAR_EARTH_DIPOLE = 5, F_PARABOLIC = 6,
  F_CURRENT_SHEET = 7, F_DIPOLE_UNIFORM = 8, F_GRID = 9, F_EXPR = 10
};

// A user-written field (a _Field subclass overriding B and/or E,
// rapt/fields.py:5-74) traced into register programs by the Python front end:
// instruction i is code[4i..4i+3] = (opcode, destination, operand a, operand
// b) over 32 double registers, 0-3 holding t, x, y, z; a program leaves its
// three components in registers 4-6.  B's program is the first nb
// instructions, E's the next ne (ne = 0: E = 0).  Opcodes: 0 const (a =
// constant index), 1 add, 2 sub, 3 mul, 4 div, 5 neg, 6 pow, 7 sqrt, 8 exp,
// 9 log, 10 sin, 11 cos, 12 tan, 13 tanh, 14 sinh, 15 cosh, 16 atan, 17 asin,
// 18 acos, 19 fabs, 20 atan2, 21 mov, 22 log10, 23 hypot, 24 maximum, 25 minimum,
// 26 sign (numpy's: NaN propagates, sign(0) = 0).
struct ExprC {
  const int32_t* code;
  const double* k;
  int32_t nb, ne;
};
constexpr int EXPR_NREG = 32;

// A gridded field (Grid, rapt/fields.py:513-814) in HBM: B and E sampled on
// a rectilinear grid (axes need not be uniform) at nt >= 1 times.  Values
// are [nt][nx][ny][nz][4] doubles (x, y, z, unused), so the two z-neighbours
// of a cell corner are one 64-byte run.  All snapshots stay resident; the
// reference's window of three snapshots is the special case nt <= 3.
struct GridC {
  const double* ax[4];  // axis coordinates: x, y, z, t (t unused when nt = 1)
  const double* B;      // point layout [nt][nx][ny][nz][4], or cell layout (brick)
  const double* E;      // same layout as B, or nullptr (E = 0)
  int32_t n[4];         // nx, ny, nz, nt
  int32_t uniform;      // bit k: axis k is uniformly spaced
  int32_t brick;        // 1: cell layout [nt][nx-1][ny-1][nz-1][32]: the 8 corners
                        // of a cell, corner (dx, dy, dz) at 3 (4 dx + 2 dy + dz)
  double a0[4], ida[4]; // uniform axes: first coordinate, 1 / spacing
  double lo[4], hi[4];  // first and last coordinate of each axis
};

// Field constants: passed by value as a kernel argument, so every lane reads
// them from scalar registers (they are uniform across the launch).
struct FieldC {
  int32_t model;
  int32_t is_static;
  double gstep;
  double tstep;
  double p[16];
  // derived on the host when the field is set (IEEE divisions, so the same
  // values the device would compute)
  double i2d;   // 1 / (2 gstep): central-difference scale in space
  double ihts;  // 0.5 / tstep: central-difference scale in time (0 for static fields)
  union {
    GridC grid;   // F_GRID only
    ExprC expr;   // F_EXPR only
  };
};

template <class T>
struct V3T {
  T x, y, z;
};
using V3 = V3T<double>;
using V3f = V3T<float>;

// 1/sqrt(x) and 1/x from the hardware estimates (v_rsq_f64, v_rcp_f64) with
// two Newton steps each: full double precision to within an ulp or two, in 8
// and 5 operations instead of the IEEE sqrt and divide sequences.  A step of
// rsqrt_nr is y + y (1/2 - (x/2) y y) with x/2 formed once: the same values as
// y + (y/2)(1 - x y y), since halving is exact.  Arguments are positive and
// normal here (squared lengths, field magnitudes).
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double hx = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  double e = __builtin_fma(-(hx * y), y, 0.5);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-(hx * y), y, 0.5);
  return __builtin_fma(y, e, y);
}
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
// single precision: v_rsq_f32 / v_rcp_f32 are accurate to about 1 ulp
__device__ __forceinline__ float rsqrt_nr(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float rcp_nr(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fmaT(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double absT(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float absT(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double sqrtT(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float sqrtT(float a) { return __builtin_sqrtf(a); }

// |r|^-5 from r^2
template <class T>
__device__ __forceinline__ T inv_r5(T r2) {
  const T rs = rsqrt_nr(r2);
  const T i2 = rs * rs;
  return i2 * i2 * rs;
}
// c |r|^-5.  In double, c * inv_r5(r2).  In float |r|^-5 underflows (4e7 m
// gives 1e-38), so the coefficient is folded in first: ((c rs) rs^2) rs^2.
template <class T>
__device__ __forceinline__ T coef_r5(T c, T r2) {
  if constexpr (sizeof(T) == sizeof(double)) {
    return c * inv_r5(r2);
  } else {
    const T rs = rsqrt_nr(r2);
    const T i2 = rs * rs;
    return ((c * rs) * i2) * i2;
  }
}

template <class T>
__device__ __forceinline__ T dot3(const V3T<T>& a, const V3T<T>& b) {
  // same association as numpy's 3-element dot: fma(a2,b2, fma(a1,b1, a0*b0))
  return fmaT(a.z, b.z, fmaT(a.y, b.y, a.x * b.x));
}
template <class T>
__device__ __forceinline__ V3T<T> cross3(const V3T<T>& a, const V3T<T>& b) {
  return V3T<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---- gridded fields ---------------------------------------------------------
// Cell of coordinate x on axis k: i in [0, n-2] with g[i] <= x <= g[i+1] and
// the fraction f = (x - g[i]) / (g[i+1] - g[i]), or -1 outside [g[0], g[n-1]]
// (scipy's RegularGridInterpolator, which the reference's Grid uses, raises
// ValueError there).  A uniform axis takes the cell and the fraction from
// its spacing alone, without reading the coordinates (one memory round trip
// less per evaluation; the fraction differs from the stored coordinates'
// by rounding only); others bisect the stored coordinates.
__device__ __forceinline__ int grid_cell(const GridC& G, int k, double x, double& f) {
  const int n = G.n[k];
  if (!(x >= G.lo[k] && x <= G.hi[k])) return -1;
  if (G.uniform & (1 << k)) {
    const double u = (x - G.a0[k]) * G.ida[k];
    int i = (int)u;
    i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
    f = u - (double)i;
    return i;
  }
  const double* g = G.ax[k];
  int i;
  {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (g[mid] <= x) lo = mid; else hi = mid;
    }
    i = lo > n - 2 ? n - 2 : lo;
  }
  f = (x - g[i]) / (g[i + 1] - g[i]);
  return i;
}

__device__ __forceinline__ double lerp(double a, double b, double f) { return a + f * (b - a); }

// Trilinear interpolation in one snapshot: c points at corner (i, j, k).
__device__ __forceinline__ V3 grid_trilinear(const double* c, int64_t sx, int64_t sy,
                                             double fx, double fy, double fz) {
  double e[2][3];
#pragma unroll
  for (int dx = 0; dx < 2; ++dx) {
    double q[2][3];
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      const double* p = c + dx * sx + dy * sy;
#pragma unroll
      for (int m = 0; m < 3; ++m) q[dy][m] = lerp(p[m], p[4 + m], fz);
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) e[dx][m] = lerp(q[0][m], q[1][m], fy);
  }
  return V3{lerp(e[0][0], e[1][0], fx), lerp(e[0][1], e[1][1], fx), lerp(e[0][2], e[1][2], fx)};
}

// The same interpolation from a cell of the cell layout (32 doubles, 256-byte
// aligned: two cache lines per evaluation), with the same arithmetic.
__device__ __forceinline__ V3 grid_trilinear_cell(const double* c, double fx, double fy,
                                                  double fz) {
  double e[2][3];
#pragma unroll
  for (int dx = 0; dx < 2; ++dx) {
    double q[2][3];
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      const double* p = c + 3 * (4 * dx + 2 * dy);
#pragma unroll
      for (int m = 0; m < 3; ++m) q[dy][m] = lerp(p[m], p[3 + m], fz);
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) e[dx][m] = lerp(q[0][m], q[1][m], fy);
  }
  return V3{lerp(e[0][0], e[1][0], fx), lerp(e[0][1], e[1][1], fx), lerp(e[0][2], e[1][2], fx)};
}

// Multilinear interpolation of V at (t, x, y, z) in either layout:
// trilinear in space, linear in time between the two bracketing snapshots
// (nt > 1).  NaN outside the grid.
__device__ __forceinline__ V3 grid_interp(const GridC& G, const double* V, double t, double x,
                                          double y, double z) {
  double fx, fy, fz, ft = 0.0;
  const int i = grid_cell(G, 0, x, fx), j = grid_cell(G, 1, y, fy), k = grid_cell(G, 2, z, fz);
  int it = 0;
  if (G.n[3] > 1) it = grid_cell(G, 3, t, ft);
  if ((i | j | k | it) < 0) {
    const double nan = __builtin_nan("");
    return V3{nan, nan, nan};
  }
  if (G.brick) {
    const int64_t cy = (int64_t)(G.n[2] - 1) * 32, cx = (int64_t)(G.n[1] - 1) * cy,
                  ct = (int64_t)(G.n[0] - 1) * cx;
    const double* c = V + it * ct + i * cx + j * cy + k * 32;
    const V3 a = grid_trilinear_cell(c, fx, fy, fz);
    if (G.n[3] == 1) return a;
    const V3 b = grid_trilinear_cell(c + ct, fx, fy, fz);
    return V3{lerp(a.x, b.x, ft), lerp(a.y, b.y, ft), lerp(a.z, b.z, ft)};
  }
  const int64_t sy = (int64_t)G.n[2] * 4, sx = (int64_t)G.n[1] * sy, st = (int64_t)G.n[0] * sx;
  const double* c = V + it * st + i * sx + j * sy + k * 4;
  const V3 a = grid_trilinear(c, sx, sy, fx, fy, fz);
  if (G.n[3] == 1) return a;
  const V3 b = grid_trilinear(c + st, sx, sy, fx, fy, fz);
  return V3{lerp(a.x, b.x, ft), lerp(a.y, b.y, ft), lerp(a.z, b.z, ft)};
}

// ---- user-written fields ------------------------------------------------------
// pow as a user field's program takes it.  The powers of squared distances
// that dipole-like fields are written with (r2**2.5, r2**1.5, r2**-0.5, ...)
// are products and one square root: a few instructions against a few
// hundred for the FP64 pow, within 2 ulp of it (x**2 and x**3 as products;
// x**2 is pow's own correctly rounded value).  Where pow and these forms
// differ beyond rounding -- a base of -0 or -inf -- no field is evaluated.
// Every other exponent calls pow.  The interpreter and the run-time compiled
// kernels both evaluate opcode 6 through this function.
__device__ __forceinline__ double expr_pow(double a, double b) {
  if (b == 2.0) return a * a;
  if (b == 2.5) return (a * a) * __builtin_sqrt(a);
  if (b == 1.5) return a * __builtin_sqrt(a);
  if (b == 0.5) return __builtin_sqrt(a);
  if (b == 3.0) return (a * a) * a;
  if (b == 1.0) return a;
  if (b == -0.5) return 1.0 / __builtin_sqrt(a);
  if (b == -1.0) return 1.0 / a;
  if (b == -1.5) return 1.0 / (a * __builtin_sqrt(a));
  if (b == -2.0) return 1.0 / (a * a);
  if (b == -2.5) return 1.0 / ((a * a) * __builtin_sqrt(a));
  if (b == -3.0) return 1.0 / ((a * a) * a);
  return pow(a, b);
}
#ifdef RAPT_JIT_FIELD
// Compiled at run time (the tracer kernels of one traced program): the
// source that includes this header defines jit_B and jit_E after it, as
// straight-line code of the program's instructions, each the same operation
// on the same values as expr_eval's case for its opcode; k is the program's
// constant table in device memory.
constexpr bool EXPR_COMPILED = true;
__device__ __forceinline__ V3 jit_B(const double* k, double t, double x, double y, double z);
__device__ __forceinline__ V3 jit_E(const double* k, double t, double x, double y, double z);
#else
constexpr bool EXPR_COMPILED = false;
// The library functions of the opcode set, out of line: inlined into the
// interpreter they gave it ~50 registers to save and restore on every call.
__device__ __noinline__ double expr_fn1(int op, double a) {
  switch (op) {
    case 8: return exp(a);
    case 9: return log(a);
    case 10: return sin(a);
    case 11: return cos(a);
    case 12: return tan(a);
    case 13: return tanh(a);
    case 14: return sinh(a);
    case 15: return cosh(a);
    case 16: return atan(a);
    case 17: return asin(a);
    case 18: return acos(a);
    default: return log10(a);  // 22
  }
}
__device__ __noinline__ double expr_fn2(int op, double a, double b) {
  switch (op) {
    case 6: return expr_pow(a, b);
    case 20: return atan2(a, b);
    default: return hypot(a, b);  // 23
  }
}
// a value every lane holds, as a scalar (64-bit)
template <class P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)u);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  return (P*)(((unsigned long long)hi << 32) | lo);
}
// One program on the registers (t, x, y, z, ...).  The program is a kernel
// argument, the same for every lane, so it is read through the scalar unit
// (constant address space): the dispatch is a scalar branch and the register
// indices are scalar offsets into the lane's scratch registers (256 bytes).
// Out of line (one copy per kernel), taking the program by value: a
// reference into the kernel's argument block would make the caller copy the
// whole block to scratch to pass its address.
__device__ __noinline__ V3 expr_eval(ExprC X, int lo, int n, double t, double x, double y,
                                     double z) {
  using CI = const __attribute__((address_space(4))) int32_t;
  using CD = const __attribute__((address_space(4))) double;
  CI* code = (CI*)uniform_ptr(X.code);
  CD* kc = (CD*)uniform_ptr(X.k);
  lo = __builtin_amdgcn_readfirstlane(lo);
  n = __builtin_amdgcn_readfirstlane(n);
  double r[EXPR_NREG];
  r[0] = t;
  r[1] = x;
  r[2] = y;
  r[3] = z;
  r[4] = r[5] = r[6] = 0.0;
  for (int i = lo; i < lo + n; ++i) {
    const int32_t op = __builtin_amdgcn_readfirstlane(code[4 * i]),
                  d = __builtin_amdgcn_readfirstlane(code[4 * i + 1]),
                  ia = __builtin_amdgcn_readfirstlane(code[4 * i + 2]),
                  ib = __builtin_amdgcn_readfirstlane(code[4 * i + 3]);
    const double a = r[ia & (EXPR_NREG - 1)], b = r[ib & (EXPR_NREG - 1)];
    double v;
    switch (op) {
      case 0: v = kc[ia]; break;
      case 1: v = a + b; break;
      case 2: v = a - b; break;
      case 3: v = a * b; break;
      case 4: v = a / b; break;
      case 5: v = -a; break;
      case 7: v = sqrt(a); break;
      case 19: v = fabs(a); break;
      case 21: v = a; break;
      case 24: v = (a > b || a != a) ? a : b; break;  // numpy maximum: a NaN propagates,
      case 25: v = (a < b || a != a) ? a : b; break;  // ties give b (numpy minimum alike)
      case 26: v = a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : (a == 0.0 ? 0.0 : a)); break;  // sign
      case 6:
      case 20:
      case 23: v = expr_fn2(op, a, b); break;
      case 8:
      case 9:
      case 10:
      case 11:
      case 12:
      case 13:
      case 14:
      case 15:
      case 16:
      case 17:
      case 18:
      case 22: v = expr_fn1(op, a); break;
      default: v = __builtin_nan(""); break;
    }
    r[d & (EXPR_NREG - 1)] = v;
  }
  return V3{r[4], r[5], r[6]};
}
#endif  // RAPT_JIT_FIELD

// ---- B(t, r) ---------------------------------------------------------------
// T = double is the tracer's arithmetic; T = float serves the FP32-field
// precision mode (constants rounded once to float).
template <int M, class T = double>
__device__ __forceinline__ V3T<T> fieldB(const FieldC& F, T t, T x, T y, T z) {
  using V = V3T<T>;
  if constexpr (M == F_EARTH_DIPOLE || M == F_DIPOLE_UNIFORM) {
    const T r2 = x * x + y * y + z * z;
    const T s = coef_r5(T(F.p[0]), r2);
    V B{s * (x * z), s * (y * z), s * (z * z - r2 * T(1.0 / 3.0))};
    if constexpr (M == F_DIPOLE_UNIFORM) {
      B.x += T(F.p[1]);
      B.y += T(F.p[2]);
      B.z += T(F.p[3]);
    }
    return B;
  } else if constexpr (M == F_DOUBLE_DIPOLE) {
    if constexpr (sizeof(T) == sizeof(double)) {
      // B = sum over the two dipoles of c_k |r_k|^-5 (3 x_k z, 3 y z, 3 z^2 - r_k^2)
      // (2 z^2 - x^2 - y^2 = 3 z^2 - r^2), with the coefficients folded into
      // the |r|^-5 factors and the sums in fused multiply-adds
      const double yz2 = __builtin_fma(y, y, z * z);
      const double r2a = __builtin_fma(x, x, yz2);
      const double xb = x - F.p[1];
      const double r2b = __builtin_fma(xb, xb, yz2);
      const double ca = F.p[0] * inv_r5(r2a);
      const double cb = (F.p[0] * F.p[2]) * inv_r5(r2b);
      const double z3 = 3.0 * z;
      const double zz3 = z3 * z;
      return V{__builtin_fma(x * z3, ca, (xb * z3) * cb), (y * z3) * (ca + cb),
               __builtin_fma(zz3 - r2a, ca, (zz3 - r2b) * cb)};
    } else {
      const T yz2 = y * y + z * z;
      const T r2a = x * x + yz2;
      const T xb = x - T(F.p[1]);
      const T r2b = xb * xb + yz2;
      const T zz2 = T(2.0) * z * z;
      const T ca = coef_r5(T(F.p[0]), r2a);
      const T cb = coef_r5(T(F.p[0]) * T(F.p[2]), r2b);
      return V{(T(3.0) * x * z) * ca + (T(3.0) * xb * z) * cb,
               (T(3.0) * y * z) * ca + (T(3.0) * y * z) * cb,
               (zz2 - x * x - y * y) * ca + (zz2 - xb * xb - y * y) * cb};
    }
  } else if constexpr (M == F_UNIFORM_BZ || M == F_UNIFORM_CROSSED_EB) {
    return V{T(0.0), T(0.0), T(F.p[0])};
  } else if constexpr (M == F_VAR_EARTH_DIPOLE) {
    const T s = T(F.p[0]) * (T(1.0) + T(F.p[1]) * sin(T(TWO_PI) * t / T(F.p[2])));
    const T r2 = x * x + y * y + z * z;
    const T iv = coef_r5(s, r2);
    return V{(T(3.0) * x * z) * iv, (T(3.0) * y * z) * iv,
             (T(2.0) * z * z - x * x - y * y) * iv};
  } else if constexpr (M == F_PARABOLIC) {
    T bx;
    if (absT(z) <= T(1.0)) {
      bx = T(F.p[0]) * z / T(F.p[2]);
    } else {
      bx = (z > T(0.0) ? T(1.0) : (z < T(0.0) ? T(-1.0) : T(0.0))) * T(F.p[3]);
    }
    return V{bx, T(0.0), T(F.p[1])};
  } else if constexpr (M == F_CURRENT_SHEET) {
    return V{T(F.p[0]) * tanh(z / T(F.p[1])), T(0.0), T(F.p[2])};
  } else if constexpr (M == F_GRID) {
    static_assert(sizeof(T) == sizeof(double), "gridded fields are FP64");
    return grid_interp(F.grid, F.grid.B, t, x, y, z);
  } else if constexpr (M == F_EXPR) {
    static_assert(sizeof(T) == sizeof(double), "user-written fields are FP64");
#ifdef RAPT_JIT_FIELD
    return jit_B(F.expr.k, t, x, y, z);
#else
    return expr_eval(F.expr, 0, F.expr.nb, t, x, y, z);
#endif
  } else {
    return V{T(0.0), T(0.0), T(0.0)};
  }
}

// E(t, r); the analytic models' E is uniform
template <int M, class T = double>
__device__ __forceinline__ V3T<T> fieldE(const FieldC& F, T t, T x, T y, T z) {
  if constexpr (M == F_GRID) {
    static_assert(sizeof(T) == sizeof(double), "gridded fields are FP64");
    if (!F.grid.E) return V3T<T>{T(0.0), T(0.0), T(0.0)};
    return grid_interp(F.grid, F.grid.E, t, x, y, z);
  } else if constexpr (M == F_EXPR) {
    static_assert(sizeof(T) == sizeof(double), "user-written fields are FP64");
    if (F.expr.ne == 0) return V3T<T>{T(0.0), T(0.0), T(0.0)};
#ifdef RAPT_JIT_FIELD
    return jit_E(F.expr.k, t, x, y, z);
#else
    return expr_eval(F.expr, F.expr.nb, F.expr.ne, t, x, y, z);
#endif
  } else if constexpr (M == F_UNIFORM_CROSSED_EB) {
    return V3T<T>{T(0.0), T(F.p[1]), T(0.0)};
  } else if constexpr (M == F_CURRENT_SHEET) {
    return V3T<T>{T(F.p[3]), T(F.p[4]), T(F.p[5])};
  } else {
    return V3T<T>{T(0.0), T(0.0), T(0.0)};
  }
}

template <int M>
constexpr bool has_E() {
  return M == F_UNIFORM_CROSSED_EB || M == F_CURRENT_SHEET || M == F_GRID || M == F_EXPR;
}

template <int M>
__device__ __forceinline__ V3 unitb(const FieldC& F, double t, double x, double y, double z) {
  V3 B = fieldB<M>(F, t, x, y, z);
  const double ib = rsqrt_nr(dot3(B, B));
  return V3{B.x * ib, B.y * ib, B.z * ib};
}

template <int M>
__device__ __forceinline__ double magB(const FieldC& F, double t, double x, double y, double z) {
  V3 B = fieldB<M>(F, t, x, y, z);
  return __builtin_sqrt(dot3(B, B));
}

// curl b from the differences of b at opposite stencil points:
// dxy = b_y(+x) - b_y(-x), and so on
__device__ __forceinline__ V3 curl_from_pairs(double dxy, double dxz, double dyx, double dyz,
                                              double dzx, double dzy, double i2d) {
  return V3{(dyz - dzy) * i2d, (dzx - dxz) * i2d, (dxy - dyx) * i2d};
}

// Stencil of _Field.gradB and _Field.curlb (rapt/fields.py:111-131, 176-200)
// from one field evaluation per point.  Points: +x, -x, +y, -y, +z, -z.
template <int M>
__device__ __forceinline__ void grad_curl(const FieldC& F, double t, double x, double y,
                                          double z, V3& gB, V3& cb) {
  const double d = F.gstep;
  const double i2d = F.i2d;
  double m[6];
  V3 u[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double s = (k & 1) ? -d : d;
    const double px = x + ((k >> 1) == 0 ? s : 0.0);
    const double py = y + ((k >> 1) == 1 ? s : 0.0);
    const double pz = z + ((k >> 1) == 2 ? s : 0.0);
    V3 B = fieldB<M>(F, t, px, py, pz);
    const double bb = dot3(B, B);
    const double ib = rsqrt_nr(bb);
    m[k] = bb * ib;
    u[k] = V3{B.x * ib, B.y * ib, B.z * ib};
  }
  gB = V3{(m[0] - m[1]) * i2d, (m[2] - m[3]) * i2d, (m[4] - m[5]) * i2d};
  // the _M1 rows (rapt/fields.py:33-36) as differences of opposite points;
  // the four-lane kernel forms the same differences in separate lanes
  cb = curl_from_pairs(u[0].y - u[1].y, u[0].z - u[1].z, u[2].x - u[3].x, u[2].z - u[3].z,
                       u[4].x - u[5].x, u[4].y - u[5].y, i2d);
}

// _Field.dbdt (rapt/fields.py:225-245); zero for static fields
template <int M>
__device__ __forceinline__ V3 dbdt(const FieldC& F, double t, double x, double y, double z) {
  const double d = F.tstep;
  V3 b1 = unitb<M>(F, t - d, x, y, z);
  V3 b2 = unitb<M>(F, t + d, x, y, z);
  const double s = F.ihts;
  return V3{(b2.x - b1.x) * s, (b2.y - b1.y) * s, (b2.z - b1.z) * s};
}

// _Field.jacobianB (rapt/fields.py:133-153): J[i][j] = dB_i/dx_j
template <int M>
__device__ __forceinline__ double jacobian_maxabs(const FieldC& F, double t, double x, double y,
                                                  double z) {
  const double d = F.gstep;
  const double i2d = F.i2d;
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    V3 Bp = fieldB<M>(F, t, x + (k == 0 ? d : 0.0), y + (k == 1 ? d : 0.0), z + (k == 2 ? d : 0.0));
    V3 Bm = fieldB<M>(F, t, x - (k == 0 ? d : 0.0), y - (k == 1 ? d : 0.0), z - (k == 2 ? d : 0.0));
    mx = fmax(mx, __builtin_fabs((Bp.x - Bm.x) * i2d));
    mx = fmax(mx, __builtin_fabs((Bp.y - Bm.y) * i2d));
    mx = fmax(mx, __builtin_fabs((Bp.z - Bm.z) * i2d));
  }
  return mx;
}

// _Field.lengthscale (rapt/fields.py:247-261)
template <int M>
__device__ __forceinline__ double lengthscale(const FieldC& F, double t, double x, double y,
                                              double z) {
  return magB<M>(F, t, x, y, z) / jacobian_maxabs<M>(F, t, x, y, z);
}

// _Field.curvature (rapt/fields.py:155-174) as written there: np.dot(gB, |B|)
// with a scalar |B| is an elementwise product, so the "perpendicular" part is
// gB_i (1 - B_i/|B|)
template <int M>
__device__ __forceinline__ double curvature(const FieldC& F, double t, double x, double y,
                                            double z) {
  const V3 B = fieldB<M>(F, t, x, y, z);
  const double Bm = __builtin_sqrt(dot3(B, B));
  V3 g, c;
  grad_curl<M>(F, t, x, y, z, g, c);
  const double s = Bm / (Bm * Bm);
  const V3 gp{g.x - (g.x * s) * B.x, g.y - (g.y * s) * B.y, g.z - (g.z * s) * B.z};
  return __builtin_sqrt(dot3(gp, gp)) / Bm;
}

// _Field.timescale (rapt/fields.py:263-280) for non-static fields
template <int M>
__device__ __forceinline__ double timescale(const FieldC& F, double t, double x, double y,
                                            double z) {
  const double d = F.tstep;
  const double B1 = magB<M>(F, t - d, x, y, z);
  const double B2 = magB<M>(F, t + d, x, y, z);
  return magB<M>(F, t, x, y, z) / __builtin_fabs((B2 - B1) / d / 2.0);
}

// ---- right-hand sides ---------------------------------------------------------
// Lorentz force, the eom closure of Particle.advance (rapt/Particle.py:284-298).
// gm is gamma*m: frozen at advance start for static fields, else recomputed.
template <int M>
__device__ __forceinline__ void rhs_lorentz(const FieldC& F, double charge, double mass,
                                            double& gm, bool eq, double t, const double* Y,
                                            double* out) {
  const V3 p{Y[3], Y[4], Y[5]};
  if (!F.is_static) gm = __builtin_sqrt(mass * mass + dot3(p, p) / C_LIGHT_SQ);
  const double ig = rcp_nr(gm);
  out[0] = p.x * ig;
  out[1] = p.y * ig;
  out[2] = p.z * ig;
  const V3 B = fieldB<M>(F, t, Y[0], Y[1], Y[2]);
  const V3 c = cross3(p, B);
  if constexpr (has_E<M>()) {
    const V3 E = fieldE<M>(F, t, Y[0], Y[1], Y[2]);
    out[3] = charge * (E.x + c.x * ig);
    out[4] = charge * (E.y + c.y * ig);
    out[5] = charge * (E.z + c.z * ig);
  } else {
    out[3] = charge * (c.x * ig);
    out[4] = charge * (c.y * ig);
    out[5] = charge * (c.z * ig);
  }
  if (eq) {
    out[2] = 0.0;
    out[5] = 0.0;
  }
}

enum Eom : int { EOM_TCB = 0, EOM_BC = 1, EOM_NT = 2 };

// Per-particle constants of the guiding-centre algebra, formed once per
// particle (or step) instead of once per right-hand side: 1/q, 1/(m c),
// 2 mu / (m c^2) (TCB) and the Lorentz factor of the speed v (BC and NT only).
struct GcK {
  double iq, imc, kmu, gamv;
};
__device__ __forceinline__ GcK gc_consts(int eom, double charge, double mass, double v,
                                         double mu) {
  GcK k{rcp_nr(charge), rcp_nr(mass * C_LIGHT), 0.0, 1.0};
  k.kmu = 2.0 * mu * (k.imc * (1.0 / C_LIGHT));
  if (eom != EOM_TCB) k.gamv = 1.0 / __builtin_sqrt(1.0 - (v / C_LIGHT) * (v / C_LIGHT));
  return k;
}

// Guiding-centre equations (rapt/GuidingCenter.py:329-395).  Y = (x, y, z, ppar).
// The algebra, given B, |B|, b = B/|B| (as mag_unit forms them), grad|B|,
// curl b and (non-static TCB only) db/dt at the guiding centre; always FP64.
// K = gc_consts(eom, charge, mass, v, mu).  The TCB Lorentz factor enters only
// as 1/gamma, formed as 1/sqrt(1 + 2 mu |B| / (m c^2) + (ppar / (m c))^2) by
// rsqrt_nr.  B*, E* and the TCB velocity are formed in fused multiply-adds
// with the per-particle factors ppar/q and mu/(gamma q) taken first: the
// same terms as the reference's, rounded fewer times.
__device__ __forceinline__ void gc_algebra_b(int eom, bool has_e, const V3& E, bool is_static,
                                             double charge, double mass, double mu, double v,
                                             const GcK& K, bool eq, const double* Y,
                                             const V3& B, double Bmag, const V3& ub,
                                             const V3& gB, const V3& cb, const V3& db,
                                             double* out) {
  const double ppar = Y[3];
  if (eom == EOM_NT) {
    const double gamma = K.gamv;
    const double gm = gamma * mass;
    const double coef = (gm * (v * v) + (ppar * ppar) / gm) / (2.0 * charge * (Bmag * Bmag));
    const V3 c = cross3(ub, gB);
    const double pg = ppar / gm;
    out[0] = coef * c.x + pg * ub.x;
    out[1] = coef * c.y + pg * ub.y;
    out[2] = coef * c.z + pg * ub.z;
    out[3] = -mu * dot3(ub, gB) / gamma;
  } else {
    const double iq = K.iq;
    const double piq = ppar * iq;
    const V3 Bs{__builtin_fma(piq, cb.x, B.x), __builtin_fma(piq, cb.y, B.y),
                __builtin_fma(piq, cb.z, B.z)};
    const double Bspar = dot3(Bs, ub);
    const double iBs = rcp_nr(Bspar);
    if (eom == EOM_BC) {
      const double gamma = K.gamv;
      const double a = ppar / (gamma * mass);
      const double bcoef = mu / (charge * gamma);
      const V3 c = cross3(ub, gB);
      out[0] = (a * Bs.x + bcoef * c.x) * iBs;
      out[1] = (a * Bs.y + bcoef * c.y) * iBs;
      out[2] = (a * Bs.z + bcoef * c.z) * iBs;
      out[3] = -mu * dot3(Bs, gB) / (gamma * Bspar);
    } else {
      const double pr = ppar * K.imc;
      const double igam = rsqrt_nr(__builtin_fma(pr, pr, __builtin_fma(K.kmu, Bmag, 1.0)));
      const double mgq = mu * igam * iq;
      V3 Es;
      if (has_e) {
        Es = V3{__builtin_fma(-mgq, gB.x, E.x), __builtin_fma(-mgq, gB.y, E.y),
                __builtin_fma(-mgq, gB.z, E.z)};
      } else {
        Es = V3{-mgq * gB.x, -mgq * gB.y, -mgq * gB.z};
      }
      if (!is_static) {
        Es.x = __builtin_fma(-piq, db.x, Es.x);
        Es.y = __builtin_fma(-piq, db.y, Es.y);
        Es.z = __builtin_fma(-piq, db.z, Es.z);
      }
      const double a = pr * C_LIGHT * igam;
      // E* x b
      const V3 c{__builtin_fma(Es.y, ub.z, -(Es.z * ub.y)), __builtin_fma(Es.z, ub.x, -(Es.x * ub.z)),
                 __builtin_fma(Es.x, ub.y, -(Es.y * ub.x))};
      out[0] = __builtin_fma(a, Bs.x, c.x) * iBs;
      out[1] = __builtin_fma(a, Bs.y, c.y) * iBs;
      out[2] = __builtin_fma(a, Bs.z, c.z) * iBs;
      out[3] = charge * dot3(Es, Bs) * iBs;
    }
  }
  if (eq) {
    out[2] = 0.0;
    out[3] = 0.0;
  }
}
// The same algebra given B alone.
__device__ __forceinline__ void gc_algebra(int eom, bool has_e, const V3& E, bool is_static,
                                           double charge, double mass, double mu, double v,
                                           const GcK& K, bool eq, const double* Y, const V3& B,
                                           const V3& gB, const V3& cb, const V3& db,
                                           double* out) {
  const double bb = dot3(B, B);
  const double ibm = rsqrt_nr(bb);
  const V3 ub{B.x * ibm, B.y * ibm, B.z * ibm};
  gc_algebra_b(eom, has_e, E, is_static, charge, mass, mu, v, K, eq, Y, B, bb * ibm, ub, gB,
               cb, db, out);
}

template <int M>
// stat: the field is static (F.is_static, or a compile-time true)
__device__ __forceinline__ void rhs_gc(const FieldC& F, int eom, bool stat, double charge,
                                       double mass, double mu, double v, const GcK& K, bool eq,
                                       double t, const double* Y, double* out) {
  const V3 B = fieldB<M>(F, t, Y[0], Y[1], Y[2]);
  V3 gB, cb, db{0.0, 0.0, 0.0};
  grad_curl<M>(F, t, Y[0], Y[1], Y[2], gB, cb);
  if (eom == EOM_TCB && !stat) db = dbdt<M>(F, t, Y[0], Y[1], Y[2]);
  const V3 Ec = fieldE<M>(F, t, Y[0], Y[1], Y[2]);
  gc_algebra(eom, has_E<M>(), Ec, stat, charge, mass, mu, v, K, eq, Y, B, gB, cb, db, out);
}

// ---- FP32-field precision mode ----------------------------------------------------
// Field evaluations and the finite-difference drift terms (grad|B|, curl b,
// db/dt) in float; the state, stage sums, error norm and the guiding-centre
// algebra stay in double (products like ppar^2 ~ 1e-44 underflow float).
// A float stencil cannot resolve the FP64 gradient step (6.4 m at 4e7 m is
// below float spacing), so it uses d = max(gradientstepsize, |r| 2^-10),
// near the optimum of truncation (~(d/L)^2) against rounding (~eps_f32 L/d),
// and divides by the distance the rounded float coordinates actually span.
template <int M>
__device__ __forceinline__ void gc_fields_f32(const FieldC& F, double t, double x, double y,
                                              double z, V3& B0, V3& gB, V3& cb, V3& db,
                                              bool need_db) {
  const float tf = (float)t, xf = (float)x, yf = (float)y, zf = (float)z;
  const V3f Bc = fieldB<M, float>(F, tf, xf, yf, zf);
  B0 = V3{Bc.x, Bc.y, Bc.z};
  const double d = fmax(F.gstep, __builtin_sqrt(x * x + y * y + z * z) * 0x1p-10);
  const double c3[3] = {x, y, z};
  float m[6];
  V3f u[6];
  double inv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float cp = (float)(c3[k] + d), cm = (float)(c3[k] - d);
    inv[k] = 1.0 / ((double)cp - (double)cm);
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      const float ck = sgn ? cm : cp;
      const V3f B = fieldB<M, float>(F, tf, k == 0 ? ck : xf, k == 1 ? ck : yf, k == 2 ? ck : zf);
      const float bb = dot3(B, B);
      const float ib = rsqrt_nr(bb);
      m[2 * k + sgn] = bb * ib;
      u[2 * k + sgn] = V3f{B.x * ib, B.y * ib, B.z * ib};
    }
  }
  gB = V3{(double)(m[0] - m[1]) * inv[0], (double)(m[2] - m[3]) * inv[1],
          (double)(m[4] - m[5]) * inv[2]};
  const double dby_dx = (double)(u[0].y - u[1].y) * inv[0];
  const double dbz_dx = (double)(u[0].z - u[1].z) * inv[0];
  const double dbx_dy = (double)(u[2].x - u[3].x) * inv[1];
  const double dbz_dy = (double)(u[2].z - u[3].z) * inv[1];
  const double dbx_dz = (double)(u[4].x - u[5].x) * inv[2];
  const double dby_dz = (double)(u[4].y - u[5].y) * inv[2];
  cb = V3{dbz_dy - dby_dz, dbx_dz - dbz_dx, dby_dx - dbx_dy};
  db = V3{0.0, 0.0, 0.0};
  if (need_db) {
    const float tp = (float)(t + F.tstep), tm = (float)(t - F.tstep);
    const double it = 1.0 / ((double)tp - (double)tm);
    const V3f b1 = fieldB<M, float>(F, tm, xf, yf, zf);
    const V3f b2 = fieldB<M, float>(F, tp, xf, yf, zf);
    const float i1 = rsqrt_nr(dot3(b1, b1)), i2 = rsqrt_nr(dot3(b2, b2));
    db = V3{(double)(b2.x * i2 - b1.x * i1) * it, (double)(b2.y * i2 - b1.y * i1) * it,
            (double)(b2.z * i2 - b1.z * i1) * it};
  }
}

template <int M>
__device__ __forceinline__ void rhs_gc_f32(const FieldC& F, int eom, bool stat, double charge,
                                           double mass, double mu, double v, const GcK& K,
                                           bool eq, double t, const double* Y, double* out) {
  V3 B, gB, cb, db;
  gc_fields_f32<M>(F, t, Y[0], Y[1], Y[2], B, gB, cb, db, eom == EOM_TCB && !stat);
  const V3 Ec = fieldE<M>(F, t, Y[0], Y[1], Y[2]);
  gc_algebra(eom, has_E<M>(), Ec, stat, charge, mass, mu, v, K, eq, Y, B, gB, cb, db, out);
}

// Lorentz force with the field and the force in float; gamma*m in double.
template <int M>
__device__ __forceinline__ void rhs_lorentz_f32(const FieldC& F, double charge, double mass,
                                                double& gm, bool eq, double t, const double* Y,
                                                double* out) {
  if (!F.is_static) {
    const V3 pd{Y[3], Y[4], Y[5]};
    gm = __builtin_sqrt(mass * mass + dot3(pd, pd) / C_LIGHT_SQ);
  }
  const float ig = (float)rcp_nr(gm);
  const V3f p{(float)Y[3], (float)Y[4], (float)Y[5]};
  out[0] = (double)(p.x * ig);
  out[1] = (double)(p.y * ig);
  out[2] = (double)(p.z * ig);
  const V3f B = fieldB<M, float>(F, (float)t, (float)Y[0], (float)Y[1], (float)Y[2]);
  const V3f c = cross3(p, B);
  const float q = (float)charge;
  if constexpr (has_E<M>()) {
    const V3f E = fieldE<M, float>(F, (float)t, (float)Y[0], (float)Y[1], (float)Y[2]);
    out[3] = (double)(q * (E.x + c.x * ig));
    out[4] = (double)(q * (E.y + c.y * ig));
    out[5] = (double)(q * (E.z + c.z * ig));
  } else {
    out[3] = (double)(q * (c.x * ig));
    out[4] = (double)(q * (c.y * ig));
    out[5] = (double)(q * (c.z * ig));
  }
  if (eq) {
    out[2] = 0.0;
    out[5] = 0.0;
  }
}

}  // namespace rapt
