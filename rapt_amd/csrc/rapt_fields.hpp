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
#include <hip/hip_runtime.h>
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
  F_CURRENT_SHEET = 7, F_DIPOLE_UNIFORM = 8
};

// Field constants: passed by value as a kernel argument, so every lane reads
// them from scalar registers (they are uniform across the launch).
struct FieldC {
  int32_t model;
  int32_t is_static;
  double gstep;
  double tstep;
  double p[16];
};

struct V3 {
  double x, y, z;
};

// 1/sqrt(x) and 1/x from the hardware estimates (v_rsq_f64, v_rcp_f64) with
// two Newton steps each: full double precision to within an ulp or two, in 9
// and 5 dependent operations instead of the IEEE sqrt and divide sequences.
// Arguments are positive and normal here (squared lengths, field magnitudes).
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = __builtin_fma(-(x * y), y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  e = __builtin_fma(-(x * y), y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
// |r|^-5 from r^2
__device__ __forceinline__ double inv_r5(double r2) {
  const double rs = rsqrt_nr(r2);
  const double i2 = rs * rs;
  return i2 * i2 * rs;
}

__device__ __forceinline__ double dot3(const V3& a, const V3& b) {
  // same association as numpy's 3-element dot: fma(a2,b2, fma(a1,b1, a0*b0))
  return __builtin_fma(a.z, b.z, __builtin_fma(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ V3 cross3(const V3& a, const V3& b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---- B(t, r) ---------------------------------------------------------------
template <int M>
__device__ __forceinline__ V3 fieldB(const FieldC& F, double t, double x, double y, double z) {
  if constexpr (M == F_EARTH_DIPOLE || M == F_DIPOLE_UNIFORM) {
    const double r2 = x * x + y * y + z * z;
    const double s = F.p[0] * inv_r5(r2);
    V3 B{s * (x * z), s * (y * z), s * (z * z - r2 * (1.0 / 3.0))};
    if constexpr (M == F_DIPOLE_UNIFORM) {
      B.x += F.p[1];
      B.y += F.p[2];
      B.z += F.p[3];
    }
    return B;
  } else if constexpr (M == F_DOUBLE_DIPOLE) {
    const double yz2 = y * y + z * z;
    const double r2a = x * x + yz2;
    const double ia = inv_r5(r2a);
    const double xb = x - F.p[1];
    const double r2b = xb * xb + yz2;
    const double ib = F.p[2] * inv_r5(r2b);
    const double c = F.p[0];
    const double zz2 = 2.0 * z * z;
    return V3{c * ((3.0 * x * z) * ia + (3.0 * xb * z) * ib),
              c * ((3.0 * y * z) * ia + (3.0 * y * z) * ib),
              c * ((zz2 - x * x - y * y) * ia + (zz2 - xb * xb - y * y) * ib)};
  } else if constexpr (M == F_UNIFORM_BZ || M == F_UNIFORM_CROSSED_EB) {
    return V3{0.0, 0.0, F.p[0]};
  } else if constexpr (M == F_VAR_EARTH_DIPOLE) {
    const double s = F.p[0] * (1.0 + F.p[1] * sin(TWO_PI * t / F.p[2]));
    const double r2 = x * x + y * y + z * z;
    const double iv = s * inv_r5(r2);
    return V3{(3.0 * x * z) * iv, (3.0 * y * z) * iv, (2.0 * z * z - x * x - y * y) * iv};
  } else if constexpr (M == F_PARABOLIC) {
    double bx;
    if (__builtin_fabs(z) <= 1.0) {
      bx = F.p[0] * z / F.p[2];
    } else {
      bx = (z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0)) * F.p[3];
    }
    return V3{bx, 0.0, F.p[1]};
  } else if constexpr (M == F_CURRENT_SHEET) {
    return V3{F.p[0] * tanh(z / F.p[1]), 0.0, F.p[2]};
  } else {
    return V3{0.0, 0.0, 0.0};
  }
}

template <int M>
__device__ __forceinline__ V3 fieldE(const FieldC& F) {
  if constexpr (M == F_UNIFORM_CROSSED_EB) {
    return V3{0.0, F.p[1], 0.0};
  } else if constexpr (M == F_CURRENT_SHEET) {
    return V3{F.p[3], F.p[4], F.p[5]};
  } else {
    return V3{0.0, 0.0, 0.0};
  }
}

template <int M>
constexpr bool has_E() {
  return M == F_UNIFORM_CROSSED_EB || M == F_CURRENT_SHEET;
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
  const double i2d = 1.0 / (2.0 * d);
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
  const double s = 0.5 / d;
  return V3{(b2.x - b1.x) * s, (b2.y - b1.y) * s, (b2.z - b1.z) * s};
}

// _Field.jacobianB (rapt/fields.py:133-153): J[i][j] = dB_i/dx_j
template <int M>
__device__ __forceinline__ double jacobian_maxabs(const FieldC& F, double t, double x, double y,
                                                  double z) {
  const double d = F.gstep;
  const double i2d = 1.0 / (2.0 * d);
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
    const V3 E = fieldE<M>(F);
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

// Guiding-centre equations (rapt/GuidingCenter.py:329-395).  Y = (x, y, z, ppar).
template <int M>
__device__ __forceinline__ void rhs_gc(const FieldC& F, int eom, double charge, double mass,
                                       double mu, double v, bool eq, double t, const double* Y,
                                       double* out) {
  const double ppar = Y[3];
  const V3 B = fieldB<M>(F, t, Y[0], Y[1], Y[2]);
  const double bb = dot3(B, B);
  const double ibm = rsqrt_nr(bb);
  const double Bmag = bb * ibm;
  const V3 ub{B.x * ibm, B.y * ibm, B.z * ibm};
  V3 gB, cb;
  if (eom == EOM_NT) {
    // Northrop-Teller needs grad|B| only; the shared stencil costs the same
    grad_curl<M>(F, t, Y[0], Y[1], Y[2], gB, cb);
    const double gamma = 1.0 / __builtin_sqrt(1.0 - (v / C_LIGHT) * (v / C_LIGHT));
    const double gm = gamma * mass;
    const double coef = (gm * (v * v) + (ppar * ppar) / gm) / (2.0 * charge * (Bmag * Bmag));
    const V3 c = cross3(ub, gB);
    const double pg = ppar / gm;
    out[0] = coef * c.x + pg * ub.x;
    out[1] = coef * c.y + pg * ub.y;
    out[2] = coef * c.z + pg * ub.z;
    out[3] = -mu * dot3(ub, gB) / gamma;
  } else {
    grad_curl<M>(F, t, Y[0], Y[1], Y[2], gB, cb);
    const double iq = rcp_nr(charge);
    const V3 Bs{B.x + ppar * cb.x * iq, B.y + ppar * cb.y * iq, B.z + ppar * cb.z * iq};
    const double Bspar = dot3(Bs, ub);
    const double iBs = rcp_nr(Bspar);
    if (eom == EOM_BC) {
      const double gamma = 1.0 / __builtin_sqrt(1.0 - (v / C_LIGHT) * (v / C_LIGHT));
      const double a = ppar / (gamma * mass);
      const double bcoef = mu / (charge * gamma);
      const V3 c = cross3(ub, gB);
      out[0] = (a * Bs.x + bcoef * c.x) * iBs;
      out[1] = (a * Bs.y + bcoef * c.y) * iBs;
      out[2] = (a * Bs.z + bcoef * c.z) * iBs;
      out[3] = -mu * dot3(Bs, gB) / (gamma * Bspar);
    } else {
      const double imc = rcp_nr(mass * C_LIGHT);
      const double pr = ppar * imc;
      const double gamma = __builtin_sqrt(1.0 + 2.0 * mu * Bmag * (imc * (1.0 / C_LIGHT)) + pr * pr);
      const double igam = rcp_nr(gamma);
      const double mg = mu * igam;
      V3 Es;
      if constexpr (has_E<M>()) {
        const V3 E = fieldE<M>(F);
        Es = V3{E.x - mg * gB.x * iq, E.y - mg * gB.y * iq, E.z - mg * gB.z * iq};
      } else {
        Es = V3{-mg * gB.x * iq, -mg * gB.y * iq, -mg * gB.z * iq};
      }
      if (!F.is_static) {
        const V3 db = dbdt<M>(F, t, Y[0], Y[1], Y[2]);
        Es.x -= ppar * db.x * iq;
        Es.y -= ppar * db.y * iq;
        Es.z -= ppar * db.z * iq;
      }
      const double a = pr * C_LIGHT * igam;
      const V3 c = cross3(Es, ub);
      out[0] = (a * Bs.x + c.x) * iBs;
      out[1] = (a * Bs.y + c.y) * iBs;
      out[2] = (a * Bs.z + c.z) * iBs;
      out[3] = charge * dot3(Es, Bs) * iBs;
    }
  }
  if (eq) {
    out[2] = 0.0;
    out[3] = 0.0;
  }
}

}  // namespace rapt
