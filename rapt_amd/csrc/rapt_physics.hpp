// rapt_physics.hpp -- adiabaticity tests and Particle <-> GuidingCenter
// conversions on the device, shared by the tracer kernels.
#pragma once
#include "rapt_fields.hpp"
#include "rapt_kargs.hpp"

namespace rapt {

// ---- adiabaticity tests ------------------------------------------------------------
// Particle.isadiabatic (rapt/Particle.py:345-384) on the row (t, r, p)
template <int M>
__device__ __forceinline__ bool particle_isadiabatic(const FieldC& F, double epss, double epst,
                                                     double t, const double* y, double mass,
                                                     double charge) {
  const V3 p{y[3], y[4], y[5]};
  const double gm = __builtin_sqrt(mass * mass + dot3(p, p) / C_LIGHT_SQ);
  const V3 v{p.x / gm, p.y / gm, p.z / gm};
  const double vsq = dot3(v, v);
  const double gamma = 1.0 / __builtin_sqrt(1.0 - vsq / C_LIGHT_SQ);
  const V3 B = fieldB<M>(F, t, y[0], y[1], y[2]);
  const double Bmag = __builtin_sqrt(dot3(B, B));
  const double vpar = dot3(v, B) / Bmag;
  const double vperp = __builtin_sqrt(vsq - vpar * vpar);
  const double rc = gamma * mass * vperp / (__builtin_fabs(charge) * Bmag);
  bool ok = rc / lengthscale<M>(F, t, y[0], y[1], y[2]) < epss;
  if (!F.is_static && ok) {
    const double cp = TWO_PI * gamma * mass / Bmag / __builtin_fabs(charge);
    ok = cp / timescale<M>(F, t, y[0], y[1], y[2]) < epst;
  }
  return ok;
}

// GuidingCenter.isadiabatic (rapt/GuidingCenter.py:287-327); the cycper()
// term keeps the reference's unnormalised ppar**2 (rapt/GuidingCenter.py:535)
template <int M>
__device__ __forceinline__ bool gc_isadiabatic(const FieldC& F, double epss, double epst, double t,
                                               const double* y, double mass, double charge,
                                               double mu) {
  const double Bmag = magB<M>(F, t, y[0], y[1], y[2]);
  const double pp = y[3];
  const double mcc = mass * C_LIGHT * C_LIGHT;
  const double pr = pp / mass / C_LIGHT;
  const double gamma = __builtin_sqrt(1.0 + 2.0 * mu * Bmag / mcc + pr * pr);
  double vp, v;
  if (gamma - 1.0 < 1e-6) {
    vp = pp / mass;
    v = __builtin_sqrt(2.0 * mu * Bmag / mass + vp * vp);
  } else {
    vp = pp / mass / gamma;
    v = C_LIGHT * __builtin_sqrt(1.0 - 1.0 / (gamma * gamma));
  }
  const double g2 = 1.0 / __builtin_sqrt(1.0 - (v / C_LIGHT) * (v / C_LIGHT));
  const double vperp = __builtin_sqrt((v - vp) * (v + vp));
  const double rc = g2 * mass * vperp / (__builtin_fabs(charge) * Bmag);
  bool ok = rc / lengthscale<M>(F, t, y[0], y[1], y[2]) < epss;
  if (!F.is_static && ok) {
    const double g3 = __builtin_sqrt(1.0 + 2.0 * mu * Bmag / mcc + pp * pp);
    double v3;
    if (g3 - 1.0 < 1e-6) {
      const double vp3 = pp / mass;
      v3 = __builtin_sqrt(2.0 * mu * Bmag / mass + vp3 * vp3);
    } else {
      v3 = C_LIGHT * __builtin_sqrt(1.0 - 1.0 / (g3 * g3));
    }
    const double g4 = 1.0 / __builtin_sqrt(1.0 - (v3 / C_LIGHT) * (v3 / C_LIGHT));
    const double cp = TWO_PI * g4 * mass / Bmag / __builtin_fabs(charge);
    ok = cp / timescale<M>(F, t, y[0], y[1], y[2]) < epst;
  }
  return ok;
}

// ---- mode conversions (ADAPTIVE) -----------------------------------------------------
// utils.guidingcenter (rapt/utils.py:251-326) + GuidingCenter.init(Particle)
// (rapt/GuidingCenter.py:168-186, __init__ :123-133).  y: Particle state in,
// GC state out (x, y, z, ppar); returns false when the iteration fails.
template <int M>
__device__ __forceinline__ bool particle_to_gc(const FieldC& F, double t, double* y, double mass, double charge,
                               double& mu, double& vout) {
  const V3 p{y[3], y[4], y[5]};
  const double gm = __builtin_sqrt(mass * mass + dot3(p, p) / C_LIGHT_SQ);
  const V3 v{p.x / gm, p.y / gm, p.z / gm};
  const double vsq = dot3(v, v);
  const double gamma = 1.0 / __builtin_sqrt(1.0 - vsq / C_LIGHT_SQ);
  const V3 r{y[0], y[1], y[2]};
  auto gyro = [&](const V3& q) {
    const V3 B = fieldB<M>(F, t, q.x, q.y, q.z);
    const double s = gamma * mass / (charge * dot3(B, B));
    const V3 c = cross3(B, v);
    return V3{s * c.x, s * c.y, s * c.z};
  };
  V3 g = gyro(r);
  V3 gold{r.x - g.x, r.y - g.y, r.z - g.z};
  for (int it = 1; it <= 20; ++it) {
    g = gyro(gold);
    const V3 gc{r.x - g.x, r.y - g.y, r.z - g.z};
    const V3 dd{gc.x - gold.x, gc.y - gold.y, gc.z - gold.z};
    if (__builtin_sqrt(dot3(dd, dd)) / __builtin_sqrt(dot3(gc, gc)) < 1e-3) {
      const V3 B = fieldB<M>(F, t, gc.x, gc.y, gc.z);
      const double Bn = __builtin_sqrt(dot3(B, B));
      const double vp = dot3(v, B) / Bn;
      const double spd = __builtin_sqrt(vsq);
      const double g2 = 1.0 / __builtin_sqrt(1.0 - (spd / C_LIGHT) * (spd / C_LIGHT));
      const double ppar = mass * g2 * vp;
      const double vpar = ppar / (mass * g2);
      mu = g2 * g2 * mass * (spd - vpar) * (spd + vpar) / (2.0 * Bn);
      vout = spd;
      y[0] = gc.x;
      y[1] = gc.y;
      y[2] = gc.z;
      y[3] = ppar;
      y[4] = y[5] = 0.0;
      return true;
    }
    gold = gc;
  }
  return false;
}

// Particle.init(GuidingCenter) (rapt/Particle.py:149-164) with utils.GCtoFP
// (rapt/utils.py:376-433) at gyrophase 0; the field is read at t = 0 there,
// the new Particle's tcur.  y: GC state in, Particle state out.
template <int M>
__device__ __forceinline__ void gc_to_particle(const FieldC& F, double t, double* y, double mass, double charge,
                               double mu) {
  (void)t;
  const double Bm = magB<M>(F, t, y[0], y[1], y[2]);
  const double pp = y[3];
  const double pr = pp / mass / C_LIGHT;
  const double gammasq = 1.0 + 2.0 * mu * Bm / (mass * C_LIGHT * C_LIGHT) + pr * pr;
  double v;
  if (__builtin_sqrt(gammasq) - 1.0 < 1e-6) {
    const double q = pp / mass;
    v = __builtin_sqrt(2.0 * mu * Bm / mass + q * q);
  } else {
    v = C_LIGHT * __builtin_sqrt(1.0 - 1.0 / gammasq);
  }
  const double vpar = pp / mass / __builtin_sqrt(gammasq);
  const V3 B = fieldB<M>(F, 0.0, y[0], y[1], y[2]);
  const double Bn = __builtin_sqrt(dot3(B, B));
  const V3 b{B.x / Bn, B.y / Bn, B.z / Bn};
  const double pa = acos(vpar / v);
  const double g2 = 1.0 / __builtin_sqrt(1.0 - (v / C_LIGHT) * (v / C_LIGHT));
  const double rc = g2 * mass * __builtin_sqrt((v - vpar) * (v + vpar)) /
                    (__builtin_fabs(charge) * Bn);
  V3 u;
  if (B.x == 0.0) {
    u = V3{1.0, 0.0, 0.0};
  } else if (B.y == 0.0) {
    u = V3{0.0, 1.0, 0.0};
  } else if (B.z == 0.0) {
    u = V3{0.0, 0.0, 1.0};
  } else {
    const double cc = -1.0 * (B.x + B.y) / B.z;
    const double nn = __builtin_sqrt(2.0 + cc * cc);
    u = V3{1.0 / nn, 1.0 / nn, cc / nn};
  }
  const double un = __builtin_sqrt(dot3(u, u));
  u = V3{u.x / un, u.y / un, u.z / un};
  const V3 w = cross3(b, u);
  const double s = charge > 0.0 ? 1.0 : (charge < 0.0 ? -1.0 : 0.0);
  const double cpa = cos(pa), spa = sin(pa);
  const V3 vel{v * (cpa * b.x - s * spa * w.x), v * (cpa * b.y - s * spa * w.y),
               v * (cpa * b.z - s * spa * w.z)};
  const double gamma = 1.0 / __builtin_sqrt(1.0 - dot3(vel, vel) / C_LIGHT_SQ);
  const double mg = mass * gamma;
  y[0] = y[0] + rc * u.x;
  y[1] = y[1] + rc * u.y;
  y[2] = y[2] + rc * u.z;
  y[3] = mg * vel.x;
  y[4] = mg * vel.y;
  y[5] = mg * vel.z;
}

}  // namespace rapt
