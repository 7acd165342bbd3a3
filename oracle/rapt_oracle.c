/*
 * rapt_oracle.c -- CPU FP64 restatement of RAPT's tracing path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, and the CPU
 * baseline ("port") that bench.py times.  The product never links this file.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Contraction
 * must stay off: every a*b+c below is two roundings, as in numpy, except where
 * an explicit fma() mirrors numpy's BLAS dot (see dot3 / curl stencil).
 *
 * Citations are to the reference (mkozturk/rapt) unless stated otherwise.
 */
#include "rapt_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* rapt/__init__.py:5-10 */
#define C_LIGHT 299792458.0
/* c**2 evaluated as a Python int then converted to float */
#define C_LIGHT_SQ 89875517873681764.0
static const double PI = 3.141592653589793;

/* np.dot of 3-vectors (OpenBLAS ddot, measured): fma chain */
static inline double dot3(const double *a, const double *b) {
    return fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0]));
}
/* np.cross (numpy/_core/numeric.py): separate products, then subtract */
static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
/* x**2 in Python and numpy calls libm pow(x, 2.0) */
static inline double sq(double x) { return pow(x, 2.0); }

/* ======================================================================
 * Field models: rapt/fields.py
 * ==================================================================== */
/* A user-written field (a _Field subclass overriding B/E,
 * rapt/fields.py:5-74): the operations its Python code performs on the
 * scalars of tpos, in order, as the reference evaluates them (numpy float64
 * scalar arithmetic; math functions of the C library). */
static void expr_run(const orc_field *f, int lo, int n, const double *tp, double *out) {
    double r[32];
    r[0] = tp[0]; r[1] = tp[1]; r[2] = tp[2]; r[3] = tp[3];
    r[4] = r[5] = r[6] = 0.0;
    for (int i = lo; i < lo + n; ++i) {
        const int32_t *c = f->code + 4 * i;
        const double a = c[0] == 0 ? 0.0 : r[c[2] & 31], b = r[c[3] & 31];
        double v;
        switch (c[0]) {
        case 0: v = f->consts[c[2]]; break;
        case 1: v = a + b; break;
        case 2: v = a - b; break;
        case 3: v = a * b; break;
        case 4: v = a / b; break;
        case 5: v = -a; break;
        case 6: v = pow(a, b); break;
        case 7: v = sqrt(a); break;
        case 8: v = exp(a); break;
        case 9: v = log(a); break;
        case 10: v = sin(a); break;
        case 11: v = cos(a); break;
        case 12: v = tan(a); break;
        case 13: v = tanh(a); break;
        case 14: v = sinh(a); break;
        case 15: v = cosh(a); break;
        case 16: v = atan(a); break;
        case 17: v = asin(a); break;
        case 18: v = acos(a); break;
        case 19: v = fabs(a); break;
        case 20: v = atan2(a, b); break;
        case 21: v = a; break;
        case 22: v = log10(a); break;
        case 23: v = hypot(a, b); break;
        case 24: v = (a > b || a != a) ? a : b; break; /* numpy maximum: ties give b */
        case 25: v = (a < b || a != a) ? a : b; break; /* numpy minimum */
        case 26: v = a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : (a == 0.0 ? 0.0 : a)); break;
        default: v = NAN; break;
        }
        r[c[1] & 31] = v;
    }
    out[0] = r[4]; out[1] = r[5]; out[2] = r[6];
}

static void field_B(const orc_field *f, const double *tp, double *B) {
    const double t = tp[0];
    double x = tp[1], y = tp[2], z = tp[3];
    switch (f->model) {
    case ORC_FIELD_EARTH_DIPOLE:
    case ORC_FIELD_DIPOLE_UNIFORM: { /* rapt/fields.py:301-317 */
        double r2 = x * x + y * y + z * z;
        double s = f->p[0] / pow(r2, 2.5);
        B[0] = s * (x * z);
        B[1] = s * (y * z);
        B[2] = s * (z * z - r2 / 3.0);
        if (f->model == ORC_FIELD_DIPOLE_UNIFORM) {
            B[0] = B[0] + f->p[1];
            B[1] = B[1] + f->p[2];
            B[2] = B[2] + f->p[3];
        }
        return;
    }
    case ORC_FIELD_DOUBLE_DIPOLE: { /* rapt/fields.py:344-362 */
        double p1 = pow(x * x + y * y + z * z, 2.5);
        double b1x = 3.0 * x * z / p1, b1y = 3.0 * y * z / p1;
        double b1z = (2.0 * z * z - x * x - y * y) / p1;
        x = x - f->p[1];
        double p2 = pow(x * x + y * y + z * z, 2.5);
        double k = f->p[2];
        double b2x = k * (3.0 * x * z) / p2, b2y = k * (3.0 * y * z) / p2;
        double b2z = k * (2.0 * z * z - x * x - y * y) / p2;
        B[0] = f->p[0] * (b1x + b2x);
        B[1] = f->p[0] * (b1y + b2y);
        B[2] = f->p[0] * (b1z + b2z);
        return;
    }
    case ORC_FIELD_UNIFORM_BZ:
    case ORC_FIELD_UNIFORM_CROSSED_EB: /* rapt/fields.py:376-390 */
        B[0] = 0.0; B[1] = 0.0; B[2] = f->p[0];
        return;
    case ORC_FIELD_VAR_EARTH_DIPOLE: { /* rapt/fields.py:455-470 */
        double s = f->p[0] * (1.0 + f->p[1] * sin(2.0 * PI * t / f->p[2]));
        double pw = pow(x * x + y * y + z * z, 2.5);
        B[0] = s * (3.0 * x * z) / pw;
        B[1] = s * (3.0 * y * z) / pw;
        B[2] = s * (2.0 * z * z - x * x - y * y) / pw;
        return;
    }
    case ORC_FIELD_PARABOLIC: /* rapt/fields.py:506-511 (module B0 when |z|>1) */
        if (fabs(z) <= 1.0) {
            B[0] = f->p[0] * z / f->p[2];
        } else {
            double sg = (z > 0.0) ? 1.0 : ((z < 0.0) ? -1.0 : 0.0);
            B[0] = sg * f->p[3];
        }
        B[1] = 0.0; B[2] = f->p[1];
        return;
    case ORC_FIELD_CURRENT_SHEET: /* B0 tanh(z/lambda) x + Bn z */
        B[0] = f->p[0] * tanh(z / f->p[1]);
        B[1] = 0.0; B[2] = f->p[2];
        return;
    case ORC_FIELD_EXPR:
        expr_run(f, 0, f->nb, tp, B);
        return;
    default:
        B[0] = B[1] = B[2] = 0.0;
        return;
    }
}

static void field_E(const orc_field *f, const double *tp, double *E) {
    E[0] = E[1] = E[2] = 0.0;
    if (f->model == ORC_FIELD_EXPR) {
        if (f->ne > 0) expr_run(f, f->nb, f->ne, tp, E);
    } else if (f->model == ORC_FIELD_UNIFORM_CROSSED_EB) { /* rapt/fields.py:413-427 */
        E[1] = f->p[1];
    } else if (f->model == ORC_FIELD_CURRENT_SHEET) {
        E[0] = f->p[3]; E[1] = f->p[4]; E[2] = f->p[5];
    }
}

/* rapt/fields.py:76-109 */
static double field_magB(const orc_field *f, const double *tp) {
    double B[3];
    field_B(f, tp, B);
    return sqrt(dot3(B, B));
}
static void field_unitb(const orc_field *f, const double *tp, double *b) {
    double B[3];
    field_B(f, tp, B);
    double m = sqrt(dot3(B, B));
    b[0] = B[0] / m; b[1] = B[1] / m; b[2] = B[2] / m;
}
static inline void shift(const double *tp, int k, double d, double *o) {
    o[0] = tp[0]; o[1] = tp[1]; o[2] = tp[2]; o[3] = tp[3];
    o[k] = tp[k] + d;
}
static inline void shiftm(const double *tp, int k, double d, double *o) {
    o[0] = tp[0]; o[1] = tp[1]; o[2] = tp[2]; o[3] = tp[3];
    o[k] = tp[k] - d;
}
/* rapt/fields.py:111-131 */
static void field_gradB(const orc_field *f, const double *tp, double *g) {
    const double d = f->gstep;
    double a[4], b[4];
    for (int k = 1; k <= 3; ++k) {
        shift(tp, k, d, a);
        shiftm(tp, k, d, b);
        g[k - 1] = (field_magB(f, a) - field_magB(f, b)) / (2.0 * d);
    }
}
/* rapt/fields.py:133-153: J[i][j] = dB_i/dx_j */
static void field_jacobianB(const orc_field *f, const double *tp, double *J) {
    const double d = f->gstep;
    double a[4], b[4], Bp[3], Bm[3];
    for (int k = 1; k <= 3; ++k) {
        shift(tp, k, d, a);
        shiftm(tp, k, d, b);
        field_B(f, a, Bp);
        field_B(f, b, Bm);
        for (int i = 0; i < 3; ++i) J[i * 3 + (k - 1)] = (Bp[i] - Bm[i]) / (2.0 * d);
    }
}
/* rapt/fields.py:155-174 (np.dot(gB, B) with scalar B is an elementwise product) */
static double field_curvature(const orc_field *f, const double *tp) {
    double Bv[3], gB[3], gp[3];
    field_B(f, tp, Bv);
    double Bm = sqrt(dot3(Bv, Bv));
    field_gradB(f, tp, gB);
    double Bsq = sq(Bm);
    for (int i = 0; i < 3; ++i) gp[i] = gB[i] - (gB[i] * Bm / Bsq) * Bv[i];
    return sqrt(dot3(gp, gp)) / Bm;
}
/* rapt/fields.py:176-200 with the _M1 stencil (:33-36); numpy dgemv order */
static void field_curlb(const orc_field *f, const double *tp, double *cb) {
    const double d = f->gstep;
    double q[4], bt[18];
    shift(tp, 1, d, q);  field_unitb(f, q, bt + 0);
    shiftm(tp, 1, d, q); field_unitb(f, q, bt + 3);
    shift(tp, 2, d, q);  field_unitb(f, q, bt + 6);
    shiftm(tp, 2, d, q); field_unitb(f, q, bt + 9);
    shift(tp, 3, d, q);  field_unitb(f, q, bt + 12);
    shiftm(tp, 3, d, q); field_unitb(f, q, bt + 15);
    /* row 0: +b8 -b11 -b13 +b16 ; row 1: -b2 +b5 +b12 -b15 ; row 2: +b1 -b4 -b6 +b9 */
    double r0 = (bt[8] + (-bt[11] - bt[13])) + bt[16];
    double r1 = (-bt[2] + bt[12]) + (bt[5] - bt[15]);
    double r2 = (-bt[4] - bt[6]) + (bt[1] + bt[9]);
    cb[0] = r0 / (2.0 * d);
    cb[1] = r1 / (2.0 * d);
    cb[2] = r2 / (2.0 * d);
}
/* rapt/fields.py:202-223 */
static double field_dBdt(const orc_field *f, const double *tp) {
    if (f->is_static) return 0.0;
    const double d = f->tstep;
    double a[4] = {tp[0] - d, tp[1], tp[2], tp[3]};
    double b[4] = {tp[0] + d, tp[1], tp[2], tp[3]};
    double B1 = field_magB(f, a), B2 = field_magB(f, b);
    return (B2 - B1) / d / 2.0;
}
/* rapt/fields.py:225-245 */
static void field_dbdt(const orc_field *f, const double *tp, double *o) {
    if (f->is_static) { o[0] = o[1] = o[2] = 0.0; return; }
    const double d = f->tstep;
    double a[4] = {tp[0] - d, tp[1], tp[2], tp[3]};
    double b[4] = {tp[0] + d, tp[1], tp[2], tp[3]};
    double b1[3], b2[3];
    field_unitb(f, a, b1);
    field_unitb(f, b, b2);
    for (int i = 0; i < 3; ++i) o[i] = (b2[i] - b1[i]) / d / 2.0;
}
/* rapt/fields.py:247-261 */
static double field_lengthscale(const orc_field *f, const double *tp) {
    double J[9];
    field_jacobianB(f, tp, J);
    double mx = fabs(J[0]);
    for (int i = 1; i < 9; ++i)
        if (fabs(J[i]) > mx) mx = fabs(J[i]);
    return field_magB(f, tp) / mx;
}
/* rapt/fields.py:263-280 (None -> NaN) */
static double field_timescale(const orc_field *f, const double *tp) {
    if (f->is_static) return NAN;
    return field_magB(f, tp) / fabs(field_dBdt(f, tp));
}

int orc_field_eval(const orc_field *f, int what, const double tp[4], double *out) {
    switch (what) {
    case ORC_EVAL_B: field_B(f, tp, out); return 3;
    case ORC_EVAL_E: field_E(f, tp, out); return 3;
    case ORC_EVAL_UNITB: field_unitb(f, tp, out); return 3;
    case ORC_EVAL_MAGB: out[0] = field_magB(f, tp); return 1;
    case ORC_EVAL_GRADB: field_gradB(f, tp, out); return 3;
    case ORC_EVAL_JACOBIANB: field_jacobianB(f, tp, out); return 9;
    case ORC_EVAL_CURVATURE: out[0] = field_curvature(f, tp); return 1;
    case ORC_EVAL_CURLB: field_curlb(f, tp, out); return 3;
    case ORC_EVAL_DBDT_MAG: out[0] = field_dBdt(f, tp); return 1;
    case ORC_EVAL_DBDT_VEC: field_dbdt(f, tp, out); return 3;
    case ORC_EVAL_LENGTHSCALE: out[0] = field_lengthscale(f, tp); return 1;
    case ORC_EVAL_TIMESCALE: out[0] = field_timescale(f, tp); return 1;
    default: return -1;
    }
}

int orc_field_eval_batch(const orc_field *f, int what, int64_t n, const double *tpos,
                         double *out) {
    double tmp[9];
    int w = orc_field_eval(f, what, tpos, tmp);
    if (w < 0) return w;
    for (int64_t i = 0; i < n; ++i) orc_field_eval(f, what, tpos + 4 * i, out + w * i);
    return w;
}

/* ======================================================================
 * Right-hand sides
 * ==================================================================== */
/* Particle.advance eom, rapt/Particle.py:284-298.  *gm_io is the closure's
 * nonlocal gm: frozen for static fields, recomputed otherwise. */
void orc_rhs_lorentz(const orc_field *f, const orc_params *p, double mass, double charge,
                     double *gm_io, double t, const double *Y, double *out) {
    double tp[4] = {t, Y[0], Y[1], Y[2]};
    const double *mom = Y + 3;
    if (!f->is_static) *gm_io = sqrt(sq(mass) + dot3(mom, mom) / C_LIGHT_SQ);
    const double gm = *gm_io;
    out[0] = Y[3] / gm; out[1] = Y[4] / gm; out[2] = Y[5] / gm;
    double E[3], B[3], cr[3];
    field_E(f, tp, E);
    field_B(f, tp, B);
    cross3(mom, B, cr);
    for (int i = 0; i < 3; ++i) out[3 + i] = charge * (E[i] + cr[i] / gm);
    if (p->enforce_equatorial) out[2] = out[5] = 0.0;
}

/* rapt/GuidingCenter.py:329-395 */
void orc_rhs_gc(const orc_field *f, const orc_params *p, int eom, double mass, double charge,
                double mu, double v, double t, const double *Y, double *out) {
    double tp[4] = {t, Y[0], Y[1], Y[2]};
    const double ppar = Y[3];
    double B[3], ub[3], cb[3], Bs[3], gB[3], cr[3];
    if (eom == ORC_EOM_TCB) { /* :329-355 */
        field_B(f, tp, B);
        double Bmag = sqrt(dot3(B, B));
        for (int i = 0; i < 3; ++i) ub[i] = B[i] / Bmag;
        double gamma = sqrt(1.0 + 2.0 * mu * Bmag / (mass * C_LIGHT * C_LIGHT) +
                            sq(ppar / (mass * C_LIGHT)));
        field_curlb(f, tp, cb);
        for (int i = 0; i < 3; ++i) Bs[i] = B[i] + ppar * cb[i] / charge;
        double Bspar = dot3(Bs, ub);
        double E[3], db[3], Es[3];
        field_E(f, tp, E);
        field_dbdt(f, tp, db);
        field_gradB(f, tp, gB);
        for (int i = 0; i < 3; ++i) Es[i] = E[i] - (ppar * db[i] + mu * gB[i] / gamma) / charge;
        cross3(Es, ub, cr);
        for (int i = 0; i < 3; ++i) out[i] = (ppar * Bs[i] / (gamma * mass) + cr[i]) / Bspar;
        out[3] = charge * dot3(Es, Bs) / Bspar;
    } else if (eom == ORC_EOM_BC) { /* :357-379 */
        double gamma = 1.0 / sqrt(1.0 - sq(v / C_LIGHT));
        field_B(f, tp, B);
        double Bmag = sqrt(dot3(B, B));
        for (int i = 0; i < 3; ++i) ub[i] = B[i] / Bmag;
        field_gradB(f, tp, gB);
        field_curlb(f, tp, cb);
        for (int i = 0; i < 3; ++i) Bs[i] = B[i] + ppar * cb[i] / charge;
        double Bspar = dot3(Bs, ub);
        cross3(ub, gB, cr);
        for (int i = 0; i < 3; ++i)
            out[i] = (ppar * Bs[i] / (gamma * mass) + mu * cr[i] / (charge * gamma)) / Bspar;
        out[3] = -mu * dot3(Bs, gB) / (gamma * Bspar);
    } else { /* Northrop-Teller :381-395 */
        double gamma = 1.0 / sqrt(1.0 - sq(v / C_LIGHT));
        double gm = gamma * mass;
        field_B(f, tp, B);
        double Bm = sqrt(dot3(B, B));
        for (int i = 0; i < 3; ++i) ub[i] = B[i] / Bm;
        field_gradB(f, tp, gB);
        cross3(ub, gB, cr);
        double coef = (gm * sq(v) + sq(ppar) / gm) / (2.0 * charge * sq(Bm));
        for (int i = 0; i < 3; ++i) out[i] = coef * cr[i] + ppar * ub[i] / gm;
        out[3] = -mu * dot3(ub, gB) / gamma;
    }
    if (p->enforce_equatorial) out[2] = out[3] = 0.0;
}

/* ======================================================================
 * DOPRI5 / DOP853 (Hairer & Wanner), one scipy integrate() call each.
 * ==================================================================== */
#define UROUND 2.3e-16
#define NMAXDIM 8

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* HINIT: initial step guess (HNW93 II.4, Hairer's dopri5/dop853 codes) */
static double hinit(orc_rhs_fn fn, void *ctx, int n, double x, const double *y, double posneg,
                    const double *f0, double *f1, double *y1, int iord, double hmax,
                    double atol, double rtol) {
    double dnf = 0.0, dny = 0.0;
    for (int i = 0; i < n; ++i) {
        double sk = atol + rtol * fabs(y[i]);
        dnf = dnf + (f0[i] / sk) * (f0[i] / sk);
        dny = dny + (y[i] / sk) * (y[i] / sk);
    }
    double h;
    if (dnf <= 1.0e-10 || dny <= 1.0e-10)
        h = 1.0e-6;
    else
        h = sqrt(dny / dnf) * 0.01;
    h = dmin(h, hmax);
    h = dsign(h, posneg);
    for (int i = 0; i < n; ++i) y1[i] = y[i] + h * f0[i];
    fn(ctx, x + h, y1, f1);
    double der2 = 0.0;
    for (int i = 0; i < n; ++i) {
        double sk = atol + rtol * fabs(y[i]);
        double q = (f1[i] - f0[i]) / sk;
        der2 = der2 + q * q;
    }
    der2 = sqrt(der2) / h;
    double der12 = dmax(fabs(der2), sqrt(dnf));
    double h1;
    if (der12 <= 1.0e-15)
        h1 = dmax(1.0e-6, fabs(h) * 1.0e-3);
    else
        h1 = pow(0.01 / der12, 1.0 / (double)iord);
    h = dmin(100.0 * fabs(h), dmin(h1, hmax));
    return dsign(h, posneg);
}

/* scipy dopri5 defaults map: beta == 0 means Hairer's default 0.04 */
static double dopri5_beta(double b) { return b == 0.0 ? 0.04 : (b < 0.0 ? 0.0 : b); }
static double dop853_beta(double b) { return b <= 0.0 ? 0.0 : b; }

int orc_dopri5_call(orc_rhs_fn fn, void *ctx, int n, double *x_io, double *y, double xend,
                    const orc_params *p, orc_ode *st) {
    static const double c2 = 0.2, c3 = 0.3, c4 = 0.8, c5 = 8.0 / 9.0;
    static const double a21 = 0.2, a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    static const double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
    static const double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0,
                        a53 = 64448.0 / 6561.0, a54 = -212.0 / 729.0;
    static const double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                        a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
    static const double a71 = 35.0 / 384.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0,
                        a75 = -2187.0 / 6784.0, a76 = 11.0 / 84.0;
    static const double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0,
                        e5 = -17253.0 / 339200.0, e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
    double k1[NMAXDIM], k2[NMAXDIM], k3[NMAXDIM], k4[NMAXDIM], k5[NMAXDIM], k6[NMAXDIM];
    double y1[NMAXDIM], ysti[NMAXDIM];
    double x = *x_io;
    const double safe = p->safety == 0.0 ? 0.9 : p->safety;
    const double fac1 = p->dfactor == 0.0 ? 0.2 : p->dfactor;
    const double fac2 = p->ifactor == 0.0 ? 10.0 : p->ifactor;
    const double beta = dopri5_beta(p->beta);
    const double atol = p->atol, rtol = p->rtol;
    const int nmax = p->nsteps > 0 ? p->nsteps : 500;
    double hmax = p->max_step == 0.0 ? xend - x : p->max_step;
    double h = st->h;
    double facold = 1.0e-4;
    const double expo1 = 0.2 - beta * 0.75;
    const double facc1 = 1.0 / fac1, facc2 = 1.0 / fac2;
    const double posneg = dsign(1.0, xend - x);
    int last = 0, reject = 0;
    int64_t nstep = 0, naccpt = 0, nrejct = 0, nfcn = 0;
    fn(ctx, x, y, k1);
    hmax = fabs(hmax);
    if (h == 0.0) h = hinit(fn, ctx, n, x, y, posneg, k1, k2, k3, 5, hmax, atol, rtol);
    nfcn += 2;
    int idid = 1;
    for (;;) {
        if (nstep > nmax) { idid = -2; break; }
        if (0.1 * fabs(h) <= fabs(x) * UROUND) { idid = -3; break; }
        if ((x + 1.01 * h - xend) * posneg > 0.0) {
            h = xend - x;
            last = 1;
        }
        nstep++;
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * a21 * k1[i];
        fn(ctx, x + c2 * h, y1, k2);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
        fn(ctx, x + c3 * h, y1, k3);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
        fn(ctx, x + c4 * h, y1, k4);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
        fn(ctx, x + c5 * h, y1, k5);
        for (int i = 0; i < n; ++i)
            ysti[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] +
                                  a65 * k5[i]);
        double xph = x + h;
        fn(ctx, xph, ysti, k6);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a71 * k1[i] + a73 * k3[i] + a74 * k4[i] + a75 * k5[i] +
                                a76 * k6[i]);
        fn(ctx, xph, y1, k2);
        for (int i = 0; i < n; ++i)
            k4[i] = (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] +
                     e7 * k2[i]) * h;
        nfcn += 6;
        double err = 0.0;
        for (int i = 0; i < n; ++i) {
            double sk = atol + rtol * dmax(fabs(y[i]), fabs(y1[i]));
            double q = k4[i] / sk;
            err = err + q * q;
        }
        err = sqrt(err / (double)n);
        double fac11 = pow(err, expo1);
        double fac = fac11 / pow(facold, beta);
        fac = dmax(facc2, dmin(facc1, fac / safe));
        double hnew = h / fac;
        if (err <= 1.0) {
            facold = dmax(err, 1.0e-4);
            naccpt++;
            for (int i = 0; i < n; ++i) { k1[i] = k2[i]; y[i] = y1[i]; }
            x = xph;
            if (last) { h = hnew; idid = 1; break; }
            if (fabs(hnew) > hmax) hnew = posneg * hmax;
            if (reject) hnew = posneg * dmin(fabs(hnew), fabs(h));
            reject = 0;
        } else {
            hnew = h / dmin(facc1, fac11 / safe);
            reject = 1;
            if (naccpt >= 1) nrejct++;
            last = 0;
        }
        h = hnew;
    }
    *x_io = x;
    st->h = h;
    st->istate = idid;
    st->cnt.naccpt += naccpt; st->cnt.nrejct += nrejct;
    st->cnt.nstep += nstep; st->cnt.nfcn += nfcn;
    return idid;
}

int orc_dop853_call(orc_rhs_fn fn, void *ctx, int n, double *x_io, double *y, double xend,
                    const orc_params *p, orc_ode *st) {
    /* Dormand & Prince 8(5,3) coefficients (HNW93; Hairer's DOP853) */
    static const double c2 = 0.526001519587677318785587544488e-01,
        c3 = 0.789002279381515978178381316732e-01, c4 = 0.118350341907227396726757197510e+00,
        c5 = 0.281649658092772603273242802490e+00, c6 = 0.333333333333333333333333333333e+00,
        c7 = 0.25e+00, c8 = 0.307692307692307692307692307692e+00,
        c9 = 0.651282051282051282051282051282e+00, c10 = 0.6e+00,
        c11 = 0.857142857142857142857142857142e+00;
    static const double a21 = 5.26001519587677318785587544488e-2,
        a31 = 1.97250569845378994544595329183e-2, a32 = 5.91751709536136983633785987549e-2,
        a41 = 2.95875854768068491816892993775e-2, a43 = 8.87627564304205475450678981324e-2,
        a51 = 2.41365134159266685502369798665e-1, a53 = -8.84549479328286085344864962717e-1,
        a54 = 9.24834003261792003115737966543e-1, a61 = 3.7037037037037037037037037037e-2,
        a64 = 1.70828608729473871279604482173e-1, a65 = 1.25467687566822425016691814123e-1,
        a71 = 3.7109375e-2, a74 = 1.70252211019544039314978060272e-1,
        a75 = 6.02165389804559606850219397283e-2, a76 = -1.7578125e-2,
        a81 = 3.70920001185047927108779319836e-2, a84 = 1.70383925712239993810214054705e-1,
        a85 = 1.07262030446373284651809199168e-1, a86 = -1.53194377486244017527936158236e-2,
        a87 = 8.27378916381402288758473766002e-3, a91 = 6.24110958716075717114429577812e-1,
        a94 = -3.36089262944694129406857109825e0, a95 = -8.68219346841726006818189891453e-1,
        a96 = 2.75920996994467083049415600797e1, a97 = 2.01540675504778934086186788979e1,
        a98 = -4.34898841810699588477366255144e1, a101 = 4.77662536438264365890433908527e-1,
        a104 = -2.48811461997166764192642586468e0, a105 = -5.90290826836842996371446475743e-1,
        a106 = 2.12300514481811942347288949897e1, a107 = 1.52792336328824235832596922938e1,
        a108 = -3.32882109689848629194453265587e1, a109 = -2.03312017085086261358222928593e-2,
        a111 = -9.3714243008598732571704021658e-1, a114 = 5.18637242884406370830023853209e0,
        a115 = 1.09143734899672957818500254654e0, a116 = -8.14978701074692612513997267357e0,
        a117 = -1.85200656599969598641566180701e1, a118 = 2.27394870993505042818970056734e1,
        a119 = 2.49360555267965238987089396762e0, a1110 = -3.0467644718982195003823669022e0,
        a121 = 2.27331014751653820792359768449e0, a124 = -1.05344954667372501984066689879e1,
        a125 = -2.00087205822486249909675718444e0, a126 = -1.79589318631187989172765950534e1,
        a127 = 2.79488845294199600508499808837e1, a128 = -2.85899827713502369474065508674e0,
        a129 = -8.87285693353062954433549289258e0, a1210 = 1.23605671757943030647266201528e1,
        a1211 = 6.43392746015763530355970484046e-1;
    static const double b1 = 5.42937341165687622380535766363e-2,
        b6 = 4.45031289275240888144113950566e0, b7 = 1.89151789931450038304281599044e0,
        b8 = -5.8012039600105847814672114227e0, b9 = 3.1116436695781989440891606237e-1,
        b10 = -1.52160949662516078556178806805e-1, b11 = 2.01365400804030348374776537501e-1,
        b12 = 4.47106157277725905176885569043e-2;
    static const double bhh1 = 0.244094488188976377952755905512e+00,
        bhh2 = 0.733846688281611857341361741547e+00, bhh3 = 0.220588235294117647058823529412e-01;
    static const double er1 = 0.1312004499419488073250102996e-01,
        er6 = -0.1225156446376204440720569753e+01, er7 = -0.4957589496572501915214079952e+00,
        er8 = 0.1664377182454986536961530415e+01, er9 = -0.3503288487499736816886487290e+00,
        er10 = 0.3341791187130174790297318841e+00, er11 = 0.8192320648511571246570742613e-01,
        er12 = -0.2235530786388629525884427845e-01;
    double k1[NMAXDIM], k2[NMAXDIM], k3[NMAXDIM], k4[NMAXDIM], k5[NMAXDIM], k6[NMAXDIM],
        k7[NMAXDIM], k8[NMAXDIM], k9[NMAXDIM], k10[NMAXDIM], y1[NMAXDIM];
    double x = *x_io;
    const double safe = p->safety == 0.0 ? 0.9 : p->safety;
    const double fac1 = p->dfactor == 0.0 ? 0.333 : p->dfactor;
    const double fac2 = p->ifactor == 0.0 ? 6.0 : p->ifactor;
    const double beta = dop853_beta(p->beta);
    const double atol = p->atol, rtol = p->rtol;
    const int nmax = p->nsteps > 0 ? p->nsteps : 500;
    double hmax = p->max_step == 0.0 ? xend - x : p->max_step;
    double h = st->h;
    double facold = 1.0e-4;
    const double expo1 = 1.0 / 8.0 - beta * 0.2;
    const double facc1 = 1.0 / fac1, facc2 = 1.0 / fac2;
    const double posneg = dsign(1.0, xend - x);
    int last = 0, reject = 0;
    int64_t nstep = 0, naccpt = 0, nrejct = 0, nfcn = 0;
    fn(ctx, x, y, k1);
    hmax = fabs(hmax);
    if (h == 0.0) h = hinit(fn, ctx, n, x, y, posneg, k1, k2, k3, 8, hmax, atol, rtol);
    nfcn += 2;
    int idid = 1;
    for (;;) {
        if (nstep > nmax) { idid = -2; break; }
        if (0.1 * fabs(h) <= fabs(x) * UROUND) { idid = -3; break; }
        if ((x + 1.01 * h - xend) * posneg > 0.0) {
            h = xend - x;
            last = 1;
        }
        nstep++;
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * a21 * k1[i];
        fn(ctx, x + c2 * h, y1, k2);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
        fn(ctx, x + c3 * h, y1, k3);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * (a41 * k1[i] + a43 * k3[i]);
        fn(ctx, x + c4 * h, y1, k4);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * (a51 * k1[i] + a53 * k3[i] + a54 * k4[i]);
        fn(ctx, x + c5 * h, y1, k5);
        for (int i = 0; i < n; ++i) y1[i] = y[i] + h * (a61 * k1[i] + a64 * k4[i] + a65 * k5[i]);
        fn(ctx, x + c6 * h, y1, k6);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a71 * k1[i] + a74 * k4[i] + a75 * k5[i] + a76 * k6[i]);
        fn(ctx, x + c7 * h, y1, k7);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a81 * k1[i] + a84 * k4[i] + a85 * k5[i] + a86 * k6[i] +
                                a87 * k7[i]);
        fn(ctx, x + c8 * h, y1, k8);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a91 * k1[i] + a94 * k4[i] + a95 * k5[i] + a96 * k6[i] +
                                a97 * k7[i] + a98 * k8[i]);
        fn(ctx, x + c9 * h, y1, k9);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a101 * k1[i] + a104 * k4[i] + a105 * k5[i] + a106 * k6[i] +
                                a107 * k7[i] + a108 * k8[i] + a109 * k9[i]);
        fn(ctx, x + c10 * h, y1, k10);
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a111 * k1[i] + a114 * k4[i] + a115 * k5[i] + a116 * k6[i] +
                                a117 * k7[i] + a118 * k8[i] + a119 * k9[i] + a1110 * k10[i]);
        fn(ctx, x + c11 * h, y1, k2);
        double xph = x + h;
        for (int i = 0; i < n; ++i)
            y1[i] = y[i] + h * (a121 * k1[i] + a124 * k4[i] + a125 * k5[i] + a126 * k6[i] +
                                a127 * k7[i] + a128 * k8[i] + a129 * k9[i] + a1210 * k10[i] +
                                a1211 * k2[i]);
        fn(ctx, xph, y1, k3);
        nfcn += 11;
        for (int i = 0; i < n; ++i) {
            k4[i] = b1 * k1[i] + b6 * k6[i] + b7 * k7[i] + b8 * k8[i] + b9 * k9[i] +
                    b10 * k10[i] + b11 * k2[i] + b12 * k3[i];
            k5[i] = y[i] + h * k4[i];
        }
        double err = 0.0, err2 = 0.0;
        for (int i = 0; i < n; ++i) {
            double sk = atol + rtol * dmax(fabs(y[i]), fabs(k5[i]));
            double erri = k4[i] - bhh1 * k1[i] - bhh2 * k9[i] - bhh3 * k3[i];
            double q = erri / sk;
            err2 = err2 + q * q;
            erri = er1 * k1[i] + er6 * k6[i] + er7 * k7[i] + er8 * k8[i] + er9 * k9[i] +
                   er10 * k10[i] + er11 * k2[i] + er12 * k3[i];
            q = erri / sk;
            err = err + q * q;
        }
        double deno = err + 0.01 * err2;
        if (deno <= 0.0) deno = 1.0;
        err = fabs(h) * err * sqrt(1.0 / ((double)n * deno));
        double fac11 = pow(err, expo1);
        double fac = fac11 / pow(facold, beta);
        fac = dmax(facc2, dmin(facc1, fac / safe));
        double hnew = h / fac;
        if (err <= 1.0) {
            facold = dmax(err, 1.0e-4);
            naccpt++;
            fn(ctx, xph, k5, k4);
            nfcn += 1;
            for (int i = 0; i < n; ++i) { k1[i] = k4[i]; y[i] = k5[i]; }
            x = xph;
            if (last) { h = hnew; idid = 1; break; }
            if (fabs(hnew) > hmax) hnew = posneg * hmax;
            if (reject) hnew = posneg * dmin(fabs(hnew), fabs(h));
            reject = 0;
        } else {
            hnew = h / dmin(facc1, fac11 / safe);
            reject = 1;
            if (naccpt >= 1) nrejct++;
            last = 0;
        }
        h = hnew;
    }
    *x_io = x;
    st->h = h;
    st->istate = idid;
    st->cnt.naccpt += naccpt; st->cnt.nrejct += nrejct;
    st->cnt.nstep += nstep; st->cnt.nfcn += nfcn;
    return idid;
}

/* rapt/rkf.py:13-143.  f(x, t) of the reference is fn(ctx, t, x, out). */
int orc_rkf(orc_rhs_fn fn, void *ctx, int n, double a, double b, const double *x0, double tol,
            double hmax, double hmin, double *tout, double *xout, int64_t cap, int64_t *npts) {
    static const double a2 = 2.500000000000000e-01, a3 = 3.750000000000000e-01,
                        a4 = 9.230769230769231e-01, a5 = 1.000000000000000e+00,
                        a6 = 5.000000000000000e-01;
    static const double b21 = 2.500000000000000e-01, b31 = 9.375000000000000e-02,
                        b32 = 2.812500000000000e-01, b41 = 8.793809740555303e-01,
                        b42 = -3.277196176604461e+00, b43 = 3.320892125625853e+00,
                        b51 = 2.032407407407407e+00, b52 = -8.000000000000000e+00,
                        b53 = 7.173489278752436e+00, b54 = -2.058966861598441e-01,
                        b61 = -2.962962962962963e-01, b62 = 2.000000000000000e+00,
                        b63 = -1.381676413255361e+00, b64 = 4.529727095516569e-01,
                        b65 = -2.750000000000000e-01;
    static const double r1 = 2.777777777777778e-03, r3 = -2.994152046783626e-02,
                        r4 = -2.919989367357789e-02, r5 = 2.000000000000000e-02,
                        r6 = 3.636363636363636e-02;
    static const double c1 = 1.157407407407407e-01, c3 = 5.489278752436647e-01,
                        c4 = 5.353313840155945e-01, c5 = -2.000000000000000e-01;
    double x[NMAXDIM], k1[NMAXDIM], k2[NMAXDIM], k3[NMAXDIM], k4[NMAXDIM], k5[NMAXDIM],
        k6[NMAXDIM], tmp[NMAXDIM];
    double t = a, h = hmax;
    int64_t m = 0;
    memcpy(x, x0, sizeof(double) * n);
    if (cap < 1) return ORC_ERR_CAP;
    tout[0] = t;
    memcpy(xout, x, sizeof(double) * n);
    m = 1;
    int rc = ORC_OK;
    while (t < b) {
        if (t + h > b) h = b - t;
        fn(ctx, t, x, k1);
        for (int i = 0; i < n; ++i) { k1[i] = h * k1[i]; tmp[i] = x[i] + b21 * k1[i]; }
        fn(ctx, t + a2 * h, tmp, k2);
        for (int i = 0; i < n; ++i) { k2[i] = h * k2[i]; tmp[i] = x[i] + b31 * k1[i] + b32 * k2[i]; }
        fn(ctx, t + a3 * h, tmp, k3);
        for (int i = 0; i < n; ++i) {
            k3[i] = h * k3[i];
            tmp[i] = x[i] + b41 * k1[i] + b42 * k2[i] + b43 * k3[i];
        }
        fn(ctx, t + a4 * h, tmp, k4);
        for (int i = 0; i < n; ++i) {
            k4[i] = h * k4[i];
            tmp[i] = x[i] + b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
        }
        fn(ctx, t + a5 * h, tmp, k5);
        for (int i = 0; i < n; ++i) {
            k5[i] = h * k5[i];
            tmp[i] = x[i] + b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
        }
        fn(ctx, t + a6 * h, tmp, k6);
        double r = 0.0;
        for (int i = 0; i < n; ++i) {
            k6[i] = h * k6[i];
            double ri = fabs(r1 * k1[i] + r3 * k3[i] + r4 * k4[i] + r5 * k5[i] + r6 * k6[i]) / h;
            if (i == 0 || ri > r) r = ri;
        }
        if (r <= tol) {
            t = t + h;
            for (int i = 0; i < n; ++i) x[i] = x[i] + c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
            if (m >= cap) { rc = ORC_ERR_CAP; break; }
            tout[m] = t;
            memcpy(xout + m * n, x, sizeof(double) * n);
            m++;
        }
        double q = pow(tol / r, 0.25);
        h = h * dmin(dmax(0.84 * q, 0.1), 4.0);
        if (h > hmax) {
            h = hmax;
        } else if (h < hmin || !(h > 0.0)) {
            rc = ORC_ERR_HMIN;
            break;
        }
    }
    *npts = m;
    return rc;
}

/* ======================================================================
 * Test problems for black-box comparison with scipy
 * ==================================================================== */
static void tp_harmonic(void *ctx, double t, const double *y, double *o) {
    (void)ctx; (void)t;
    o[0] = y[1];
    o[1] = -y[0];
}
/* Arenstorf orbit (HNW93 II.0, eq. 0.1) */
static void tp_arenstorf(void *ctx, double t, const double *y, double *o) {
    (void)ctx; (void)t;
    const double mu = 0.012277471, mup = 1.0 - mu;
    double d1 = pow((y[0] + mu) * (y[0] + mu) + y[1] * y[1], 1.5);
    double d2 = pow((y[0] - mup) * (y[0] - mup) + y[1] * y[1], 1.5);
    o[0] = y[2];
    o[1] = y[3];
    o[2] = y[0] + 2.0 * y[3] - mup * (y[0] + mu) / d1 - mu * (y[0] - mup) / d2;
    o[3] = y[1] - 2.0 * y[2] - mup * y[1] / d1 - mu * y[1] / d2;
}

int orc_test_problem_run(int problem, int stepper, const orc_params *p, int n, double t0,
                         const double *y0, double dt, int64_t ncalls, double *rows,
                         orc_counters *cnt, int32_t *istate) {
    orc_rhs_fn fn = problem == 0 ? tp_harmonic : tp_arenstorf;
    double y[NMAXDIM];
    memcpy(y, y0, sizeof(double) * n);
    double x = t0;
    orc_ode st;
    memset(&st, 0, sizeof st);
    st.h = p->first_step;
    int64_t k = 0;
    int rc = 1;
    for (; k < ncalls; ++k) {
        double xend = x + dt;
        if (stepper == ORC_STEP_DOP853)
            rc = orc_dop853_call(fn, NULL, n, &x, y, xend, p, &st);
        else
            rc = orc_dopri5_call(fn, NULL, n, &x, y, xend, p, &st);
        rows[k * (n + 1)] = x;
        memcpy(rows + k * (n + 1) + 1, y, sizeof(double) * n);
        if (rc < 0) { k++; break; }
    }
    if (cnt) *cnt = st.cnt;
    if (istate) *istate = rc;
    return (int)k;
}

/* ======================================================================
 * Physics helpers, rapt/utils.py
 * ==================================================================== */
/* :29-66 */
double orc_cyclotron_period(const orc_field *f, double t, const double pos[3],
                            const double vel[3], double mass, double charge) {
    double gamma = 1.0 / sqrt(1.0 - dot3(vel, vel) / C_LIGHT_SQ);
    double tp[4] = {t, pos[0], pos[1], pos[2]};
    double B = field_magB(f, tp);
    return 2.0 * PI * gamma * mass / B / fabs(charge);
}
/* :68-104 */
double orc_cyclotron_period2(const orc_field *f, double t, const double pos[3], double speed,
                             double mass, double charge) {
    double gamma = 1.0 / sqrt(1.0 - sq(speed / C_LIGHT));
    double tp[4] = {t, pos[0], pos[1], pos[2]};
    double B = field_magB(f, tp);
    return 2.0 * PI * gamma * mass / B / fabs(charge);
}
/* :106-145 */
double orc_cyclotron_radius(const orc_field *f, double t, const double pos[3],
                            const double vel[3], double mass, double charge) {
    double vsq = dot3(vel, vel);
    double gamma = 1.0 / sqrt(1.0 - vsq / C_LIGHT_SQ);
    double tp[4] = {t, pos[0], pos[1], pos[2]};
    double B[3];
    field_B(f, tp, B);
    double Bmag = sqrt(dot3(B, B));
    double vpar = dot3(vel, B) / Bmag;
    double vperp = sqrt(vsq - sq(vpar));
    return gamma * mass * vperp / (fabs(charge) * Bmag);
}
/* :147-187 */
double orc_cyclotron_radius2(const orc_field *f, double t, const double pos[3], double vpar,
                             double v, double mass, double charge) {
    double gamma = 1.0 / sqrt(1.0 - sq(v / C_LIGHT));
    double tp[4] = {t, pos[0], pos[1], pos[2]};
    double B[3];
    field_B(f, tp, B);
    double Bmag = sqrt(dot3(B, B));
    double vperp = sqrt((v - vpar) * (v + vpar));
    return gamma * mass * vperp / (fabs(charge) * Bmag);
}
/* :189-216 */
double orc_magnetic_moment(const orc_field *f, double t, const double pos[3], double vpar,
                           double v, double mass) {
    double gamma = 1.0 / sqrt(1.0 - sq(v / C_LIGHT));
    double tp[4] = {t, pos[0], pos[1], pos[2]};
    double Bmag = field_magB(f, tp);
    return sq(gamma) * mass * (v - vpar) * (v + vpar) / (2.0 * Bmag);
}
/* :251-326 gyrovector + fixed point R = r - rho(R) */
static void gyrovector(const orc_field *f, double t, const double *r, const double *v,
                       double mass, double charge, double *o) {
    double vsq = dot3(v, v);
    double gamma = 1.0 / sqrt(1.0 - vsq / C_LIGHT_SQ);
    double tp[4] = {t, r[0], r[1], r[2]};
    double B[3], cr[3];
    field_B(f, tp, B);
    double Bsq = dot3(B, B);
    cross3(B, v, cr);
    double s = gamma * mass / (charge * Bsq);
    o[0] = s * cr[0]; o[1] = s * cr[1]; o[2] = s * cr[2];
}
static inline double norm3(const double *v) { return sqrt(dot3(v, v)); }

int orc_guidingcenter(const orc_field *f, double t, const double r[3], const double v[3],
                      double mass, double charge, double tol, int maxiter, double gc_out[3],
                      double *vp, double *spd) {
    double gold[3], gc[3], g[3], dd[3];
    gyrovector(f, t, r, v, mass, charge, g);
    for (int i = 0; i < 3; ++i) gold[i] = r[i] - g[i];
    for (int it = 1; it <= maxiter; ++it) {
        gyrovector(f, t, gold, v, mass, charge, g);
        for (int i = 0; i < 3; ++i) gc[i] = r[i] - g[i];
        for (int i = 0; i < 3; ++i) dd[i] = gc[i] - gold[i];
        if (norm3(dd) / norm3(gc) < tol) {
            double tp[4] = {t, gc[0], gc[1], gc[2]};
            double B[3];
            field_B(f, tp, B);
            *vp = dot3(v, B) / norm3(B);
            *spd = norm3(v);
            memcpy(gc_out, gc, sizeof gc);
            return ORC_OK;
        }
        memcpy(gold, gc, sizeof gc);
    }
    return ORC_ERR_GCITER;
}
/* :328-374 */
static void getperp(const double *v, double *o) {
    if (v[0] == 0.0) { o[0] = 1; o[1] = 0; o[2] = 0; return; }
    if (v[1] == 0.0) { o[0] = 0; o[1] = 1; o[2] = 0; return; }
    if (v[2] == 0.0) { o[0] = 0; o[1] = 0; o[2] = 1; return; }
    double cc = -1.0 * (v[0] + v[1]) / v[2];
    double nn = sqrt(2.0 + sq(cc));
    o[0] = 1.0 / nn; o[1] = 1.0 / nn; o[2] = cc / nn;
}
/* :376-433 */
void orc_gctofp(const orc_field *f, double t, const double R[3], double vp, double speed,
                double mass, double charge, double gyrophase, double pos[3], double vel[3]) {
    double tp[4] = {t, R[0], R[1], R[2]};
    double B[3], b[3], u[3], w[3];
    field_B(f, tp, B);
    double Bsq = dot3(B, B);
    double sB = sqrt(Bsq);
    for (int i = 0; i < 3; ++i) b[i] = B[i] / sB;
    double pa = acos(vp / speed);
    double rc = orc_cyclotron_radius2(f, t, R, vp, speed, mass, charge);
    getperp(B, u);
    double un = sqrt(dot3(u, u));
    for (int i = 0; i < 3; ++i) u[i] = u[i] / un;
    cross3(b, u, w);
    double s = (charge > 0.0) ? 1.0 : ((charge < 0.0) ? -1.0 : 0.0);
    double cg = cos(gyrophase), sg = sin(gyrophase), cp = cos(pa), sp = sin(pa);
    for (int i = 0; i < 3; ++i) {
        pos[i] = R[i] + rc * (cg * u[i] + sg * w[i]);
        vel[i] = speed * (cp * b[i] + s * sp * sg * u[i] - s * sp * cg * w[i]);
    }
}

/* ======================================================================
 * Tracers
 * ==================================================================== */
typedef struct {
    const orc_field *f;
    const orc_params *p;
    double mass, charge, gm;
} lorentz_ctx;
static void lorentz_fn(void *c, double t, const double *y, double *o) {
    lorentz_ctx *L = (lorentz_ctx *)c;
    orc_rhs_lorentz(L->f, L->p, L->mass, L->charge, &L->gm, t, y, o);
}
typedef struct {
    const orc_field *f;
    const orc_params *p;
    int eom;
    double mass, charge, mu, v;
} gc_ctx;
static void gc_fn(void *c, double t, const double *y, double *o) {
    gc_ctx *G = (gc_ctx *)c;
    orc_rhs_gc(G->f, G->p, G->eom, G->mass, G->charge, G->mu, G->v, t, y, o);
}

static int finite_vec(const double *v, int n) {
    for (int i = 0; i < n; ++i)
        if (!isfinite(v[i])) return 0;
    return 1;
}

/* Particle.isadiabatic, rapt/Particle.py:345-384 (cycrad :483-488, cycper :489-494) */
static int particle_isadiabatic(const orc_field *f, const orc_params *p, const double *row,
                                double mass, double charge) {
    const double *mom = row + 4;
    double gm = sqrt(sq(mass) + dot3(mom, mom) / C_LIGHT_SQ);
    double v[3] = {mom[0] / gm, mom[1] / gm, mom[2] / gm};
    double cr = orc_cyclotron_radius(f, row[0], row + 1, v, mass, charge);
    int ok = cr / field_lengthscale(f, row) < p->epss;
    if (f->is_static || !ok) return ok;
    double cp = orc_cyclotron_period(f, row[0], row + 1, v, mass, charge);
    return cp / field_timescale(f, row) < p->epst;
}

/* GuidingCenter.isadiabatic, rapt/GuidingCenter.py:287-327 (cycrad :517-529,
 * cycper :531-541 with its pp**2 term kept as written) */
static int gc_isadiabatic(const orc_field *f, const orc_params *p, const double *row,
                          double mass, double charge, double mu) {
    double Bmag = field_magB(f, row);
    double pp = row[4];
    double gamma = sqrt(1.0 + 2.0 * mu * Bmag / (mass * C_LIGHT * C_LIGHT) +
                        sq(pp / mass / C_LIGHT));
    double vp, v;
    if (gamma - 1.0 < 1e-6) {
        vp = pp / mass;
        v = sqrt(2.0 * mu * Bmag / mass + sq(vp));
    } else {
        vp = pp / mass / gamma;
        v = C_LIGHT * sqrt(1.0 - 1.0 / sq(gamma));
    }
    double cr = orc_cyclotron_radius2(f, row[0], row + 1, vp, v, mass, charge);
    int ok = cr / field_lengthscale(f, row) < p->epss;
    if (f->is_static || !ok) return ok;
    double g2 = sqrt(1.0 + 2.0 * mu * Bmag / (mass * C_LIGHT * C_LIGHT) + sq(pp));
    double v2;
    if (g2 - 1.0 < 1e-6) {
        double vp2 = pp / mass;
        v2 = sqrt(2.0 * mu * Bmag / mass + sq(vp2));
    } else {
        v2 = C_LIGHT * sqrt(1.0 - 1.0 / sq(g2));
    }
    double cp = orc_cyclotron_period2(f, row[0], row + 1, v2, mass, charge);
    return cp / field_timescale(f, row) < p->epst;
}

/* RKF45 stepper for the tracers: one rkf() call (rapt/rkf.py:13-143) per
 * output interval, h starting at hmax, keeping only the end point. */
static int rkf_interval(orc_rhs_fn fn, void *ctx, int n, double *x, double *y, double xend,
                        const orc_params *p, orc_ode *st) {
    static const double a2 = 2.500000000000000e-01, a3 = 3.750000000000000e-01,
                        a4 = 9.230769230769231e-01, a5 = 1.000000000000000e+00,
                        a6 = 5.000000000000000e-01;
    static const double b21 = 2.500000000000000e-01, b31 = 9.375000000000000e-02,
                        b32 = 2.812500000000000e-01, b41 = 8.793809740555303e-01,
                        b42 = -3.277196176604461e+00, b43 = 3.320892125625853e+00,
                        b51 = 2.032407407407407e+00, b52 = -8.000000000000000e+00,
                        b53 = 7.173489278752436e+00, b54 = -2.058966861598441e-01,
                        b61 = -2.962962962962963e-01, b62 = 2.000000000000000e+00,
                        b63 = -1.381676413255361e+00, b64 = 4.529727095516569e-01,
                        b65 = -2.750000000000000e-01;
    static const double r1 = 2.777777777777778e-03, r3 = -2.994152046783626e-02,
                        r4 = -2.919989367357789e-02, r5 = 2.000000000000000e-02,
                        r6 = 3.636363636363636e-02;
    static const double c1 = 1.157407407407407e-01, c3 = 5.489278752436647e-01,
                        c4 = 5.353313840155945e-01, c5 = -2.000000000000000e-01;
    double k1[NMAXDIM], k2[NMAXDIM], k3[NMAXDIM], k4[NMAXDIM], k5[NMAXDIM], k6[NMAXDIM],
        tmp[NMAXDIM];
    const double tol = p->rkf_tol, hmax = p->rkf_hmax, hmin = p->rkf_hmin;
    double t = *x, h = hmax;
    int rc = 1;
    while (t < xend) {
        if (t + h > xend) h = xend - t;
        fn(ctx, t, y, k1);
        for (int i = 0; i < n; ++i) { k1[i] = h * k1[i]; tmp[i] = y[i] + b21 * k1[i]; }
        fn(ctx, t + a2 * h, tmp, k2);
        for (int i = 0; i < n; ++i) { k2[i] = h * k2[i]; tmp[i] = y[i] + b31 * k1[i] + b32 * k2[i]; }
        fn(ctx, t + a3 * h, tmp, k3);
        for (int i = 0; i < n; ++i) {
            k3[i] = h * k3[i];
            tmp[i] = y[i] + b41 * k1[i] + b42 * k2[i] + b43 * k3[i];
        }
        fn(ctx, t + a4 * h, tmp, k4);
        for (int i = 0; i < n; ++i) {
            k4[i] = h * k4[i];
            tmp[i] = y[i] + b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
        }
        fn(ctx, t + a5 * h, tmp, k5);
        for (int i = 0; i < n; ++i) {
            k5[i] = h * k5[i];
            tmp[i] = y[i] + b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
        }
        fn(ctx, t + a6 * h, tmp, k6);
        double r = 0.0;
        for (int i = 0; i < n; ++i) {
            k6[i] = h * k6[i];
            double ri = fabs(r1 * k1[i] + r3 * k3[i] + r4 * k4[i] + r5 * k5[i] + r6 * k6[i]) / h;
            if (i == 0 || ri > r) r = ri;
        }
        st->cnt.nfcn += 6;
        st->cnt.nstep += 1;
        if (r <= tol) {
            t = t + h;
            for (int i = 0; i < n; ++i)
                y[i] = y[i] + c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i];
            st->cnt.naccpt += 1;
        } else {
            st->cnt.nrejct += 1;
        }
        h = h * dmin(dmax(0.84 * pow(tol / r, 0.25), 0.1), 4.0);
        if (h > hmax) {
            h = hmax;
        } else if (h < hmin || !(h > 0.0)) {
            rc = ORC_ERR_HMIN;
            break;
        }
    }
    *x = t;
    st->istate = rc;
    return rc;
}

static int step_call(int stepper, orc_rhs_fn fn, void *ctx, int n, double *x, double *y,
                     double xend, const orc_params *p, orc_ode *st) {
    if (stepper == ORC_STEP_DOP853) return orc_dop853_call(fn, ctx, n, x, y, xend, p, st);
    if (stepper == ORC_STEP_DOPRI5) return orc_dopri5_call(fn, ctx, n, x, y, xend, p, st);
    return rkf_interval(fn, ctx, n, x, y, xend, p, st);
}

int64_t orc_particle_advance(const orc_field *f, const orc_params *p, const double last[7],
                             double mass, double charge, double delta, double *rows,
                             int64_t cap, double *tcur, int32_t *status, int32_t *switched,
                             orc_counters *cnt) {
    const double t0 = last[0];
    const double *mom = last + 4;
    lorentz_ctx L = {f, p, mass, charge, 0.0};
    L.gm = sqrt(sq(mass) + dot3(mom, mom) / C_LIGHT_SQ);
    double vel[3] = {mom[0] / L.gm, mom[1] / L.gm, mom[2] / L.gm};
    /* rapt/Particle.py:282 (Ptimestep is overridden by the cyclotron period) */
    double dt = orc_cyclotron_period(f, t0, last + 1, vel, mass, charge) / p->cyclotronresolution;
    double y[6];
    memcpy(y, last + 1, sizeof y);
    double x = t0;
    orc_ode st;
    memset(&st, 0, sizeof st);
    st.h = p->first_step;
    st.istate = 1;
    int64_t n = 0;
    *status = ORC_OK;
    *switched = 0;
    if (tcur) *tcur = t0;
    const double tend = t0 + delta;
    while (st.istate >= 0 && x < tend) {
        double tnext = x + dt;
        int rc = step_call(p->stepper, lorentz_fn, &L, 6, &x, y, tnext, p, &st);
        if (n >= cap) { *status = ORC_ERR_CAP; break; }
        double *row = rows + 7 * n;
        row[0] = tnext;
        memcpy(row + 1, y, sizeof y);
        n++;
        if (tcur) *tcur = x + dt;
        if (rc < 0) { *status = rc; break; }
        if (!finite_vec(y, 6)) { *status = ORC_ERR_NAN; break; }
        if (p->check_adiabaticity && particle_isadiabatic(f, p, row, mass, charge)) {
            *switched = 1;
            break;
        }
    }
    if (cnt) {
        cnt->naccpt += st.cnt.naccpt; cnt->nrejct += st.cnt.nrejct;
        cnt->nstep += st.cnt.nstep; cnt->nfcn += st.cnt.nfcn;
    }
    return n;
}

int64_t orc_gc_advance(const orc_field *f, const orc_params *p, int eom, const double last[5],
                       double mass, double charge, double mu, double v, double dt,
                       double delta, double *rows, int64_t cap, double *tcur,
                       int32_t *status, int32_t *switched, orc_counters *cnt) {
    const double t0 = last[0];
    gc_ctx G = {f, p, eom, mass, charge, mu, v};
    double y[4];
    memcpy(y, last + 1, sizeof y);
    double x = t0;
    orc_ode st;
    memset(&st, 0, sizeof st);
    st.h = p->first_step;
    st.istate = 1;
    int64_t n = 0;
    *status = ORC_OK;
    *switched = 0;
    if (tcur) *tcur = t0;
    const double tend = t0 + delta;
    int stepper = p->stepper == ORC_STEP_DOP853 ? ORC_STEP_DOPRI5 : p->stepper;
    while (st.istate >= 0 && x < tend) {
        int rc = step_call(stepper, gc_fn, &G, 4, &x, y, x + dt, p, &st);
        if (n >= cap) { *status = ORC_ERR_CAP; break; }
        double *row = rows + 5 * n;
        row[0] = x;
        memcpy(row + 1, y, sizeof y);
        n++;
        if (tcur) *tcur = x;
        if (rc < 0) { *status = rc; break; }
        if (!finite_vec(y, 4)) { *status = ORC_ERR_NAN; break; }
        if (p->check_adiabaticity && !gc_isadiabatic(f, p, row, mass, charge, mu)) {
            *switched = 1;
            break;
        }
    }
    if (cnt) {
        cnt->naccpt += st.cnt.naccpt; cnt->nrejct += st.cnt.nrejct;
        cnt->nstep += st.cnt.nstep; cnt->nfcn += st.cnt.nfcn;
    }
    return n;
}

/* Particle.__init__ (rapt/Particle.py:105-109) */
static void particle_init_row(double t0, const double pos[3], const double vel[3], double mass,
                              double *row7) {
    double gamma = 1.0 / sqrt(1.0 - dot3(vel, vel) / C_LIGHT_SQ);
    row7[0] = t0;
    row7[1] = pos[0]; row7[2] = pos[1]; row7[3] = pos[2];
    for (int i = 0; i < 3; ++i) row7[4 + i] = mass * gamma * vel[i];
}
/* GuidingCenter.init(Particle) (rapt/GuidingCenter.py:168-186) then __init__
 * with ppar (:123-133).  row5 out, *mu and *v out. */
static int gc_from_particle(const orc_field *f, const double *prow, double mass, double charge,
                            double *row5, double *mu, double *vout) {
    const double *mom = prow + 4;
    double gm = sqrt(sq(mass) + dot3(mom, mom) / C_LIGHT_SQ);
    double vel[3] = {mom[0] / gm, mom[1] / gm, mom[2] / gm};
    double pos[3], vp, v;
    int rc = orc_guidingcenter(f, prow[0], prow + 1, vel, mass, charge, 1e-3, 20, pos, &vp, &v);
    if (rc != ORC_OK) return rc;
    double gamma = 1.0 / sqrt(1.0 - sq(v / C_LIGHT));
    double ppar = mass * gamma * vp;
    /* __init__: gamma recomputed, mu from ppar/(mass*gamma) */
    double g2 = 1.0 / sqrt(1.0 - sq(v / C_LIGHT));
    *mu = orc_magnetic_moment(f, prow[0], pos, ppar / (mass * g2), v, mass);
    *vout = v;
    row5[0] = prow[0];
    row5[1] = pos[0]; row5[2] = pos[1]; row5[3] = pos[2];
    row5[4] = ppar;
    return ORC_OK;
}
/* Particle.init(GuidingCenter) (rapt/Particle.py:149-164); GCtoFP is evaluated
 * at the new object's tcur, which is 0 for Particle(). */
static void particle_from_gc(const orc_field *f, const double *grow, double mass,
                             double charge, double mu, double *row7) {
    double B = field_magB(f, grow);
    double pp = grow[4];
    double gammasq = 1.0 + 2.0 * mu * B / (mass * C_LIGHT * C_LIGHT) + sq(pp / mass / C_LIGHT);
    double v;
    if (sqrt(gammasq) - 1.0 < 1e-6)
        v = sqrt(2.0 * mu * B / mass + sq(pp / mass));
    else
        v = C_LIGHT * sqrt(1.0 - 1.0 / gammasq);
    double vpar = pp / mass / sqrt(gammasq);
    double pos[3], vel[3];
    orc_gctofp(f, 0.0, grow + 1, vpar, v, mass, charge, 0.0, pos, vel);
    particle_init_row(grow[0], pos, vel, mass, row7);
}

static void put_row8(double *dst, int mode, const double *src, double mu) {
    dst[0] = (double)mode;
    if (mode == 0) {
        memcpy(dst + 1, src, 7 * sizeof(double));
    } else {
        memcpy(dst + 1, src, 5 * sizeof(double));
        dst[6] = mu;
        dst[7] = 0.0;
    }
}

int orc_particle_isadiabatic(const orc_field *f, const orc_params *p, const double row7[7],
                             double mass, double charge) {
    return particle_isadiabatic(f, p, row7, mass, charge);
}
int orc_gc_isadiabatic(const orc_field *f, const orc_params *p, const double row5[5],
                       double mass, double charge, double mu) {
    return gc_isadiabatic(f, p, row5, mass, charge, mu);
}
int orc_adaptive_to_gc(const orc_field *f, const double row7[7], double mass, double charge,
                       double row5[5], double *mu, double *v) {
    return gc_from_particle(f, row7, mass, charge, row5, mu, v);
}
void orc_adaptive_to_particle(const orc_field *f, const double row5[5], double mass,
                              double charge, double mu, double row7[7]) {
    particle_from_gc(f, row5, mass, charge, mu, row7);
}

/* The switching loop of Adaptive.advance (rapt/Adaptive.py:187-222) from the
 * last tracer of trajlist: mode 0 a Particle with last row prow, mode 1 a
 * GuidingCenter with last row grow, magnetic moment mu and speed v.  The
 * first segment's rows continue that tracer (no start row); every switch
 * appends the new tracer's start row.  As written there, t starts at 0 and
 * becomes the new tracer's absolute tcur after a switch. */
static int64_t adaptive_loop(const orc_field *f, const orc_params *pp, int mode, double *prow,
                             double *grow, double mu, double v, double mass, double charge,
                             double delta, double *rows, int64_t n, int64_t cap,
                             int32_t *nswitch, int32_t *status, orc_counters *cnt) {
    double t = 0.0;
    double *buf = (double *)malloc(sizeof(double) * 7 * (size_t)(cap > 0 ? cap : 1));
    while (t < delta) {
        int32_t st = ORC_OK, sw = 0;
        double tc = 0.0;
        int64_t m;
        if (mode == 0) {
            m = orc_particle_advance(f, pp, prow, mass, charge, delta - t, buf, cap - n, &tc,
                                     &st, &sw, cnt);
            for (int64_t i = 0; i < m; ++i) put_row8(rows + 8 * (n + i), 0, buf + 7 * i, 0.0);
            n += m;
            if (m > 0) memcpy(prow, buf + 7 * (m - 1), 7 * sizeof(double));
        } else {
            /* GuidingCenter.advance sets up scipy dopri5 with its defaults
             * (rapt/GuidingCenter.py:450) */
            orc_params pg = *pp;
            pg.stepper = ORC_STEP_DOPRI5;
            pg.safety = 0.9; pg.dfactor = 0.2; pg.ifactor = 10.0; pg.beta = 0.0;
            m = orc_gc_advance(f, &pg, ORC_EOM_TCB, grow, mass, charge, mu, v, pp->gctimestep,
                               delta - t, buf, cap - n, &tc, &st, &sw, cnt);
            for (int64_t i = 0; i < m; ++i) put_row8(rows + 8 * (n + i), 1, buf + 5 * i, mu);
            n += m;
            if (m > 0) memcpy(grow, buf + 5 * (m - 1), 5 * sizeof(double));
        }
        if (st != ORC_OK) { *status = st; break; }
        if (sw) {
            if (n >= cap) { *status = ORC_ERR_CAP; break; }
            if (mode == 0) {
                int rc = gc_from_particle(f, prow, mass, charge, grow, &mu, &v);
                if (rc != ORC_OK) { *status = rc; break; }
                mode = 1;
                put_row8(rows + 8 * n++, 1, grow, mu);
                t = grow[0];
            } else {
                particle_from_gc(f, grow, mass, charge, mu, prow);
                mode = 0;
                put_row8(rows + 8 * n++, 0, prow, 0.0);
                t = prow[0];
            }
            (*nswitch)++;
        } else {
            t = tc;
        }
    }
    free(buf);
    return n;
}

int64_t orc_adaptive_run(const orc_field *f, const orc_params *p, const double pos[3],
                         const double vel[3], double t0, double mass, double charge,
                         double delta, double *rows, int64_t cap, int32_t *nswitch,
                         int32_t *status, orc_counters *cnt) {
    orc_params pp = *p;
    pp.check_adiabaticity = 1;
    double prow[7], grow[5], mu = 0.0, v = 0.0;
    int64_t n = 0;
    int mode;
    *nswitch = 0;
    *status = ORC_OK;
    if (pp.gctimestep <= 0.0) { *status = -10; return 0; }
    /* Adaptive.__init__ (rapt/Adaptive.py:70-104) */
    particle_init_row(t0, pos, vel, mass, prow);
    if (particle_isadiabatic(f, &pp, prow, mass, charge)) {
        int rc = gc_from_particle(f, prow, mass, charge, grow, &mu, &v);
        if (rc != ORC_OK) { *status = rc; return 0; }
        mode = 1;
    } else {
        mode = 0;
    }
    if (n >= cap) { *status = ORC_ERR_CAP; return n; }
    put_row8(rows + 8 * n++, mode, mode == 0 ? prow : grow, mu);
    return adaptive_loop(f, &pp, mode, prow, grow, mu, v, mass, charge, delta, rows, n, cap,
                         nswitch, status, cnt);
}

int64_t orc_adaptive_continue(const orc_field *f, const orc_params *p, int mode,
                              const double *last, double mu, double v, double mass,
                              double charge, double delta, double *rows, int64_t cap,
                              int32_t *nswitch, int32_t *status, orc_counters *cnt) {
    orc_params pp = *p;
    pp.check_adiabaticity = 1;
    double prow[7] = {0}, grow[5] = {0};
    *nswitch = 0;
    *status = ORC_OK;
    if (pp.gctimestep <= 0.0) { *status = -10; return 0; }
    if (mode == 0) memcpy(prow, last, sizeof prow);
    else memcpy(grow, last, sizeof grow);
    return adaptive_loop(f, &pp, mode, prow, grow, mu, v, mass, charge, delta, rows, 0, cap,
                         nswitch, status, cnt);
}

/* ======================================================================
 * Batch drivers (CPU baseline): OpenMP over particles
 * ==================================================================== */
int orc_batch_particle(const orc_field *f, const orc_params *p, int64_t n, const double *last7,
                       const double *mass, const double *charge, double delta, double *rows,
                       const int64_t *row_off, const int64_t *row_cap, int64_t *nrows,
                       int32_t *status, orc_counters *cnt, int nthreads) {
    int64_t acc = 0, rej = 0, nst = 0, nfc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : acc, rej, nst, nfc)
    for (int64_t i = 0; i < n; ++i) {
        orc_counters c = {0, 0, 0, 0};
        int32_t sw = 0;
        double tc;
        nrows[i] = orc_particle_advance(f, p, last7 + 7 * i, mass[i], charge[i], delta,
                                        rows + 7 * row_off[i], row_cap[i], &tc, &status[i], &sw,
                                        &c);
        acc += c.naccpt; rej += c.nrejct; nst += c.nstep; nfc += c.nfcn;
    }
    if (cnt) { cnt->naccpt = acc; cnt->nrejct = rej; cnt->nstep = nst; cnt->nfcn = nfc; }
    return 0;
}

int orc_batch_gc(const orc_field *f, const orc_params *p, int eom, int64_t n, const double *last5,
                 const double *mass, const double *charge, const double *mu, const double *v,
                 const double *dt, double delta, double *rows, const int64_t *row_off,
                 const int64_t *row_cap, int64_t *nrows, int32_t *status, orc_counters *cnt,
                 int nthreads) {
    int64_t acc = 0, rej = 0, nst = 0, nfc = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : acc, rej, nst, nfc)
    for (int64_t i = 0; i < n; ++i) {
        orc_counters c = {0, 0, 0, 0};
        int32_t sw = 0;
        double tc;
        nrows[i] = orc_gc_advance(f, p, eom, last5 + 5 * i, mass[i], charge[i], mu[i], v[i],
                                  dt[i], delta, rows + 5 * row_off[i], row_cap[i], &tc,
                                  &status[i], &sw, &c);
        acc += c.naccpt; rej += c.nrejct; nst += c.nstep; nfc += c.nfcn;
    }
    if (cnt) { cnt->naccpt = acc; cnt->nrejct = rej; cnt->nstep = nst; cnt->nfcn = nfc; }
    return 0;
}

/* ======================================================================
 * BounceCenter, equatorial branch: rapt/BounceCenter.py:154-204, 237-238
 * ==================================================================== */
typedef struct {
    const orc_field *f;
    double mu, mass, charge, gamma;
} eqctx;

/* _v_grad (:154-172): mu (B x gradB) / (q gamma |B|^2) */
static void eq_vgrad(const eqctx *c, const double *tp, double *out) {
    double B[3], g[3], x[3];
    field_B(c->f, tp, B);
    const double bsq = dot3(B, B);
    field_gradB(c->f, tp, g);
    cross3(B, g, x);
    const double den = c->charge * c->gamma * bsq;
    for (int k = 0; k < 3; ++k) out[k] = c->mu * x[k] / den;
}

/* _v_inert (:174-204): the 4-vector tpos shifted by d (t, vg/|vg|) */
static void eq_vinert(const eqctx *c, const double *tp, double *out) {
    double vg[3], v1[3], v2[3], B[3], dv[3], x[3], dir[4], tm[4], tq[4];
    eq_vgrad(c, tp, vg);
    const double vm = sqrt(dot3(vg, vg));
    dir[0] = tp[0];
    for (int k = 0; k < 3; ++k) dir[k + 1] = vg[k] / vm;
    const double d = c->f->gstep;
    for (int k = 0; k < 4; ++k) {
        tm[k] = tp[k] - d * dir[k];
        tq[k] = tp[k] + d * dir[k];
    }
    eq_vgrad(c, tm, v1);
    eq_vgrad(c, tq, v2);
    field_B(c->f, tp, B);
    const double bsq = dot3(B, B);
    for (int k = 0; k < 3; ++k) dv[k] = v2[k] - v1[k];
    cross3(B, dv, x);
    const double s = c->mass * vm / (c->charge * bsq * 2 * d);
    for (int k = 0; k < 3; ++k) out[k] = s * x[k];
}

/* deriv's equatorial branch with tpos = (t, Y): _v_grad + _v_inert */
static void eq_rhs(void *vc, double t, const double *y, double *out) {
    const eqctx *c = (const eqctx *)vc;
    const double tp[4] = {t, y[0], y[1], y[2]};
    double vg[3], vi[3];
    eq_vgrad(c, tp, vg);
    eq_vinert(c, tp, vi);
    for (int k = 0; k < 3; ++k) out[k] = vg[k] + vi[k];
}

/* _v_grad and _v_inert at one 4-vector tpos: out[0..2], out[3..5] */
void orc_bc_equatorial_drifts(const orc_field *f, const double tpos[4], double mu, double mass,
                              double charge, double gamma, double *out) {
    eqctx c = {f, mu, mass, charge, gamma};
    eq_vgrad(&c, tpos, out);
    eq_vinert(&c, tpos, out + 3);
}

int64_t orc_bc_equatorial(const orc_field *f, const orc_params *p, double tpos[4], double mu,
                          double mass, double charge, double gamma, double dt, double tcur,
                          int64_t nrow, double *rows, int32_t *status, orc_counters *cnt) {
    eqctx c = {f, mu, mass, charge, gamma};
    double x = tpos[0], y[3] = {tpos[1], tpos[2], tpos[3]};
    orc_ode st;
    memset(&st, 0, sizeof st);
    st.h = p->first_step;
    int rc = 1;
    int64_t r = 0;
    for (; r < nrow; ++r) {
        rc = orc_dopri5_call(eq_rhs, &c, 3, &x, y, x + dt, p, &st);
        if (rc < 0) break;
        double *row = rows + 4 * r;
        row[0] = tcur + (double)r * dt;
        row[1] = y[0]; row[2] = y[1]; row[3] = y[2];
    }
    tpos[0] = x; tpos[1] = y[0]; tpos[2] = y[1]; tpos[3] = y[2];
    if (status) *status = rc;
    if (cnt) *cnt = st.cnt;
    return r;
}
