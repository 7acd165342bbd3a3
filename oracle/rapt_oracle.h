/*
 * rapt_oracle.h -- CPU FP64 restatement of RAPT's tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (rapt_amd/ + librapt_hip.so) never links or calls it.
 *
 * Everything here follows the reference's Python arithmetic operation by
 * operation (citations are to the reference's files).  Two numpy behaviours
 * measured in the container that generated the golden fixtures are mirrored:
 *   - np.dot of two 3-vectors evaluates fma(a2,b2, fma(a1,b1, a0*b0));
 *   - np.dot(_M1, beta) (rapt/fields.py:200) accumulates the first 16 terms in
 *     two interleaved partial sums (even / odd j), adds them, then adds j=16,17.
 * The embedded Runge-Kutta codes restate Hairer & Wanner's DOPRI5 / DOP853
 * (Solving ODEs I, 2nd ed., 1993) with scipy.integrate.ode's call semantics
 * (scipy/integrate/_ode.py:1121-1250): the step size persists across
 * integrate() calls (HINIT only on the first call), FACOLD resets per call,
 * at most nsteps+1 steps per call.
 */
#ifndef RAPT_ORACLE_H
#define RAPT_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* field models (same numbering as the product's descriptor) */
enum {
    ORC_FIELD_ZERO = 0,
    ORC_FIELD_EARTH_DIPOLE = 1,      /* p0 = coeff (-3 B0 Re^3)                        */
    ORC_FIELD_DOUBLE_DIPOLE = 2,     /* p0 = coeff (-B0 Re^3), p1 = distance, p2 = k   */
    ORC_FIELD_UNIFORM_BZ = 3,        /* p0 = Bz                                        */
    ORC_FIELD_UNIFORM_CROSSED_EB = 4,/* p0 = Bz, p1 = Ey                               */
    ORC_FIELD_VAR_EARTH_DIPOLE = 5,  /* p0 = -B0*Re**3, p1 = amp, p2 = period          */
    ORC_FIELD_PARABOLIC = 6,         /* p0 = B0, p1 = Bn, p2 = d, p3 = module B0        */
    ORC_FIELD_CURRENT_SHEET = 7,     /* p0 = B0, p1 = lambda, p2 = Bn, p3..5 = E        */
    ORC_FIELD_DIPOLE_UNIFORM = 8,    /* p0 = coeff (-3 B0 Re^3), p1..3 = uniform B      */
    ORC_FIELD_EXPR = 10              /* user-written B/E as register programs (code)    */
};

typedef struct {
    int32_t model;
    int32_t is_static;
    double gstep;   /* gradientstepsize  */
    double tstep;   /* timederivstepsize */
    double p[16];
    /* ORC_FIELD_EXPR: instructions (opcode, dest, a, b) over 32 registers,
     * 0-3 = t, x, y, z, results in 4-6; B's program is code[0, nb), E's
     * code[nb, nb + ne) (ne = 0: E = 0) */
    const int32_t *code;
    const double *consts;
    int32_t nb, ne;
} orc_field;

/* what= selector for orc_field_eval */
enum {
    ORC_EVAL_B = 0, ORC_EVAL_E = 1, ORC_EVAL_UNITB = 2, ORC_EVAL_MAGB = 3,
    ORC_EVAL_GRADB = 4, ORC_EVAL_JACOBIANB = 5, ORC_EVAL_CURVATURE = 6,
    ORC_EVAL_CURLB = 7, ORC_EVAL_DBDT_MAG = 8, ORC_EVAL_DBDT_VEC = 9,
    ORC_EVAL_LENGTHSCALE = 10, ORC_EVAL_TIMESCALE = 11
};

enum { ORC_EOM_TCB = 0, ORC_EOM_BC = 1, ORC_EOM_NT = 2 };
enum { ORC_STEP_DOP853 = 0, ORC_STEP_DOPRI5 = 1, ORC_STEP_RKF45 = 2 };

/* status codes (per particle) */
enum {
    ORC_OK = 1, ORC_ERR_NSTEPS = -2, ORC_ERR_HSMALL = -3, ORC_ERR_NAN = -4,
    ORC_ERR_GCITER = -5, ORC_ERR_CAP = -6, ORC_ERR_HMIN = -7
};

/* rapt.params (rapt/__init__.py:21-34) plus the scipy integrator knobs */
typedef struct {
    double rtol, atol;                 /* solvertolerances                */
    double cyclotronresolution;
    double bounceresolution;
    double gctimestep;
    double epss, epst;
    int32_t enforce_equatorial;
    int32_t nsteps;                    /* scipy nsteps (500)              */
    int32_t stepper;                   /* ORC_STEP_*                      */
    int32_t check_adiabaticity;
    double safety, dfactor, ifactor, beta;   /* scipy set_integrator args */
    double max_step, first_step;
    double rkf_tol, rkf_hmax, rkf_hmin;      /* RKF45 stepper             */
} orc_params;

typedef struct {
    int64_t naccpt, nrejct, nstep, nfcn;
} orc_counters;

/* persistent integrator state of one scipy ode object */
typedef struct {
    double h;          /* work[6]: step carried between integrate() calls */
    int32_t istate;
    int32_t pad;
    orc_counters cnt;  /* cumulative over calls */
} orc_ode;

typedef void (*orc_rhs_fn)(void *ctx, double t, const double *y, double *out);

/* ---- fields --------------------------------------------------------- */
int orc_field_eval(const orc_field *f, int what, const double tpos[4], double *out);
int orc_field_eval_batch(const orc_field *f, int what, int64_t n, const double *tpos, double *out);

/* ---- right-hand sides ------------------------------------------------ */
void orc_rhs_lorentz(const orc_field *f, const orc_params *p, double mass, double charge,
                     double *gm_io, double t, const double *y, double *out);
void orc_rhs_gc(const orc_field *f, const orc_params *p, int eom, double mass, double charge,
                double mu, double v, double t, const double *y, double *out);

/* ---- integrators (one scipy integrate() call) ----------------------- */
int orc_dopri5_call(orc_rhs_fn fn, void *ctx, int n, double *x, double *y, double xend,
                    const orc_params *p, orc_ode *st);
int orc_dop853_call(orc_rhs_fn fn, void *ctx, int n, double *x, double *y, double xend,
                    const orc_params *p, orc_ode *st);
/* rapt/rkf.py:13-143; returns accepted points (incl. x0) into tout/xout (cap rows) */
int orc_rkf(orc_rhs_fn fn, void *ctx, int n, double a, double b, const double *x0,
            double tol, double hmax, double hmin, double *tout, double *xout, int64_t cap,
            int64_t *npts);

/* built-in test problems for the integrators (harmonic oscillator, Arenstorf) */
int orc_test_problem_run(int problem, int stepper, const orc_params *p, int n, double t0,
                         const double *y0, double dt, int64_t ncalls, double *rows,
                         orc_counters *cnt, int32_t *istate);

/* ---- physics helpers (rapt/utils.py) -------------------------------- */
double orc_cyclotron_period(const orc_field *f, double t, const double pos[3],
                            const double vel[3], double mass, double charge);
double orc_cyclotron_period2(const orc_field *f, double t, const double pos[3], double speed,
                             double mass, double charge);
double orc_cyclotron_radius(const orc_field *f, double t, const double pos[3],
                            const double vel[3], double mass, double charge);
double orc_cyclotron_radius2(const orc_field *f, double t, const double pos[3], double vpar,
                             double v, double mass, double charge);
double orc_magnetic_moment(const orc_field *f, double t, const double pos[3], double vpar,
                           double v, double mass);
int orc_guidingcenter(const orc_field *f, double t, const double r[3], const double v[3],
                      double mass, double charge, double tol, int maxiter,
                      double gc_out[3], double *vp, double *spd);
void orc_gctofp(const orc_field *f, double t, const double R[3], double vp, double speed,
                double mass, double charge, double gyrophase, double pos[3], double vel[3]);

/* ---- tracers ---------------------------------------------------------- */
/* Particle.advance (rapt/Particle.py:230-309).  last[7] = (t,x,y,z,px,py,pz).
 * rows: appended rows (7 doubles each).  Returns number of rows appended. */
int64_t orc_particle_advance(const orc_field *f, const orc_params *p, const double last[7],
                             double mass, double charge, double delta, double *rows,
                             int64_t cap, double *tcur, int32_t *status, int32_t *switched,
                             orc_counters *cnt);
/* GuidingCenter.advance (rapt/GuidingCenter.py:397-458).  last[5] = (t,x,y,z,ppar). */
int64_t orc_gc_advance(const orc_field *f, const orc_params *p, int eom, const double last[5],
                       double mass, double charge, double mu, double v, double dt,
                       double delta, double *rows, int64_t cap, double *tcur,
                       int32_t *status, int32_t *switched, orc_counters *cnt);

/* Adaptive (rapt/Adaptive.py:70-104,187-222).  Rows are 8 wide:
 * (mode, t, x, y, z, a, b, c) with mode 0 = Particle (a,b,c = px,py,pz) and
 * mode 1 = GuidingCenter (a = ppar, b = mu, c = 0).  The first row of every
 * segment is its initial state, as in the reference's trajlist. */
int64_t orc_adaptive_run(const orc_field *f, const orc_params *p, const double pos[3],
                         const double vel[3], double t0, double mass, double charge,
                         double delta, double *rows, int64_t cap, int32_t *nswitch,
                         int32_t *status, orc_counters *cnt);
/* A further Adaptive.advance(delta) from the last tracer of trajlist: mode 0 a
 * Particle whose last row is last[7], mode 1 a GuidingCenter with last row
 * last[5], moment mu and speed v.  Rows as orc_adaptive_run, without a start
 * row for the continued tracer. */
int64_t orc_adaptive_continue(const orc_field *f, const orc_params *p, int mode,
                              const double *last, double mu, double v, double mass,
                              double charge, double delta, double *rows, int64_t cap,
                              int32_t *nswitch, int32_t *status, orc_counters *cnt);
/* Particle.isadiabatic (rapt/Particle.py:345-384) of a Particle whose last row
 * is row7, and GuidingCenter.isadiabatic (rapt/GuidingCenter.py:287-327) of a
 * guiding centre with last row row5 and moment mu: 1 adiabatic, 0 not. */
int orc_particle_isadiabatic(const orc_field *f, const orc_params *p, const double row7[7],
                             double mass, double charge);
int orc_gc_isadiabatic(const orc_field *f, const orc_params *p, const double row5[5],
                       double mass, double charge, double mu);
/* The conversions of a switch: GuidingCenter.init(Particle)
 * (rapt/GuidingCenter.py:168-186, then __init__ :123-133) gives row5, mu and
 * the speed v (returns 1, or the fixed-point iteration's error); and
 * Particle.init(GuidingCenter) (rapt/Particle.py:149-164, GCtoFP at
 * gyrophase 0 and the new Particle's t = 0) gives row7. */
int orc_adaptive_to_gc(const orc_field *f, const double row7[7], double mass, double charge,
                       double row5[5], double *mu, double *v);
void orc_adaptive_to_particle(const orc_field *f, const double row5[5], double mass,
                              double charge, double mu, double row7[7]);

/* batch drivers with OpenMP over particles (CPU baseline).  SoA-free: state is
 * an array of per-particle records. */
int orc_batch_particle(const orc_field *f, const orc_params *p, int64_t n, const double *last7,
                       const double *mass, const double *charge, double delta, double *rows,
                       const int64_t *row_off, const int64_t *row_cap, int64_t *nrows,
                       int32_t *status, orc_counters *cnt, int nthreads);
int orc_batch_gc(const orc_field *f, const orc_params *p, int eom, int64_t n, const double *last5,
                 const double *mass, const double *charge, const double *mu, const double *v,
                 const double *dt, double delta, double *rows, const int64_t *row_off,
                 const int64_t *row_cap, int64_t *nrows, int32_t *status, orc_counters *cnt,
                 int nthreads);

/* BounceCenter.advance with isequatorial (rapt/BounceCenter.py:154-204,
 * 237-238): dopri5 (p) on _v_grad + _v_inert at (t, x, y, z), nrow
 * integrate() calls of dt; rows [nrow][4] with the reference's row times
 * tcur + r dt.  tpos in/out; returns rows written. */
void orc_bc_equatorial_drifts(const orc_field *f, const double tpos[4], double mu, double mass,
                              double charge, double gamma, double *out);
int64_t orc_bc_equatorial(const orc_field *f, const orc_params *p, double tpos[4], double mu,
                          double mass, double charge, double gamma, double dt, double tcur,
                          int64_t nrow, double *rows, int32_t *status, orc_counters *cnt);

#ifdef __cplusplus
}
#endif
#endif
