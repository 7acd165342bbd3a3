/*
 * rapt_hip.h -- C-ABI of librapt_hip.so, the MI355X (gfx950) particle tracer.
 *
 * The reference (mkozturk/rapt) is pure Python; its only native boundary on
 * the tracing path is scipy.integrate.ode driving a Python right-hand side
 * (rapt/Particle.py:300-305, rapt/GuidingCenter.py:450-453).  This library
 * replaces that boundary for a whole batch of particles at once: the Python
 * front end (rapt_amd/) packs the tracer objects' last trajectory rows and
 * rapt.params into the structs below and calls these entry points via ctypes.
 *
 * Conventions: plain pointers and sizes, no C++ types; every function returns
 * 0 on success and a negative RAPT_E_* code on failure (message via
 * rapt_last_error).  Host pointers are caller-owned; device buffers are owned
 * by the library (rapt_batch).  A context is bound to one device; calls on one
 * context are not thread-safe.
 */
#ifndef RAPT_HIP_H
#define RAPT_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define RAPT_OK 0
#define RAPT_E_ARG (-1)
#define RAPT_E_HIP (-2)
#define RAPT_E_OOM (-3)
#define RAPT_E_STATE (-4)

/* ---- field models (rapt/fields.py:282-511 plus two user-analytic forms) */
enum rapt_field_model {
    RAPT_FIELD_ZERO = 0,
    RAPT_FIELD_EARTH_DIPOLE = 1,       /* p0 = -3 B0 Re^3          rapt/fields.py:282-317 */
    RAPT_FIELD_DOUBLE_DIPOLE = 2,      /* p0 = -B0 Re^3, p1 = distance, p2 = imagestrength
                                                                   rapt/fields.py:319-362 */
    RAPT_FIELD_UNIFORM_BZ = 3,         /* p0 = Bz                  rapt/fields.py:364-390 */
    RAPT_FIELD_UNIFORM_CROSSED_EB = 4, /* p0 = Bz, p1 = Ey         rapt/fields.py:392-427 */
    RAPT_FIELD_VAR_EARTH_DIPOLE = 5,   /* p0 = -B0 Re^3, p1 = amp, p2 = period
                                                                   rapt/fields.py:429-470 */
    RAPT_FIELD_PARABOLIC = 6,          /* p0 = B0, p1 = Bn, p2 = d, p3 = module B0
                                                                   rapt/fields.py:472-511 */
    RAPT_FIELD_CURRENT_SHEET = 7,      /* B = (p0 tanh(z/p1), 0, p2), E = (p3, p4, p5) */
    RAPT_FIELD_DIPOLE_UNIFORM = 8,     /* EarthDipole(p0) + (p1, p2, p3)                */
    RAPT_FIELD_GRID = 9,               /* B, E sampled on a grid (rapt_grid_set)
                                                                   rapt/fields.py:513-814 */
    RAPT_FIELD_EXPR = 10,              /* a user-written field (rapt_field_expr_set)
                                                                   rapt/fields.py:5-74 */
    RAPT_FIELD_NMODELS = 11
};
#define RAPT_EXPR_MAXCODE 256
#define RAPT_EXPR_MAXCONST 128

/* Field descriptor: replaces the _Field plugin object (rapt/fields.py:5-41). */
typedef struct {
    int32_t model;      /* enum rapt_field_model */
    int32_t is_static;  /* _Field.static */
    double gstep;       /* _Field.gradientstepsize */
    double tstep;       /* _Field.timederivstepsize */
    double p[16];       /* model constants */
} rapt_field_desc;

/* ---- tracing modes ---------------------------------------------------- */
enum rapt_mode {
    RAPT_MODE_PARTICLE = 0, /* Particle.advance     rapt/Particle.py:230-309 */
    RAPT_MODE_GC = 1,       /* GuidingCenter.advance rapt/GuidingCenter.py:397-458 */
    RAPT_MODE_ADAPTIVE = 2  /* Adaptive.advance     rapt/Adaptive.py:187-222 */
};
enum rapt_eom { RAPT_EOM_TCB = 0, RAPT_EOM_BC = 1, RAPT_EOM_NT = 2 };
enum rapt_stepper { RAPT_STEP_DOP853 = 0, RAPT_STEP_DOPRI5 = 1, RAPT_STEP_RKF45 = 2 };
/* FP32_FIELD: field evaluations and the finite-difference drift terms in
 * float (gradient step max(gradientstepsize, |r|/1024)), state, stage sums,
 * error norm and step control in double.  The error norm scales each
 * component by the largest component of its 3-vector (position, momentum)
 * instead of componentwise, so the float rounding of the forces stays below
 * tolerance where a component passes through zero, and the step controller
 * runs with beta = 0 (I-controller): a PI term settles on an error level
 * below the float rounding floor and shrinks the step without bound.  Not a
 * reference mode. */
enum rapt_precision { RAPT_PREC_FP64 = 0, RAPT_PREC_FP32_FIELD = 1 };

/* Per-launch settings: rapt.params (rapt/__init__.py:21-34) plus the scipy
 * set_integrator arguments the tracers pass (rapt/Particle.py:300,
 * rapt/GuidingCenter.py:450). */
typedef struct {
    int32_t mode;                /* enum rapt_mode */
    int32_t eom;                 /* enum rapt_eom (GC segments) */
    int32_t stepper;             /* stepper of the mode's primary segment */
    int32_t precision;           /* enum rapt_precision */
    int32_t nsteps;              /* scipy nsteps (500) */
    int32_t enforce_equatorial;  /* params["enforce equatorial"] */
    int32_t check_adiabaticity;  /* raise Adiabatic / NonAdiabatic -> stop (PARTICLE/GC) */
    int32_t block;               /* threads per block (0 = default: 64 for launches of
                                    fewer than four waves per CU, else 256) */
    double rtol, atol;           /* params["solvertolerances"] */
    double safety, dfactor, ifactor, beta;  /* scipy dop853/dopri5 knobs */
    double max_step, first_step;
    double gc_rtol, gc_atol;     /* ADAPTIVE: tolerances of GC segments */
    double rkf_tol, rkf_hmax, rkf_hmin;     /* RKF45 stepper (rapt/rkf.py) */
    double cyclotronresolution;  /* params["cyclotronresolution"] */
    double gctimestep;           /* params["GCtimestep"] (ADAPTIVE GC segments) */
    double epss, epst;           /* params["epss"], params["epst"] */
    double delta;                /* advance(delta) */
    int64_t lanes;               /* resident lanes of the persistent grid (0 = auto);
                                    fewer lanes than particles turns on work refill */
    int32_t lanes_per_particle;  /* GC with DOPRI5 in FP64: 4 (default, 0) spreads one
                                    guiding centre over a lane quad, 2 over a lane pair,
                                    8 over a lane octet; 1 = one lane per particle;
                                    3 = mixed: the first mixed_quads particles of the
                                    work order on lane quads, the rest on lane pairs;
                                    5 = mixed: the first mixed_octets on octets, the
                                    next mixed_quads on quads, the rest on pairs.
                                    Results do not depend on it */
    int32_t work_order;          /* 0: the batch's queue order (shortest output step
                                    first, or by cost after rapt_batch_set_cost); 1:
                                    input order.  Results per particle do not depend
                                    on it */
    int64_t mixed_quads;         /* lanes_per_particle 3, 5: particles on quads (0: n/8) */
    int64_t mixed_octets;        /* lanes_per_particle 5: particles on octets (0: n/64) */
} rapt_trace_desc;

/* Per-particle state, slot-major (state[slot * n + i]).  Filled by the host
 * from the tracer's last trajectory row; updated in place by rapt_batch_run so
 * a run stopped at its row capacity continues where it stopped. */
enum rapt_slot {
    RAPT_S_T = 0,                   /* r.t (integrator time) */
    RAPT_S_Y0 = 1,                  /* x, y, z, px, py, pz | x, y, z, ppar */
    RAPT_S_MASS = 7, RAPT_S_CHARGE = 8,
    RAPT_S_MU = 9,                  /* GC: magnetic moment */
    RAPT_S_V = 10,                  /* GC: speed (BrizardChan / NorthropTeller) */
    RAPT_S_DT = 11,                 /* output step (GC input; Particle: computed) */
    RAPT_S_H = 12,                  /* step size carried between integrate() calls */
    RAPT_S_TEND = 13,               /* t0 + delta of the current advance */
    RAPT_S_GM = 14,                 /* gamma*m frozen at advance start (static fields) */
    RAPT_S_TCUR = 15,               /* tracer.tcur */
    RAPT_NSLOT = 16
};
/* per-particle status */
enum rapt_status {
    RAPT_ST_PENDING = 0, RAPT_ST_DONE = 1, RAPT_ST_SWITCH = 2,
    RAPT_ST_NSTEPS = -2, RAPT_ST_HSMALL = -3, RAPT_ST_NAN = -4, RAPT_ST_GCITER = -5,
    RAPT_ST_CAP = -6, RAPT_ST_HMIN = -7, RAPT_ST_CONFIG = -10
};

typedef struct {
    int64_t naccpt, nrejct, nstep, nfcn;  /* summed over the batch */
    int64_t nswitch, nrows;
    double kernel_ms;                     /* device time of the launch (HIP events) */
} rapt_counters;

typedef struct rapt_ctx rapt_ctx;
typedef struct rapt_batch rapt_batch;

/* ---- context ---------------------------------------------------------- */
int rapt_ctx_create(int device, rapt_ctx **out);
int rapt_ctx_destroy(rapt_ctx *ctx);
const char *rapt_last_error(rapt_ctx *ctx);
/* replaces the field object handed to the tracer constructors */
int rapt_field_set(rapt_ctx *ctx, const rapt_field_desc *field);

/* Gridded field for RAPT_FIELD_GRID: replaces Grid's parsed files and its
 * RegularGridInterpolator (rapt/fields.py:513-814).  Axes x[nx], y[ny],
 * z[nz] strictly increasing (uniform spacing not required), t[nt] the
 * snapshot times (ignored when nt = 1: a static 3-D grid).  B, E: [nt][nx]
 * [ny][nz][3] in SI units, C order (E may be NULL: E = 0).  The library
 * copies everything into device memory owned by the context and keeps all
 * snapshots resident (the reference keeps a window of three).  Evaluation is
 * trilinear in space and linear in time between the bracketing snapshots;
 * outside the grid B and E are NaN (the reference raises ValueError), which
 * stops a traced particle with a non-finite or step-size status.  The
 * device copy is cell-major (the 8 corners of each cell contiguous, 8x the
 * point data) when that takes at most 64 GiB and half the free device
 * memory, else point-major; environment RAPT_GRID_POINTS=1 forces the point
 * layout.  Both give identical values.  Then rapt_field_set with model
 * RAPT_FIELD_GRID (params unused) selects it. */
int rapt_grid_set(rapt_ctx *ctx, int32_t nx, int32_t ny, int32_t nz, int32_t nt, const double *x,
                  const double *y, const double *z, const double *t, const double *B,
                  const double *E);

/* A user-written field for RAPT_FIELD_EXPR: replaces a _Field subclass that
 * overrides B and/or E in Python (rapt/fields.py:5-74, the contract of the
 * reference's "Creating new fields" example), whose methods the reference's
 * integrator calls per right-hand side.  The Python front end traces those
 * methods into register programs (rapt_amd/fieldexpr.py); the device runs
 * them in FP64, compiled into the tracer kernels or interpreted
 * (rapt_field_expr_mode).  code: (nb + ne) x 4 int32 (opcode, destination,
 * operand a, operand b) over 32 registers, 0-3 = t, x, y, z; each program
 * leaves its vector in registers 4-6; B's program is the first nb
 * instructions, E's the next ne (ne = 0: E = 0).  Opcodes: 0 register[d] =
 * consts[a], 1 a+b, 2 a-b, 3 a*b, 4 a/b, 5 -a, 6 pow(a,b), 7 sqrt, 8 exp,
 * 9 log, 10 sin, 11 cos, 12 tan, 13 tanh, 14 sinh, 15 cosh, 16 atan, 17 asin,
 * 18 acos, 19 fabs, 20 atan2(a,b), 21 a (move), 22 log10, 23 hypot(a,b),
 * 24 maximum(a,b), 25 minimum(a,b) (numpy's: a NaN operand propagates, ties
 * give b),
 * 26 sign(a) (numpy's: sign(0) = 0).
 * At most RAPT_EXPR_MAXCODE instructions and RAPT_EXPR_MAXCONST constants;
 * the library copies both into device memory owned by the context.  Then
 * rapt_field_set with model RAPT_FIELD_EXPR (params unused) selects it.
 * The field-line entry points (rapt_halfbounce, rapt_eye, rapt_bounce_*)
 * take it (compiled or interpreted as rapt_field_expr_mode says); FP32-field
 * precision does not; interpreted GuidingCenter
 * runs use one lane per particle whatever lanes_per_particle asks (results
 * are the same), compiled ones every layout. */
int rapt_field_expr_set(rapt_ctx *ctx, int32_t nb, int32_t ne, const int32_t *code,
                        int32_t nconst, const double *consts);

/* How the tracer kernels evaluate the context's user-written field.  No
 * counterpart in the reference, whose integrator calls the Python methods
 * (rapt/fields.py:5-74); both modes give the same trajectories bit for bit.
 * RAPT_EXPR_COMPILED (the default): the first Particle, GuidingCenter or
 * Adaptive run of a program compiles the one-lane tracer kernels for it at
 * run time (hipRTC, a few seconds; code cached for the process by program,
 * constants excluded) and loads them on the device, so the field is inline
 * code; rapt_trace_device_noalloc uses them only once an earlier run has
 * loaded them (loading allocates), and interprets otherwise.
 * RAPT_EXPR_INTERPRETED: the built-in kernels interpret the program.
 * The other entry points that take the field (rapt_field_eval,
 * rapt_rhs_eval, rapt_particle_gc, rapt_halfbounce, rapt_eye,
 * rapt_bounce_*) compile their kernels in the same way. */
enum rapt_expr_mode { RAPT_EXPR_INTERPRETED = 0, RAPT_EXPR_COMPILED = 1 };
int rapt_field_expr_mode(rapt_ctx *ctx, int32_t mode);

/* Compiles kernels of a program (as rapt_field_expr_set takes it) for arch
 * ("gfx950" when null or empty) and caches the code for the process, without
 * a device or a context.  kernels: bit mask, 0 = all: bits 0-3 the one-lane
 * Particle, GuidingCenter, static Tao-Chan-Brizard GuidingCenter and
 * Adaptive kernels; 4-9 GuidingCenter lane pairs, quads and octets (generic,
 * static TCB); 10-11 the mixed layout; 12-15 half-bounce path / eye, labels,
 * BounceCenter, equatorial BounceCenter; 16-18 field evaluation, right-hand
 * sides, per-row diagnostics.  The message (the compiler's log on
 * failure) goes to log (log_size bytes, NUL-terminated).  No counterpart in
 * the reference: a check, or a warm-up ahead of the first run. */
int rapt_field_expr_compile(int32_t nb, int32_t ne, const int32_t *code, const char *arch,
                            int32_t kernels, char *log, int64_t log_size);

/* Batched field evaluation on the device (the _Field methods,
 * rapt/fields.py:76-280).  what: 0 B, 1 E, 2 unitb, 3 magB, 4 gradB,
 * 5 jacobianB, 6 curvature, 7 curlb, 8 dBdt, 9 dbdt, 10 lengthscale,
 * 11 timescale.  tpos: n x 4 host array; out: n x width(what) host array. */
int rapt_field_eval(rapt_ctx *ctx, int what, int64_t n, const double *tpos, double *out);

/* The tracers' right-hand sides at n points: the Lorentz equations of
 * Particle.advance's eom (rapt/Particle.py:284-298) for mode
 * RAPT_MODE_PARTICLE, ty [n][7] = (t, x, y, z, px, py, pz), k [n][3] = (mass,
 * charge, gamma m; a static field keeps the given gamma m, as the reference
 * freezes it, a time-dependent one recomputes it), out [n][6]; the
 * guiding-centre equations GuidingCenter._TaoChanBrizardEOM /
 * _BrizardChanEOM / _NorthropTellerEOM (rapt/GuidingCenter.py:329-395, eom
 * as enum rapt_eom) for RAPT_MODE_GC, ty [n][5] = (t, x, y, z, ppar), k [n][4]
 * = (mass, charge, mu, v), out [n][4].  enforce_equatorial zeroes the z
 * components as params["enforce equatorial"] does. */
int rapt_rhs_eval(rapt_ctx *ctx, int32_t mode, int32_t eom, int32_t enforce_equatorial,
                  int64_t n, const double *ty, const double *k, double *out);

/* Half-bounce path length S_b for n guiding centres (flutils.halfbouncepath,
 * rapt/flutils.py:254-316, as GuidingCenter.bounceperiod uses it,
 * rapt/GuidingCenter.py:593-606): field-line trace, mirror points, integral.
 * tpos: n x 4 host array; Bm: mirror fields; sb: out (m); status: 1 ok, < 0
 * on failure.  Needs a dipole-type field (EarthDipole, DoubleDipole,
 * VarEarthDipole, EarthDipole + uniform), a gridded or a user-written field
 * (interpreted). */
int rapt_halfbounce(rapt_ctx *ctx, int64_t n, const double *tpos, const double *Bm,
                    double fieldlineresolution, double *sb, int32_t *status);

/* Second invariant I = int sqrt(1 - B/Bm) ds between the mirror points for n
 * guiding centres (flutils.eye, rapt/flutils.py:65-151, as
 * GuidingCenter.geteye calls it once per row, rapt/GuidingCenter.py:608-624):
 * the field-line trace, spline and mirror points of rapt_halfbounce, the
 * integral by the same end-point substitution and Gauss-Legendre panels (the
 * reference: Simpson on the trace points below 70 deg equatorial pitch,
 * QUADPACK epsrel 1e-4 above).  0 where min |B| on the line reaches Bm.
 * Arguments and status as rapt_halfbounce. */
int rapt_eye(rapt_ctx *ctx, int64_t n, const double *tpos, const double *Bm,
             double fieldlineresolution, double *eye, int32_t *status);

/* Surrogate labels, one bounce-averaged drift evaluation per query: replaces
 * BounceCenter(pos, v, 0, pa, mass, charge, field) followed by the drift of
 * its deriv (rapt/BounceCenter.py:74-115, 235-244), flutils.bounceperiod
 * (rapt/flutils.py:232-252) and flutils.gradI / eye (rapt/flutils.py:65-229).
 * query[3i..3i+2] = (L [Re], equatorial pitch angle [deg], log10 K [eV]), the
 * particle starting on the x axis at L Re in the equatorial plane;
 * out[3i..3i+2] = (bounce period [s], bounce-averaged drift angular velocity
 * about z [rad/s], +b mirror-point latitude [deg]).  Re and eV are the
 * caller's constants (rapt/__init__.py:4-10).  Static dipole-type fields
 * (models 1, 2, 5, 8), gridded and user-written ones.  status[i] as
 * rapt_halfbounce. */
int rapt_bounce_labels(rapt_ctx *ctx, int64_t n, const double *query, double mass,
                       double charge, double Re, double eV, double eyegradientstep,
                       double fieldlineresolution, double *out, int32_t *status);

/* BounceCenter.advance (rapt/BounceCenter.py:206-251) for n bounce centres:
 * dopri5 with scipy's call semantics on the drift gamma m v^2 / (q S_b B^2)
 * grad I x B (the reference's deriv, :235-244), one integrate() per output
 * row.  tpos [n][4]: trajectory[-1] in, final state out; Bm: mirror field;
 * coef: gamma m v^2 / q; dt: output step (BCtimestep x bounce period);
 * tcur: first row time; nrow: rows per centre (len(arange(tcur, tcur +
 * delta, dt))); rows: [sum nrow][4] (t, x, y, z) with the reference's row
 * times (the arange value, one step behind the state); status 1 ok, -2
 * nsteps, -3 step too small, < -19 field-line failure.  Static dipole-type
 * fields (models 1, 2, 8) and static user-written ones (interpreted). */
int rapt_bounce_center_advance(rapt_ctx *ctx, int64_t n, double *tpos, const double *Bm,
                               const double *coef, const double *dt, const double *tcur,
                               const int64_t *nrow, double *rows, int32_t *status,
                               int64_t *naccpt, double rtol, double atol,
                               double eyegradientstep, double fieldlineresolution,
                               int32_t nsteps);

/* The equatorial branch of BounceCenter.advance (isequatorial = True,
 * rapt/BounceCenter.py:237-238): dopri5 with scipy's call semantics on the
 * gradient plus inertial drift, _v_grad + _v_inert (:154-204), one
 * integrate() per output row, rows as rapt_bounce_center_advance.
 * k [n][4] = (mu, mass, charge, gamma) per centre.  The reference calls the
 * two drifts with the 3-vector (x, y, z), which every field's `t, x, y, z =
 * tpos` rejects, so its branch raises; here they take (t, x, y, z).  Any
 * static field (models 0-4 and 6-10); status 1 ok, -2 nsteps, -3 step too
 * small. */
int rapt_bounce_center_equatorial(rapt_ctx *ctx, int64_t n, double *tpos, const double *k,
                                  const double *dt, const double *tcur, const int64_t *nrow,
                                  double *rows, int32_t *status, int64_t *naccpt, double rtol,
                                  double atol, int32_t nsteps);

/* Guiding centre and first invariant of every row of Particle trajectories:
 * Particle.guidingcenter() and Particle.mu() (rapt/Particle.py:463-482), i.e.
 * utils.guidingcenter (rapt/utils.py:251-326) and utils.magnetic_moment
 * (rapt/utils.py:189-216) per row.  rows [n][7] = (t, x, y, z, px, py, pz);
 * mass, charge: per row (per_row != 0) or one value each; tol, maxiter: the
 * fixed-point iteration's (reference defaults 1e-3, 20).  out [n][6] = (X,
 * Y, Z, v_par, speed, mu); status[i] = iterations used, or RAPT_GC_NOCONV
 * (out NaN) where maxiter iterations do not reach tol (the reference prints
 * a message and returns None). */
#define RAPT_GC_NOCONV (-30)
int rapt_particle_gc(rapt_ctx *ctx, int64_t n, const double *rows, const double *mass,
                     const double *charge, int32_t per_row, double tol, int32_t maxiter,
                     double *out, int32_t *status);

/* ---- bounce-drift surrogate (no counterpart in the reference) ------------------
 * An ensemble of MLPs 3 -> 256 -> 256 -> 256 -> 3 (SiLU; BF16 weights, FP32
 * accumulation; hidden 256 x 256 layers on MFMA) that maps a query (L [Re],
 * equatorial pitch angle [deg], log10 K [eV]) to (bounce period, |bounce-
 * averaged drift angular velocity|, mirror latitude) of rapt_bounce_labels.
 * Outputs are learned as normalised log-residuals against dipole closed
 * forms (Walt's bounce and drift periods, rapt/utils.py:435-513, and the
 * exact dipole mirror latitude).  Parameters per member: RAPT_MLP_NPARAM
 * floats, W1[256][3], b1[256], W2[256][256] (out, in), b2, W3, b3, W4[3][256],
 * b4[3], in base-2 units: the plain SiLU network's parameters are
 * (ln2 W1, ln2 b1, W2, ln2 b2, W3, ln2 b3, log2(e) W4, b4), and each hidden
 * unit computes a / (1 + 2^-a) of its pre-activation a (= log2(e) silu(z)). */
#define RAPT_MLP_NPARAM 133379
typedef struct rapt_mlp rapt_mlp;

/* members (1..4) ensemble members from theta0 [members][RAPT_MLP_NPARAM]; mass [kg]
 * and eV [J] set the species and units of the dipole baselines. */
int rapt_mlp_create(rapt_ctx *ctx, int32_t members, const float *theta0, double mass,
                    double eV, rapt_mlp **out);
int rapt_mlp_destroy(rapt_mlp *m);
/* output normalisation: shift[3] then scale[3] (scale > 0) of the residuals */
int rapt_mlp_set_norm(rapt_mlp *m, const float *shift_scale);
/* Regime of the drift label: the reference's gradI differences I across the
 * field line (rapt/flutils.py:196-227) and falls back to a one-sided
 * difference where the inner line (L Re - eyegradientstep on the query's x
 * axis) has no mirror point -- a jump of O(eyegradientstep / L) in omega_d
 * across the curve B(L Re - eyegradientstep) = B_m.  With eyegradientstep > 0
 * the features of the two sides move apart (the pitch-angle feature by 2.5)
 * so each side is a smooth function for the network, and rapt_mlp_infer flags
 * queries within 1e-5 of the curve for the traced fallback; 0 turns it off.
 * Uses the context's field at the time of the call (dipole-type models 1, 2,
 * 5, 8); applies to rapt_mlp_targets and rapt_mlp_infer. */
int rapt_mlp_set_split(rapt_mlp *m, double eyegradientstep, double Re);
/* features X [n][3] and targets T [n][3] from queries [n][3] and traced labels
 * [n][3] (rapt_bounce_labels); normalised != 0 applies the shift/scale. */
int rapt_mlp_targets(rapt_mlp *m, int64_t n, const float *query, const double *labels,
                     int normalised, float *X, float *T);
/* steps Adam steps (all members together) on mean squared error, batch rows
 * per step taken from order[steps * batch] (indices into X/T); learning rate
 * cosine-decayed from lr0 to lr1 over the call (lr0 = 0: gradients only).
 * loss_out [members]: last step's loss; grad_out [members][NPARAM]: last
 * step's gradients (either may be NULL).  batch: a multiple of 64. */
int rapt_mlp_train(rapt_mlp *m, int64_t n, const float *X, const float *T, int32_t batch,
                   int32_t steps, const int64_t *order, float lr0, float lr1,
                   double *loss_out, float *grad_out);
int rapt_mlp_get_params(rapt_mlp *m, float *theta);
/* training-path forward of features X [n][3]: Y [members][n][3] normalised */
int rapt_mlp_forward(rapt_mlp *m, int64_t n, const float *X, float *Y);
/* fused ensemble inference: out [n][3] physical values (ensemble mean of the
 * normalised outputs), flag[i] = 1 where a member deviates from the
 * ensemble mean by more than thresh (relative) on any output (the caller traces those: rapt_bounce_labels);
 * raw [members][n][3] optional; kernel_ms: inference kernel time. */
int rapt_mlp_infer(rapt_mlp *m, int64_t n, const float *query, float thresh, float *out,
                   int32_t *flag, float *raw, double *kernel_ms);

/* ---- device-resident batch (replaces one scipy ode loop per particle) -- */
int rapt_batch_create(rapt_ctx *ctx, int64_t n, rapt_batch **out);
int rapt_batch_destroy(rapt_batch *b);
/* state: RAPT_NSLOT x n slot-major host array; flags: per-particle int32
 * (bit 0: advance already started; bits 1-2: adaptive segment mode). */
int rapt_batch_upload(rapt_batch *b, const double *state, const int32_t *flags);
/* Work order by cost: with work_order 0, runs take particles in descending
 * cost[i] (e.g. each particle's accepted steps in an earlier or trial
 * launch) instead of by output step.  The next upload restores the output-
 * step order.  No reference counterpart: scheduling of the batched tracer.
 * Results per particle do not depend on the order. */
int rapt_batch_set_cost(rapt_batch *b, const double *cost);
/* Restore the state and flags of the last upload (device to device). */
int rapt_batch_reset(rapt_batch *b);
/* Row storage: row_cap[i] rows of `width` doubles for particle i (width 7
 * PARTICLE, 5 GC, 8 ADAPTIVE).  Resets the per-particle row counters. */
int rapt_batch_set_rows(rapt_batch *b, int32_t width, const int64_t *row_cap);
/* Run the advance on the device; inputs stay in HBM.  counters may be NULL. */
int rapt_batch_run(rapt_batch *b, const rapt_trace_desc *desc, rapt_counters *counters);
/* Copy results to the host (any pointer may be NULL). */
int rapt_batch_download(rapt_batch *b, double *state, int32_t *flags, int32_t *status,
                        int64_t *nrows, double *rows, int64_t *naccpt);

/* Output epochs: the batch runs in epochs of at most rows_per_epoch rows per
 * particle, each epoch one launch of a captured hipGraph (counter reset,
 * tracer kernel, copies of the epoch's rows, row counts and statuses to pinned
 * host buffers).  Particles that stop on the full epoch buffer continue
 * exactly in the next epoch; finished ones are skipped.  Device row memory is
 * n x rows_per_epoch rows however long the run.  begin captures the graph for
 * the descriptor (the batch's state, flags and work order as uploaded);
 * next runs one epoch and returns pointers to the pinned buffers (valid until
 * the next call): rows [n][rows_per_epoch][width], nrows [n] (rows written by
 * this epoch), status [n], running = particles still to continue; end frees
 * the epoch resources (also done by rapt_batch_destroy). */
int rapt_batch_epochs_begin(rapt_batch *b, const rapt_trace_desc *desc, int32_t rows_per_epoch);
int rapt_batch_epoch_next(rapt_batch *b, const double **rows, const int64_t **nrows,
                          const int32_t **status, int32_t *running, rapt_counters *counters);
int rapt_batch_epochs_end(rapt_batch *b);

/* Allocation-free advance on caller-owned device buffers.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels as rapt_batch_run; every pointer except d and
 * cnt is device memory of the context's device: state [RAPT_NSLOT][n]
 * (slot-major, updated in place), flags [n] (in/out), row_off/row_cap [n]
 * (rows of particle i at rows + width * row_off[i], at most row_cap[i] of
 * them; width 7/5/8 by mode), rows, nrows/status/naccpt [n] (out), order [n]
 * (optional work order, particle indices; NULL: input order).  The launch
 * calls no hipMalloc/hipFree: its work-queue and step counters live in a
 * 64-byte scratch array that each device holds until the process ends, so
 * calls that overlap on one device race (serialize them); each call clears
 * the counters first.  Runs on the context's stream and returns when the
 * advance is done; cnt (host, may be NULL) as rapt_batch_run. */
int rapt_trace_device_noalloc(rapt_ctx *ctx, const rapt_trace_desc *desc, int64_t n,
                              double *state, int32_t *flags, const int64_t *row_off,
                              const int64_t *row_cap, double *rows, int64_t *nrows,
                              int32_t *status, int64_t *naccpt, const int64_t *order,
                              rapt_counters *counters);

/* One-shot convenience: upload, run, download (host buffers). */
int rapt_trace(rapt_ctx *ctx, const rapt_trace_desc *desc, int64_t n, double *state,
               int32_t *flags, const int64_t *row_cap, double *rows, int64_t *nrows,
               int32_t *status, rapt_counters *counters);

#ifdef __cplusplus
}
#endif
#endif
