"""ctypes front end of the oracle (oracle/librapt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RAPT_ORACLE_LIB names another build of the oracle (e.g. one with a planted
# mistake, to see that the pins fail)
_LIB_PATH = os.environ.get("RAPT_ORACLE_LIB") or os.path.join(_HERE, "librapt_oracle.so")

ZERO, EARTH_DIPOLE, DOUBLE_DIPOLE, UNIFORM_BZ, UNIFORM_CROSSED_EB, VAR_EARTH_DIPOLE, \
    PARABOLIC, CURRENT_SHEET, DIPOLE_UNIFORM = range(9)
EXPR = 10
EVAL = dict(B=0, E=1, unitb=2, magB=3, gradB=4, jacobianB=5, curvature=6, curlb=7,
            dBdt=8, dbdt=9, lengthscale=10, timescale=11)
EOM = dict(TaoChanBrizardEOM=0, BrizardChanEOM=1, NorthropTellerEOM=2)
DOP853, DOPRI5, RKF45 = 0, 1, 2


class Field(C.Structure):
    _fields_ = [("model", C.c_int32), ("is_static", C.c_int32), ("gstep", C.c_double),
                ("tstep", C.c_double), ("p", C.c_double * 16), ("code", C.c_void_p),
                ("consts", C.c_void_p), ("nb", C.c_int32), ("ne", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double),
                ("cyclotronresolution", C.c_double), ("bounceresolution", C.c_double),
                ("gctimestep", C.c_double), ("epss", C.c_double), ("epst", C.c_double),
                ("enforce_equatorial", C.c_int32), ("nsteps", C.c_int32),
                ("stepper", C.c_int32), ("check_adiabaticity", C.c_int32),
                ("safety", C.c_double), ("dfactor", C.c_double), ("ifactor", C.c_double),
                ("beta", C.c_double), ("max_step", C.c_double), ("first_step", C.c_double),
                ("rkf_tol", C.c_double), ("rkf_hmax", C.c_double), ("rkf_hmin", C.c_double)]


class Counters(C.Structure):
    _fields_ = [("naccpt", C.c_int64), ("nrejct", C.c_int64), ("nstep", C.c_int64),
                ("nfcn", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_field_eval_batch.argtypes = [C.POINTER(Field), C.c_int, C.c_int64, dp, dp]
        L.orc_field_eval_batch.restype = C.c_int
        L.orc_test_problem_run.argtypes = [C.c_int, C.c_int, C.POINTER(Params), C.c_int,
                                           C.c_double, dp, C.c_double, C.c_int64, dp,
                                           C.POINTER(Counters), C.POINTER(C.c_int32)]
        L.orc_test_problem_run.restype = C.c_int
        L.orc_particle_advance.argtypes = [C.POINTER(Field), C.POINTER(Params), dp, C.c_double,
                                           C.c_double, C.c_double, dp, C.c_int64, dp,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                           C.POINTER(Counters)]
        L.orc_particle_advance.restype = C.c_int64
        L.orc_gc_advance.argtypes = [C.POINTER(Field), C.POINTER(Params), C.c_int, dp,
                                     C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, dp, C.c_int64, dp,
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(Counters)]
        L.orc_gc_advance.restype = C.c_int64
        L.orc_adaptive_run.argtypes = [C.POINTER(Field), C.POINTER(Params), dp, dp, C.c_double,
                                       C.c_double, C.c_double, C.c_double, dp, C.c_int64,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.POINTER(Counters)]
        L.orc_adaptive_run.restype = C.c_int64
        L.orc_bc_equatorial.argtypes = [C.POINTER(Field), C.POINTER(Params), dp, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int64, dp, C.POINTER(C.c_int32),
                                        C.POINTER(Counters)]
        L.orc_bc_equatorial.restype = C.c_int64
        L.orc_bc_equatorial_drifts.argtypes = [C.POINTER(Field), dp, C.c_double, C.c_double,
                                               C.c_double, C.c_double, dp]
        L.orc_bc_equatorial_drifts.restype = None
        L.orc_adaptive_continue.argtypes = [C.POINTER(Field), C.POINTER(Params), C.c_int, dp,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, dp, C.c_int64, C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32), C.POINTER(Counters)]
        L.orc_adaptive_continue.restype = C.c_int64
        L.orc_particle_isadiabatic.argtypes = [C.POINTER(Field), C.POINTER(Params), dp,
                                               C.c_double, C.c_double]
        L.orc_particle_isadiabatic.restype = C.c_int
        L.orc_gc_isadiabatic.argtypes = [C.POINTER(Field), C.POINTER(Params), dp, C.c_double,
                                         C.c_double, C.c_double]
        L.orc_gc_isadiabatic.restype = C.c_int
        L.orc_adaptive_to_gc.argtypes = [C.POINTER(Field), dp, C.c_double, C.c_double, dp, dp,
                                         dp]
        L.orc_adaptive_to_gc.restype = C.c_int
        L.orc_adaptive_to_particle.argtypes = [C.POINTER(Field), dp, C.c_double, C.c_double,
                                               C.c_double, dp]
        L.orc_adaptive_to_particle.restype = None
        L.orc_rkf.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, dp,
                              C.c_double, C.c_double, C.c_double, dp, dp, C.c_int64,
                              C.POINTER(C.c_int64)]
        L.orc_rkf.restype = C.c_int
        for nm in ("orc_cyclotron_period", "orc_cyclotron_radius"):
            getattr(L, nm).argtypes = [C.POINTER(Field), C.c_double, dp, dp, C.c_double,
                                       C.c_double]
            getattr(L, nm).restype = C.c_double
        L.orc_cyclotron_period2.argtypes = [C.POINTER(Field), C.c_double, dp, C.c_double,
                                            C.c_double, C.c_double]
        L.orc_cyclotron_period2.restype = C.c_double
        L.orc_cyclotron_radius2.argtypes = [C.POINTER(Field), C.c_double, dp, C.c_double,
                                            C.c_double, C.c_double, C.c_double]
        L.orc_cyclotron_radius2.restype = C.c_double
        L.orc_magnetic_moment.argtypes = [C.POINTER(Field), C.c_double, dp, C.c_double,
                                          C.c_double, C.c_double]
        L.orc_magnetic_moment.restype = C.c_double
        L.orc_guidingcenter.argtypes = [C.POINTER(Field), C.c_double, dp, dp, C.c_double,
                                        C.c_double, C.c_double, C.c_int, dp, dp, dp]
        L.orc_guidingcenter.restype = C.c_int
        L.orc_gctofp.argtypes = [C.POINTER(Field), C.c_double, dp, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_gctofp.restype = None
        ip = C.POINTER(C.c_int64)
        i32p = C.POINTER(C.c_int32)
        L.orc_batch_particle.argtypes = [C.POINTER(Field), C.POINTER(Params), C.c_int64, dp, dp,
                                         dp, C.c_double, dp, ip, ip, ip, i32p,
                                         C.POINTER(Counters), C.c_int]
        L.orc_batch_gc.argtypes = [C.POINTER(Field), C.POINTER(Params), C.c_int, C.c_int64, dp,
                                   dp, dp, dp, dp, dp, C.c_double, dp, ip, ip, ip, i32p,
                                   C.POINTER(Counters), C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_field(model, params=(), is_static=True, gstep=1e-6, tstep=1e-3):
    f = Field()
    f.model = model
    f.is_static = 1 if is_static else 0
    f.gstep = gstep
    f.tstep = tstep
    for i, v in enumerate(params):
        f.p[i] = float(v)
    return f


def field_from_desc(desc):
    """desc: dict with model, params, static, gstep, tstep (the product's field
    descriptor), and for a user-written field its traced program ``expr``
    (code [n][4] int32, consts, nb, ne)."""
    f = make_field(desc["model"], desc["params"], desc["static"], desc["gstep"],
                   desc["tstep"])
    prog = desc.get("expr")
    if prog is not None:
        code = np.ascontiguousarray(prog.code, dtype=np.int32)
        consts = np.ascontiguousarray(prog.consts, dtype=np.float64)
        if consts.size == 0:
            consts = np.zeros(1)
        f._keep = (code, consts)  # the struct holds raw pointers into these
        f.code, f.consts = code.ctypes.data, consts.ctypes.data
        f.nb, f.ne = prog.nb, prog.ne
    return f


def make_params(rtol=1.49012e-8, atol=1.49012e-8, cyclotronresolution=10, bounceresolution=10,
                gctimestep=0.0, epss=5e-2, epst=5e-2, enforce_equatorial=False, nsteps=500,
                stepper=DOP853, check_adiabaticity=False, safety=0.9, dfactor=0.3,
                ifactor=6.0, beta=0.1, max_step=0.0, first_step=0.0, rkf_tol=1e-3,
                rkf_hmax=1.0, rkf_hmin=1e-6):
    p = Params()
    p.rtol, p.atol = rtol, atol
    p.cyclotronresolution = cyclotronresolution
    p.bounceresolution = bounceresolution
    p.gctimestep = gctimestep
    p.epss, p.epst = epss, epst
    p.enforce_equatorial = int(bool(enforce_equatorial))
    p.nsteps = nsteps
    p.stepper = stepper
    p.check_adiabaticity = int(bool(check_adiabaticity))
    p.safety, p.dfactor, p.ifactor, p.beta = safety, dfactor, ifactor, beta
    p.max_step, p.first_step = max_step, first_step
    p.rkf_tol, p.rkf_hmax, p.rkf_hmin = rkf_tol, rkf_hmax, rkf_hmin
    return p


def particle_params(**kw):
    """scipy dop853 as set by Particle.advance (rapt/Particle.py:300)."""
    d = dict(stepper=DOP853, safety=0.9, dfactor=0.3, ifactor=6.0, beta=0.1)
    d.update(kw)
    return make_params(**d)


def gc_params(**kw):
    """scipy dopri5 as set by GuidingCenter.advance (rapt/GuidingCenter.py:450)."""
    d = dict(stepper=DOPRI5, safety=0.9, dfactor=0.2, ifactor=10.0, beta=0.0)
    d.update(kw)
    return make_params(**d)


def field_eval(field, what, tpos):
    tpos = np.ascontiguousarray(np.atleast_2d(tpos), dtype=np.float64)
    n = tpos.shape[0]
    out = np.zeros(n * 9)
    w = lib().orc_field_eval_batch(C.byref(field), EVAL[what], n, _dp(tpos), _dp(out))
    if w < 0:
        raise ValueError(what)
    return out[:n * w].reshape(n, w)


def test_problem(problem, stepper, params, y0, t0, dt, ncalls):
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    n = y0.size
    rows = np.zeros((ncalls, n + 1))
    cnt = Counters()
    ist = C.c_int32()
    k = lib().orc_test_problem_run(problem, stepper, C.byref(params), n, t0, _dp(y0), dt,
                                   ncalls, _dp(rows), C.byref(cnt), C.byref(ist))
    return rows[:k], cnt, ist.value


def particle_advance(field, params, last, mass, charge, delta, cap=200000):
    last = np.ascontiguousarray(last, dtype=np.float64)
    rows = np.zeros((cap, 7))
    tcur = C.c_double()
    st = C.c_int32()
    sw = C.c_int32()
    cnt = Counters()
    n = lib().orc_particle_advance(C.byref(field), C.byref(params), _dp(last), mass, charge,
                                   delta, _dp(rows), cap, C.byref(tcur), C.byref(st),
                                   C.byref(sw), C.byref(cnt))
    return rows[:n].copy(), dict(tcur=tcur.value, status=st.value, switched=sw.value,
                                 naccpt=cnt.naccpt, nrejct=cnt.nrejct, nfcn=cnt.nfcn)


def gc_advance(field, params, eom, last, mass, charge, mu, v, dt, delta, cap=400000):
    last = np.ascontiguousarray(last, dtype=np.float64)
    rows = np.zeros((cap, 5))
    tcur = C.c_double()
    st = C.c_int32()
    sw = C.c_int32()
    cnt = Counters()
    e = EOM[eom] if isinstance(eom, str) else int(eom)
    n = lib().orc_gc_advance(C.byref(field), C.byref(params), e, _dp(last), mass, charge, mu,
                             v, dt, delta, _dp(rows), cap, C.byref(tcur), C.byref(st),
                             C.byref(sw), C.byref(cnt))
    return rows[:n].copy(), dict(tcur=tcur.value, status=st.value, switched=sw.value,
                                 naccpt=cnt.naccpt, nrejct=cnt.nrejct, nfcn=cnt.nfcn)


def bc_equatorial_drifts(field, tpos, mu, mass, charge, gamma):
    """(_v_grad, _v_inert) of BounceCenter at one (t, x, y, z)."""
    tp = np.ascontiguousarray(tpos, dtype=np.float64)
    out = np.zeros(6)
    lib().orc_bc_equatorial_drifts(C.byref(field), _dp(tp), mu, mass, charge, gamma, _dp(out))
    return out[:3].copy(), out[3:].copy()


def bc_equatorial(field, params, tpos, mu, mass, charge, gamma, dt, tcur, nrow):
    """BounceCenter.advance with isequatorial: rows (nrow, 4), end state, info."""
    tp = np.ascontiguousarray(tpos, dtype=np.float64).copy()
    rows = np.zeros((max(int(nrow), 1), 4))
    st = C.c_int32()
    cnt = Counters()
    n = lib().orc_bc_equatorial(C.byref(field), C.byref(params), _dp(tp), mu, mass, charge,
                                gamma, dt, tcur, int(nrow), _dp(rows), C.byref(st),
                                C.byref(cnt))
    return rows[:n].copy(), tp, dict(status=st.value, naccpt=cnt.naccpt, nfcn=cnt.nfcn)


def adaptive_run(field, params, pos, vel, t0, mass, charge, delta, cap=400000):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    rows = np.zeros((cap, 8))
    nsw = C.c_int32()
    st = C.c_int32()
    cnt = Counters()
    n = lib().orc_adaptive_run(C.byref(field), C.byref(params), _dp(pos), _dp(vel), t0, mass,
                               charge, delta, _dp(rows), cap, C.byref(nsw), C.byref(st),
                               C.byref(cnt))
    return rows[:n].copy(), dict(nswitch=nsw.value, status=st.value, naccpt=cnt.naccpt,
                                 nrejct=cnt.nrejct, nfcn=cnt.nfcn)


def adaptive_continue(field, params, mode, last, mu, v, mass, charge, delta, cap=400000):
    """A further Adaptive.advance(delta) from the last tracer (mode 0 Particle
    with last row (t, x, y, z, px, py, pz); mode 1 GuidingCenter with last row
    (t, x, y, z, ppar), moment mu, speed v)."""
    last = np.ascontiguousarray(last, dtype=np.float64)
    rows = np.zeros((cap, 8))
    nsw = C.c_int32()
    st = C.c_int32()
    cnt = Counters()
    n = lib().orc_adaptive_continue(C.byref(field), C.byref(params), int(mode), _dp(last),
                                    float(mu), float(v), mass, charge, delta, _dp(rows), cap,
                                    C.byref(nsw), C.byref(st), C.byref(cnt))
    return rows[:n].copy(), dict(nswitch=nsw.value, status=st.value, naccpt=cnt.naccpt,
                                 nrejct=cnt.nrejct, nfcn=cnt.nfcn)


def particle_isadiabatic(field, params, row7, mass, charge):
    return bool(lib().orc_particle_isadiabatic(C.byref(field), C.byref(params),
                                               _dp(np.ascontiguousarray(row7, float)), mass,
                                               charge))


def gc_isadiabatic(field, params, row5, mass, charge, mu):
    return bool(lib().orc_gc_isadiabatic(C.byref(field), C.byref(params),
                                         _dp(np.ascontiguousarray(row5, float)), mass, charge,
                                         mu))


def adaptive_to_gc(field, row7, mass, charge):
    """GuidingCenter.init(Particle): (row5, mu, v), or None without convergence."""
    row5 = np.zeros(5)
    mu, v = C.c_double(), C.c_double()
    rc = lib().orc_adaptive_to_gc(C.byref(field), _dp(np.ascontiguousarray(row7, float)), mass,
                                  charge, _dp(row5), C.byref(mu), C.byref(v))
    if rc != 1:
        return None
    return row5, mu.value, v.value


def adaptive_to_particle(field, row5, mass, charge, mu):
    """Particle.init(GuidingCenter): row7."""
    row7 = np.zeros(7)
    lib().orc_adaptive_to_particle(C.byref(field), _dp(np.ascontiguousarray(row5, float)), mass,
                                   charge, mu, _dp(row7))
    return row7


def guidingcenter(field, t, r, v, mass, charge, tol=1e-3, maxiter=20):
    r = np.ascontiguousarray(r, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    gc = np.zeros(3)
    vp = C.c_double()
    spd = C.c_double()
    rc = lib().orc_guidingcenter(C.byref(field), t, _dp(r), _dp(v), mass, charge, tol, maxiter,
                                 _dp(gc), C.byref(vp), C.byref(spd))
    if rc != 1:
        return None
    return gc, vp.value, spd.value


def gctofp(field, t, R, vp, speed, mass, charge, gyrophase=0.0):
    R = np.ascontiguousarray(R, dtype=np.float64)
    pos = np.zeros(3)
    vel = np.zeros(3)
    lib().orc_gctofp(C.byref(field), t, _dp(R), vp, speed, mass, charge, gyrophase, _dp(pos),
                     _dp(vel))
    return pos, vel


def cyclotron_period(field, t, pos, vel, mass, charge):
    return lib().orc_cyclotron_period(C.byref(field), t, _dp(np.asarray(pos, float)),
                                      _dp(np.asarray(vel, float)), mass, charge)


def cyclotron_radius(field, t, pos, vel, mass, charge):
    return lib().orc_cyclotron_radius(C.byref(field), t, _dp(np.asarray(pos, float)),
                                      _dp(np.asarray(vel, float)), mass, charge)


def magnetic_moment(field, t, pos, vpar, v, mass):
    return lib().orc_magnetic_moment(C.byref(field), t, _dp(np.asarray(pos, float)), vpar, v,
                                     mass)


def batch_particle(field, params, last7, mass, charge, delta, row_cap, nthreads=0):
    """OpenMP Particle.advance over a batch; returns (rows list, counters, status)."""
    n = last7.shape[0]
    last7 = np.ascontiguousarray(last7, dtype=np.float64)
    mass = np.ascontiguousarray(np.broadcast_to(mass, (n,)), dtype=np.float64)
    charge = np.ascontiguousarray(np.broadcast_to(charge, (n,)), dtype=np.float64)
    row_cap = np.ascontiguousarray(row_cap, dtype=np.int64)
    row_off = np.zeros(n, np.int64)
    row_off[1:] = np.cumsum(row_cap)[:-1]
    rows = np.zeros((int(row_cap.sum()), 7))
    nrows = np.zeros(n, np.int64)
    status = np.zeros(n, np.int32)
    cnt = Counters()
    lib().orc_batch_particle(C.byref(field), C.byref(params), n, _dp(last7), _dp(mass),
                             _dp(charge), delta, _dp(rows),
                             row_off.ctypes.data_as(C.POINTER(C.c_int64)),
                             row_cap.ctypes.data_as(C.POINTER(C.c_int64)),
                             nrows.ctypes.data_as(C.POINTER(C.c_int64)),
                             status.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(cnt),
                             nthreads)
    return rows, row_off, nrows, status, cnt


def batch_gc(field, params, eom, last5, mass, charge, mu, v, dt, delta, row_cap, nthreads=0):
    n = last5.shape[0]
    f64 = lambda a: np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
    last5 = np.ascontiguousarray(last5, dtype=np.float64)
    mass, charge, mu, v, dt = f64(mass), f64(charge), f64(mu), f64(v), f64(dt)
    row_cap = np.ascontiguousarray(row_cap, dtype=np.int64)
    row_off = np.zeros(n, np.int64)
    row_off[1:] = np.cumsum(row_cap)[:-1]
    rows = np.zeros((int(row_cap.sum()), 5))
    nrows = np.zeros(n, np.int64)
    status = np.zeros(n, np.int32)
    cnt = Counters()
    e = EOM[eom] if isinstance(eom, str) else int(eom)
    lib().orc_batch_gc(C.byref(field), C.byref(params), e, n, _dp(last5), _dp(mass),
                       _dp(charge), _dp(mu), _dp(v), _dp(dt), delta, _dp(rows),
                       row_off.ctypes.data_as(C.POINTER(C.c_int64)),
                       row_cap.ctypes.data_as(C.POINTER(C.c_int64)),
                       nrows.ctypes.data_as(C.POINTER(C.c_int64)),
                       status.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(cnt), nthreads)
    return rows, row_off, nrows, status, cnt
