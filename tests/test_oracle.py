"""Pins of the CPU oracle (oracle/rapt_oracle.c) before it is trusted as the checker.

- golden fixtures written by running the reference (tests/golden/make_golden.py):
  field values and derivatives, full Particle.advance and GuidingCenter.advance
  trajectories, utils conversions, rkf;
- scipy's own DOPRI5/DOP853 run as a black box on the same problems;
- closed forms the mathematics fixes (dipole field, exact gyration, E x B drift,
  central-difference truncation, energy conservation).
"""
import ctypes as C

import numpy as np
import pytest
from scipy.integrate import ode

import rapt_amd.fields as F
from rapt_amd import constants as K

FIELDS = {
    "EarthDipole": lambda: F.EarthDipole(),
    "DoubleDipole": lambda: F.DoubleDipole(),
    "DoubleDipole_k2": lambda: F.DoubleDipole(distance=15 * K.Re, imagestrength=2),
    "UniformBz": lambda: F.UniformBz(Bz=2e-7),
    "UniformCrossedEB": lambda: F.UniformCrossedEB(Ey=1e-3, Bz=1e-7),
    "VarEarthDipole": lambda: F.VarEarthDipole(amp=0.2, period=7.0),
    "Parabolic": lambda: F.Parabolic(B0=10.0, Bn=1.0, d=0.2),
    "CurrentSheet": lambda: F.CurrentSheet(),
    "DipolePlusUniform": lambda: F.DipolePlusUniform(),
}
WHAT = ["B", "E", "unitb", "magB", "gradB", "jacobianB", "curvature", "curlb", "dBdt", "dbdt",
        "lengthscale", "timescale"]


@pytest.mark.parametrize("name", sorted(FIELDS))
def test_fields_match_reference(oracle, golden, name):
    g = golden("fields")
    f = oracle.field_from_desc(FIELDS[name]().descriptor())
    tp = g[name + "/tpos"]
    for w in WHAT:
        ref = g[name + "/" + w]
        got = oracle.field_eval(f, w, tp)
        if ref.shape != got.shape:  # static dbdt is the scalar 0 in the reference
            assert np.all(ref == 0) and np.all(got == 0), w
            continue
        if name == "CurrentSheet":
            # numpy's tanh differs from libm's by an ulp; finite differences
            # amplify it by lambda/gradientstepsize
            np.testing.assert_allclose(got, ref, rtol=2e-8, atol=0, err_msg=w)
        else:
            np.testing.assert_array_equal(got, ref, err_msg="%s.%s" % (name, w))


def _run_particle(oracle, desc, row0, mass, charge, delta, **kw):
    prm = oracle.particle_params(**kw)
    rows, info = oracle.particle_advance(oracle.field_from_desc(desc), prm, row0, mass, charge,
                                         delta)
    return np.vstack([row0[None, :], rows]), info


def test_particle_advance_config1_bitexact(oracle, golden):
    """Six 1 MeV protons of the dipole workload, full advance(10) at
    cyclotronresolution 20: identical to the reference, row by row."""
    P = golden("particles")
    lens = P["c1/len"]
    off = np.r_[0, np.cumsum(lens)]
    desc = F.EarthDipole().descriptor()
    for i in range(len(lens)):
        ref = P["c1/traj"][off[i]:off[i + 1]]
        got, info = _run_particle(oracle, desc, ref[0], K.m_pr, K.e, 10.0,
                                  cyclotronresolution=20)
        np.testing.assert_array_equal(got, ref)
        assert info["tcur"] == P["c1/tcur"][i]


@pytest.mark.parametrize("name,field,eq,tol", [
    ("crossedEB", lambda: F.UniformCrossedEB(Ey=1e-3, Bz=1e-7), False, 0),
    ("vardipole", lambda: F.VarEarthDipole(amp=0.2, period=7.0), False, 0),
    ("equatorial", lambda: F.EarthDipole(), True, 0),
    ("dipunif_e", lambda: F.DipolePlusUniform(), False, 0),
    ("dipunif_p", lambda: F.DipolePlusUniform(), False, 0),
    ("sheet", lambda: F.CurrentSheet(), False, 1e-12),
])
def test_particle_advance_fields(oracle, golden, name, field, eq, tol):
    P = golden("particles")
    ref = P[name + "/traj"]
    mass, q, delta = P[name + "/mq"]
    got, info = _run_particle(oracle, field().descriptor(), ref[0], mass, q, delta,
                              enforce_equatorial=eq)
    if tol == 0:
        np.testing.assert_array_equal(got, ref)
        assert info["tcur"] == P[name + "/tcur"][0]
    else:
        np.testing.assert_allclose(got, ref, rtol=tol, atol=0)


def test_gc_advance_config2_bitexact(oracle, golden):
    """Six DoubleDipole electrons, GuidingCenter.advance(3) with the
    reference's bounce-period output step: identical to the reference."""
    G = golden("gcs")
    lens = G["c2/len"]
    off = np.r_[0, np.cumsum(lens)]
    f = oracle.field_from_desc(F.DoubleDipole().descriptor())
    prm = oracle.gc_params(rtol=1e-8, atol=1e-8)
    for i in range(len(lens)):
        ref = G["c2/traj"][off[i]:off[i + 1]]
        rows, info = oracle.gc_advance(f, prm, "TaoChanBrizardEOM", ref[0], K.m_el, -K.e,
                                       G["c2/mu"][i], G["c2/v"][i], G["c2/dt"][i], 3.0)
        np.testing.assert_array_equal(np.vstack([ref[:1], rows]), ref)
        assert info["tcur"] == G["c2/tcur"][i]


@pytest.mark.parametrize("name,field,eom,tol", [
    ("bc", lambda: F.DoubleDipole(), "BrizardChanEOM", 0),
    ("nt", lambda: F.DoubleDipole(), "NorthropTellerEOM", 0),
    ("earth", lambda: F.EarthDipole(), "TaoChanBrizardEOM", 0),
    # numpy BLAS rounds one RHS in ~3000 differently for this field; 1e-7 is
    # the adaptive-run parity bar
    ("dipunif", lambda: F.DipolePlusUniform(), "TaoChanBrizardEOM", 1e-7),
    ("sheet", lambda: F.CurrentSheet(), "TaoChanBrizardEOM", 1e-12),
])
def test_gc_advance_variants(oracle, golden, name, field, eom, tol):
    G = golden("gcs")
    ref = G[name + "/traj"]
    v, pa, mu, dt, delta, tcur = G[name + "/in"]
    tolr = (1e-8, 1e-8) if name in ("bc", "nt") else (1.49012e-8, 1.49012e-8)
    m, q = (K.m_pr, K.e) if name == "sheet" else (K.m_el, -K.e)
    rows, info = oracle.gc_advance(oracle.field_from_desc(field().descriptor()),
                                   oracle.gc_params(rtol=tolr[0], atol=tolr[1]), eom, ref[0],
                                   m, q, mu, v, dt, delta)
    got = np.vstack([ref[:1], rows])
    assert got.shape == ref.shape
    if tol == 0:
        np.testing.assert_array_equal(got, ref)
    else:
        scale = np.max(np.abs(ref), axis=0)
        assert np.max(np.abs(got - ref) / scale) < tol


def _harm(t, y):
    return np.array([y[1], -y[0]])


def _aren(t, y):
    mu = 0.012277471
    mup = 1.0 - mu
    d1 = ((y[0] + mu) * (y[0] + mu) + y[1] * y[1]) ** 1.5
    d2 = ((y[0] - mup) * (y[0] - mup) + y[1] * y[1]) ** 1.5
    return np.array([y[2], y[3], y[0] + 2.0 * y[3] - mup * (y[0] + mu) / d1 - mu * (y[0] - mup) / d2,
                     y[1] - 2.0 * y[2] - mup * y[1] / d1 - mu * y[1] / d2])


@pytest.mark.parametrize("stepper", ["dop853", "dopri5"])
@pytest.mark.parametrize("tol", [1e-6, 1.49012e-8, 1e-10])
def test_integrators_match_scipy_bitexact(oracle, stepper, tol):
    """Hairer's DOP853/DOPRI5 with scipy.integrate.ode's call semantics
    (step carried between integrate() calls), against scipy itself."""
    kw = dict(safety=0.9, dfactor=0.3, ifactor=6.0, beta=0.1) if stepper == "dop853" else \
        dict(safety=0.9, dfactor=0.2, ifactor=10.0, beta=0.0)
    st = oracle.DOP853 if stepper == "dop853" else oracle.DOPRI5
    for prob, fn, y0, dt, nc in [(0, _harm, [1.0, 0.0], 0.7, 30),
                                 (1, _aren, [0.994, 0.0, 0.0, -2.00158510637908252240537862224],
                                  0.5, 34)]:
        r = ode(fn).set_integrator(stepper, rtol=tol, atol=tol, **kw)
        r.set_initial_value(y0, 0.0)
        ref = []
        for _ in range(nc):
            y = r.integrate(r.t + dt)
            ref.append([r.t] + list(y))
        rows, cnt, ist = oracle.test_problem(prob, st, oracle.make_params(rtol=tol, atol=tol,
                                                                          stepper=st, **kw),
                                             y0, 0.0, dt, nc)
        np.testing.assert_array_equal(rows, np.array(ref))


def test_integrator_nsteps_limit_matches_scipy(oracle):
    """At most nsteps+1 steps per integrate() call, then istate -2."""
    import warnings
    for stepper, st in (("dop853", oracle.DOP853), ("dopri5", oracle.DOPRI5)):
        r = ode(_harm).set_integrator(stepper, rtol=1e-12, atol=1e-12, nsteps=5)
        r.set_initial_value([1.0, 0.0], 0.0)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            y = r.integrate(10.0)
        rows, cnt, ist = oracle.test_problem(0, st, oracle.make_params(
            rtol=1e-12, atol=1e-12, stepper=st, nsteps=5,
            **(dict(dfactor=0.3, ifactor=6.0, beta=0.0) if st == oracle.DOP853 else
               dict(dfactor=0.2, ifactor=10.0, beta=0.0))), [1.0, 0.0], 0.0, 10.0, 1)
        assert ist == -2 and r._integrator.istate == -2
        assert rows[0, 0] == r.t
        np.testing.assert_array_equal(rows[0, 1:], y)
        assert cnt.nstep == 6


RKF_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                     C.POINTER(C.c_double))


def test_rkf_matches_reference(oracle, golden):
    """rapt/rkf.py on y'' = -y, including its hmin RuntimeError."""
    R = golden("rkf")

    @RKF_CB
    def cb(ctx, t, y, out):
        out[0] = y[1]
        out[1] = -y[0]
    L = oracle.lib()
    for j in range(4):
        tol, hmax, hmin, raised = R["osc%d/cfg" % j]
        T, X = R["osc%d/T" % j], R["osc%d/X" % j]
        cap = 100000
        tout = np.zeros(cap)
        xout = np.zeros((cap, 2))
        npts = C.c_int64()
        x0 = np.array([1.0, 0.0])
        rc = L.orc_rkf(C.cast(cb, C.c_void_p), None, 2, 0.0, 3.0, oracle._dp(x0), tol, hmax,
                       hmin, oracle._dp(tout), oracle._dp(xout), cap, C.byref(npts))
        if raised:
            assert rc == -7
        else:
            assert rc == 1
            np.testing.assert_array_equal(tout[:npts.value], T)
            np.testing.assert_array_equal(xout[:npts.value], X)


def test_utils_match_reference(oracle, golden):
    U = golden("utils")
    f = oracle.field_from_desc(F.DoubleDipole().descriptor())
    tp, vel, v, th = U["tpos"], U["vel"], U["v"], U["th"]
    k = 0
    for i in range(len(tp)):
        t, pos = tp[i, 0], tp[i, 1:]
        assert oracle.cyclotron_period(f, t, pos, vel[i], K.m_el, -K.e) == U["cyclotron_period"][i]
        assert oracle.cyclotron_radius(f, t, pos, vel[i], K.m_el, -K.e) == U["cyclotron_radius"][i]
        gc, vp, spd = oracle.guidingcenter(f, t, pos, vel[i], K.m_el, -K.e)
        np.testing.assert_array_equal(np.r_[gc, vp, spd], U["guidingcenter"][i])
        vpar = v[i] * np.cos(th[i])
        for gp in (0.0, 1.3):
            p2, v2 = oracle.gctofp(f, t, pos, vpar, v[i], K.m_el, -K.e, gp)
            np.testing.assert_allclose(np.r_[p2, v2], U["gctofp"][k], rtol=1e-14, atol=0)
            k += 1
        assert oracle.magnetic_moment(f, t, pos, vpar, v[i], K.m_el) == U["mu"][i]


# ---- closed forms ------------------------------------------------------------------
def test_dipole_closed_form(oracle):
    """EarthDipole equals the textbook dipole B = B0 Re^3/r^3 (3 (m.r)r/r^2 - m)."""
    f = oracle.field_from_desc(F.EarthDipole().descriptor())
    rng = np.random.default_rng(5)
    for _ in range(50):
        r = rng.standard_normal(3) * 5 * K.Re
        B = oracle.field_eval(f, "B", np.r_[0, r])[0]
        rr = np.linalg.norm(r)
        m = np.array([0, 0, -1.0])  # moment direction (field +z at the equator)
        Bt = K.B0 * K.Re ** 3 / rr ** 3 * (3 * np.dot(m, r) * r / rr ** 2 - m)
        np.testing.assert_allclose(B, Bt, rtol=1e-13, atol=1e-13 * np.linalg.norm(Bt))


def test_fd_gradient_truncation(oracle):
    """grad|B| by central differences agrees with the analytic dipole
    gradient to the O((d/r)^2) truncation bound."""
    f = oracle.field_from_desc(F.DoubleDipole(distance=1e30).descriptor())
    d = K.Re / 1000
    rng = np.random.default_rng(6)
    for _ in range(20):
        r = rng.standard_normal(3)
        r = r / np.linalg.norm(r) * rng.uniform(3, 8) * K.Re
        g = oracle.field_eval(f, "gradB", np.r_[0, r])[0]
        rr = np.linalg.norm(r)
        k = K.B0 * K.Re ** 3
        q = 1 + 3 * r[2] ** 2 / rr ** 2
        dq = 3 * (2 * r[2] * np.array([0, 0, 1.0]) / rr ** 2 - 2 * r[2] ** 2 * r / rr ** 4)
        # |B| = k r^-3 sqrt(1 + 3 z^2/r^2), differentiated analytically
        ga = k * (-3 * r / rr ** 5 * np.sqrt(q) + rr ** -3 * 0.5 / np.sqrt(q) * dq)
        assert np.linalg.norm(g - ga) / np.linalg.norm(ga) < 10 * (d / rr) ** 2 + 1e-9


def test_uniform_gyration_exact(oracle):
    """In UniformBz a proton gyrates on a circle of radius gamma m v / (qB)
    with energy conserved: the Lorentz RHS and DOP853 driver together."""
    Bz = 1e-7
    f = oracle.field_from_desc(F.UniformBz(Bz=Bz).descriptor())
    v = 1e6
    g = 1 / np.sqrt(1 - (v / K.c) ** 2)
    row0 = np.array([0.0, 0.0, 0.0, 0.0, K.m_pr * g * v, 0.0, 0.0])
    rows, _ = oracle.particle_advance(f, oracle.particle_params(cyclotronresolution=20), row0,
                                      K.m_pr, K.e, 5.0)
    rc = g * K.m_pr * v / (K.e * Bz)
    center = np.array([0.0, -rc])  # positive charge, B along +z, v along +x
    dist = np.hypot(rows[:, 1] - center[0], rows[:, 2] - center[1])
    np.testing.assert_allclose(dist, rc, rtol=1e-7)
    p = np.linalg.norm(rows[:, 4:7], axis=1)
    np.testing.assert_allclose(p, K.m_pr * g * v, rtol=1e-9)


def test_crossed_eb_drift(oracle):
    """Guiding centre in UniformCrossedEB drifts at E x B / B^2 = Ey/Bz along x."""
    Ey, Bz = 1e-3, 1e-7
    f = oracle.field_from_desc(F.UniformCrossedEB(Ey=Ey, Bz=Bz).descriptor())
    row0 = np.array([0.0, 0.0, 0.0, 0.0, 0.0])
    mu = 1e-20
    rows, _ = oracle.gc_advance(f, oracle.gc_params(), "TaoChanBrizardEOM", row0, K.m_pr, K.e,
                                mu, 0.0, 0.1, 1.0)
    vx = (rows[-1, 1] - row0[1]) / (rows[-1, 0] - row0[0])
    np.testing.assert_allclose(vx, Ey / Bz, rtol=1e-9)
    assert abs(rows[-1, 2]) < 1e-6 * abs(rows[-1, 1])


def test_planted_mistake_is_caught(oracle):
    """A wrong sign in the Lorentz force (planted by flipping the charge)
    breaks the config-1 golden comparison: the check has teeth."""
    P_ = np.load(__import__("os").path.join(
        __import__("os").path.dirname(__file__), "golden", "particles.npz"))
    ref = P_["c1/traj"][:P_["c1/len"][0]]
    got, _ = _run_particle(oracle, F.EarthDipole().descriptor(), ref[0], K.m_pr, -K.e, 10.0,
                           cyclotronresolution=20)
    n = min(len(got), len(ref))
    assert np.max(np.abs(got[:n, 1:4] - ref[:n, 1:4])) > 1e3


@pytest.mark.parametrize("tol,hdiv,hmin", [(1e-3, 2.0, 1e-9), (np.inf, 3.0, 0.0)])
def test_rkf_tracer_stepper_is_rkf_per_output_step(oracle, tol, hdiv, hmin):
    """The RKF45 tracer stepper (a tracer advance with stepper RKF45) is
    rapt/rkf.py's rkf(f, t, t + dt, y, tol, hmax, hmin) once per output row,
    keeping the end point: bit-identical to orc_rkf (pinned against the
    reference's rkf above) chained over the rows, adaptive and in the
    fixed-step mode tol = inf (every step accepted, h = hmax; hmin = 0, since
    the step that ends an interval can be a rounding remainder far below any
    positive hmin, which rkf turns into its RuntimeError)."""
    L = oracle.lib()
    dp = C.POINTER(C.c_double)
    L.orc_rhs_lorentz.argtypes = [C.POINTER(oracle.Field), C.POINTER(oracle.Params), C.c_double,
                                  C.c_double, dp, C.c_double, dp, dp]
    L.orc_rhs_lorentz.restype = None
    f = F.EarthDipole()
    fo = oracle.field_from_desc(f.descriptor())
    from rapt_amd import workloads as W
    wl = W.dipole_protons_1mev(n=2)
    m, q = K.m_pr, K.e
    g = 1 / np.sqrt(1 - wl["vel"][1] @ wl["vel"][1] / float(K.c) ** 2)
    last = np.r_[0.0, wl["pos"][1], m * g * wl["vel"][1]]
    Tc = oracle.cyclotron_period(fo, 0.0, last[1:4], wl["vel"][1], m, q)
    hmax = Tc / 20 / hdiv
    prm = oracle.particle_params(stepper=oracle.RKF45, cyclotronresolution=20, rkf_tol=tol,
                                 rkf_hmax=hmax, rkf_hmin=hmin)
    rows, info = oracle.particle_advance(fo, prm, last, m, q, 2.0)
    assert info["status"] == 1 and len(rows) > 50
    gm = C.c_double(np.sqrt(m ** 2 + last[4:] @ last[4:] / float(K.c) ** 2))

    @RKF_CB
    def cb(ctx, t, y, out):
        L.orc_rhs_lorentz(C.byref(fo), C.byref(prm), m, q, C.byref(gm), t, y, out)
    y, t = last[1:].copy(), last[0]
    cap = 4096
    tout, xout, npts = np.zeros(cap), np.zeros((cap, 6)), C.c_int64()
    nacc = 0
    for row in rows:
        rc = L.orc_rkf(C.cast(cb, C.c_void_p), None, 6, t, row[0], oracle._dp(y), tol, hmax,
                       hmin, oracle._dp(tout), oracle._dp(xout), cap, C.byref(npts))
        assert rc == 1
        k = npts.value
        t, y = tout[k - 1], xout[k - 1].copy()
        nacc += k - 1
        np.testing.assert_array_equal(y, row[1:])
    assert nacc == info["naccpt"]
