"""The tracers' right-hand sides on the device (rapt_rhs_eval) against the
oracle's (orc_rhs_lorentz, orc_rhs_gc: the functions whose trajectories
tests/test_oracle.py pins bit for bit against the reference's), point by
point, and the uniform-field Lorentz force in closed form.

Tolerances: Lorentz 1e-13 of the vector's norm (a Newton-refined reciprocal
of gamma m and fused products against the reference's divisions: a few
ulp); guiding centre 1e-9 (the finite-difference gradient and curl amplify
those ulps by the field scale over the stencil step, ~1e3 for DoubleDipole).
"""
import ctypes as C

import numpy as np
import pytest

import rapt_amd as rapt
import rapt_amd.fields as F
from rapt_amd import _native as N
from rapt_amd import constants as K

pytestmark = pytest.mark.gpu

FIELDS = {"EarthDipole": F.EarthDipole, "DoubleDipole": F.DoubleDipole,
          "UniformCrossedEB": F.UniformCrossedEB, "CurrentSheet": F.CurrentSheet,
          "VarEarthDipole": F.VarEarthDipole}


def _points(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(3, 8, n) * K.Re
    th = np.arccos(rng.uniform(-0.7, 0.7, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.c_[rng.uniform(0, 10, n), r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph),
                 r * np.cos(th)]


def _bind(oracle):
    L = oracle.lib()
    dp = C.POINTER(C.c_double)
    L.orc_rhs_lorentz.argtypes = [C.POINTER(oracle.Field), C.POINTER(oracle.Params), C.c_double,
                                  C.c_double, dp, C.c_double, dp, dp]
    L.orc_rhs_lorentz.restype = None
    L.orc_rhs_gc.argtypes = [C.POINTER(oracle.Field), C.POINTER(oracle.Params), C.c_int,
                             C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
    L.orc_rhs_gc.restype = None
    return L


def _rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.mark.parametrize("name", sorted(FIELDS))
def test_lorentz_rhs_vs_oracle(oracle, name):
    L = _bind(oracle)
    f = FIELDS[name]()
    n = 256
    tp = _points(n, 3)
    rng = np.random.default_rng(4)
    m, q = K.m_pr, K.e
    p = rng.normal(size=(n, 3)) * m * 1e7
    gm = np.sqrt(m ** 2 + np.sum(p * p, axis=1) / float(K.c) ** 2)
    ty = np.c_[tp, p]
    got = N.context().rhs_eval(f.descriptor(), N.MODE_PARTICLE, ty,
                               np.c_[np.full(n, m), np.full(n, q), gm])
    fo = oracle.field_from_desc(f.descriptor())
    prm = oracle.particle_params()
    ref = np.zeros((n, 6))
    out = np.zeros(6)
    for i in range(n):
        g = C.c_double(gm[i])
        y = np.ascontiguousarray(ty[i, 1:])
        L.orc_rhs_lorentz(C.byref(fo), C.byref(prm), m, q, C.byref(g), ty[i, 0], oracle._dp(y),
                          oracle._dp(out))
        ref[i] = out
    assert np.max(_rel(got[:, :3], ref[:, :3])) < 1e-13
    assert np.max(_rel(got[:, 3:], ref[:, 3:])) < 1e-13


def test_lorentz_rhs_uniform_field_closed_form():
    """UniformBz: dx/dt = p/(gamma m), dp/dt = q (p/(gamma m)) x B exactly."""
    Bz = 7e-8
    f = F.UniformBz(Bz)
    n = 64
    tp = _points(n, 5)
    rng = np.random.default_rng(6)
    m, q = K.m_el, -K.e
    p = rng.normal(size=(n, 3)) * m * 3e7
    gm = np.sqrt(m ** 2 + np.sum(p * p, axis=1) / float(K.c) ** 2)
    got = N.context().rhs_eval(f.descriptor(), N.MODE_PARTICLE, np.c_[tp, p],
                               np.c_[np.full(n, m), np.full(n, q), gm])
    v = p / gm[:, None]
    want = np.c_[v, q * np.cross(v, np.array([0.0, 0.0, Bz]))]
    assert np.max(_rel(got[:, :3], want[:, :3])) < 1e-15
    assert np.max(_rel(got[:, 3:], want[:, 3:])) < 1e-14


# BrizardChan and NorthropTeller are the static, E-free equations: those two
# run in the static fields only
@pytest.mark.parametrize("name,eom", [
    (f, e) for f in ("EarthDipole", "DoubleDipole")
    for e in ("TaoChanBrizardEOM", "BrizardChanEOM", "NorthropTellerEOM")] + [
    ("UniformCrossedEB", "TaoChanBrizardEOM"), ("CurrentSheet", "TaoChanBrizardEOM"),
    ("VarEarthDipole", "TaoChanBrizardEOM")])
def test_gc_rhs_vs_oracle(oracle, name, eom):
    L = _bind(oracle)
    f = FIELDS[name]()
    n = 256
    tp = _points(n, 7)
    rng = np.random.default_rng(8)
    m, q = K.m_el, -K.e
    v = rng.uniform(1e7, 2.5e8, n)
    g = 1 / np.sqrt(1 - (v / float(K.c)) ** 2)
    pa = np.radians(rng.uniform(10, 80, n))
    ppar = g * m * v * np.cos(pa)
    B = N.context().field_eval(f.descriptor(), "magB", tp)[:, 0]
    mu = g ** 2 * m * (v * np.sin(pa)) ** 2 / (2 * B)
    ty = np.c_[tp, ppar]
    got = N.context().rhs_eval(f.descriptor(), N.MODE_GC, ty, np.c_[np.full(n, m),
                               np.full(n, q), mu, v], eom=eom)
    fo = oracle.field_from_desc(f.descriptor())
    prm = oracle.gc_params()
    e = N.EOM[eom]
    ref = np.zeros((n, 4))
    out = np.zeros(4)
    for i in range(n):
        y = np.ascontiguousarray(ty[i, 1:])
        L.orc_rhs_gc(C.byref(fo), C.byref(prm), e, m, q, mu[i], v[i], ty[i, 0], oracle._dp(y),
                     oracle._dp(out))
        ref[i] = out
    assert np.max(_rel(got[:, :3], ref[:, :3])) < 1e-9
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=1e-9,
                               atol=1e-9 * np.max(np.abs(ref[:, 3])))


def test_guiding_center_eom_methods(oracle):
    """GuidingCenter._TaoChanBrizardEOM and its two siblings (the reference's
    private equations, rapt/GuidingCenter.py:329-395) through the object."""
    f = F.DoubleDipole()
    gc = rapt.GuidingCenter([5 * K.Re, 0.3 * K.Re, 0.2 * K.Re], 5e7, 40.0, None, 0.0, K.m_el,
                            -K.e, f)
    Y = gc.trajectory[-1, 1:5]
    for eom in ("TaoChanBrizardEOM", "BrizardChanEOM", "NorthropTellerEOM"):
        a = getattr(gc, "_" + eom)(0.0, Y)
        b = N.context().rhs_eval(f.descriptor(), N.MODE_GC, np.r_[0.0, Y][None, :],
                                 [K.m_el, -K.e, gc.mu, gc.v], eom=eom)[0]
        np.testing.assert_array_equal(a, b)
        assert a.shape == (4,) and np.all(np.isfinite(a))
