"""Surrogate labels (rapt_bounce_labels) in a pure dipole against the closed
forms along a dipole field line r = L Re sin^2(theta):
  B(theta) = B0 sqrt(1 + 3 cos^2 theta) / (L^3 sin^6 theta),
  ds = L Re sin(theta) sqrt(1 + 3 cos^2 theta) dtheta,
  S_b = int ds / sqrt(1 - B/Bm), I = int sqrt(1 - B/Bm) ds between the mirror
  colatitudes (B = Bm = B_eq / sin^2 alpha), T_b = 2 S_b / v, and the
  bounce-averaged drift omega_d = -gamma m v^2 / (q S_b B_eq) dI/dr / r with
  dI/dr the same central difference (step 0.03 Re) of the exact I.
Tolerances (measured errors are about 1.3e-5, 1.6e-5 and 1.6e-6): T_b 5e-5
relative (the device integrates a quadratic spline through traced points, as
the reference does), mirror latitude 1e-5, omega_d 1e-4 (a difference
quotient of I over 0.03 Re amplifies the quadrature error)."""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import brentq

import rapt_amd as rapt
from rapt_amd import _native as N
from rapt_amd import constants as K
from rapt_amd import fields as F

pytestmark = pytest.mark.gpu


def _dipole(L, alpha_deg, logK, m, q, d=0.03):
    Re, B0 = K.Re, K.B0
    a = np.radians(alpha_deg)
    Kj = 10.0 ** logK * K.e
    g = 1 + Kj / (m * K.c ** 2)
    v = K.c * np.sqrt(1 - 1 / g ** 2)

    def parts(Lx, Bm):
        Bth = lambda th: B0 * np.sqrt(1 + 3 * np.cos(th) ** 2) / (Lx ** 3 * np.sin(th) ** 6)
        th_m = brentq(lambda th: Bth(th) - Bm, 1e-3, np.pi / 2 - 1e-12)
        dsd = lambda th: Lx * Re * np.sin(th) * np.sqrt(1 + 3 * np.cos(th) ** 2)
        # theta = pi/2 - (pi/2 - th_m) cos(u): removes the end-point singularities
        w = np.pi / 2 - th_m
        fS = lambda u: dsd(np.pi / 2 - w * np.cos(u)) * w * np.sin(u) / np.sqrt(
            max(1 - Bth(np.pi / 2 - w * np.cos(u)) / Bm, 1e-300))
        fI = lambda u: dsd(np.pi / 2 - w * np.cos(u)) * w * np.sin(u) * np.sqrt(
            max(1 - Bth(np.pi / 2 - w * np.cos(u)) / Bm, 0.0))
        S = quad(fS, 0, np.pi, epsabs=0, epsrel=1e-12, limit=200)[0]
        I = quad(fI, 0, np.pi, epsabs=0, epsrel=1e-12, limit=200)[0]
        return S, I, th_m

    Beq = B0 / L ** 3
    Bm = Beq / np.sin(a) ** 2
    S, I, th_m = parts(L, Bm)
    Ip = parts(L + d, Bm)[1]
    Im = parts(L - d, Bm)[1]
    dIdr = (Ip - Im) / (2 * d * Re)
    wd = -g * m * v * v / (q * S * Beq) * dIdr / (L * Re)
    return 2 * S / v, wd, 90 - np.degrees(th_m)


def test_labels_match_dipole_closed_forms():
    rng = np.random.default_rng(5)
    n = 24
    qry = np.c_[rng.uniform(2.5, 7.5, n), rng.uniform(15, 80, n), rng.uniform(4, 7, n)]
    f = F.EarthDipole()
    out, st = N.context().bounce_labels(f.descriptor(), qry, K.m_pr, K.e,
                                        rapt.params["eyegradientstep"],
                                        rapt.params["fieldlineresolution"])
    assert np.all(st == 1), st
    ref = np.array([_dipole(*q, K.m_pr, K.e) for q in qry])
    rel = np.abs(out - ref) / np.abs(ref)
    print("label rel err max (Tb, wd, lat):", rel.max(axis=0))
    assert np.all(rel[:, 0] < 5e-5)
    assert np.all(rel[:, 2] < 1e-5)
    assert np.all(rel[:, 1] < 1e-4)
    assert np.all(out[:, 1] < 0)   # protons drift westward


def test_labels_reject_time_dependent_field():
    with pytest.raises(RuntimeError, match="static"):
        N.context().bounce_labels(F.VarEarthDipole(amp=0.1, period=5.0).descriptor(),
                                  np.array([[4.0, 45.0, 5.0]]), K.m_pr, K.e, 0.03 * K.Re, 50)
