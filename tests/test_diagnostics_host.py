"""The reference's debugging diagnostics on the host (rapt/flutils.py:21-63,
318-376, rapt/utils.py:583-606), pinned by the dipole's closed forms:

- criticalpoints: a dipole field line through the equator has one extremum
  of |B|, the minimum at the start point, where B = B0 / L^3 and, along the
  line, d^2B/ds^2 = 9 B_eq / r0^2 (B = B_eq (1 + 9/2 lat^2 + ...) and
  ds = r0 dlat at the equator);
- fl_error: the traced points lie on r = r0 cos^2(lat) to the tracer's
  accuracy;
- eye_error: the traced I against the closed form to the reference's
  quadrature error (Simpson on the trace points below 70 deg: ~1e-4);
- drift_error: exact on a constructed trajectory.
"""
import numpy as np

import rapt_amd as rapt
from rapt_amd import flutils, utils


def _dipole_line(L=5.0, mirror=8.0):
    f = rapt.fields.EarthDipole()
    tpos = np.array([0.0, L * rapt.Re, 0.0, 0.0])
    beq = f.magB(tpos)
    return f, tpos, beq, mirror * beq


def test_criticalpoints_dipole_minimum():
    f, tpos, beq, bm = _dipole_line()
    cp = flutils.criticalpoints(tpos, f, bm)
    assert len(cp) == 1
    s, b, b2 = cp[0]
    r0 = tpos[1]
    assert abs(s) < 1e-6 * r0
    assert abs(b / beq - 1) < 1e-12
    assert abs(b2 / (9 * beq / r0 ** 2) - 1) < 1e-3


def test_fl_error_dipole():
    f, tpos, beq, bm = _dipole_line(L=6.0)
    e_max, e_avg = flutils.fl_error(tpos, f, bm)
    assert 0 <= e_avg <= e_max < 1e-8


def test_eye_error_dipole():
    f, tpos, beq, bm = _dipole_line(L=4.0, mirror=5.0)
    I = flutils.eye(tpos, f, bm, usedipole=True)
    assert abs(flutils.eye_error(tpos, f, bm)) < 3e-4 * I


def test_drift_error_constructed():
    class _B:
        pass
    b = _B()
    ph = np.linspace(0, 2, 5)
    b.trajectory = np.c_[np.arange(5.0), 3 * np.cos(ph), 3 * np.sin(ph), np.zeros(5)]
    b.trajectory[3, 1:3] *= 1.25
    e = utils.drift_error(b)
    np.testing.assert_allclose(e, [0, 0, 0, 0.25, 0], atol=1e-15)
