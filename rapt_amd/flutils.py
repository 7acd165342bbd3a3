"""Bounce integrals along field lines, interface of rapt/flutils.py (host side).

halfbouncepath S_b = int ds / sqrt(1 - B/Bm) between the mirror points,
bounceperiod = 2 S_b / v, eye (second invariant I) and gradI, computed from a
traced field line with quadratic interpolation, root finding and adaptive
quadrature as rapt/flutils.py:65-316 does.  They set up the guiding-centre
output step; they are not part of the traced loop.
"""
import numpy as np
from scipy.integrate import quad, simpson
from scipy.interpolate import interp1d
from scipy.optimize import brentq

from .fieldline import Fieldline
from .utils import dipoleI, dipole_gradI


def _trim(s, b, Bm, inclusive):
    n = len(b)
    inside = np.where(b <= Bm if inclusive else b < Bm)[0]
    if len(inside) == 0:
        i1, i2 = int((n - 3) / 2), int((n + 1) / 2)
    else:
        i1, i2 = inside[0] - 1, inside[-1] + 1
    return s[i1:i2 + 1], b[i1:i2 + 1]


def halfbouncepath(tpos, field, Bm):
    """Half-bounce path length S_b (rapt/flutils.py:254-316)."""
    fl = Fieldline(tpos, field, Bmax=Bm)
    fl.trace()
    s, b = _trim(fl.gets(), fl.getB(), Bm, True)
    if len(b) == 3:
        # equatorial particle: S_b = pi sqrt(2 Bm / B'') with B'' from three points
        s12, s23, s13 = s[0] - s[1], s[1] - s[2], s[0] - s[2]
        B2s = 2 * (b[0] * s23 - b[1] * s13 + b[2] * s12) / (s12 * s13 * s23)
        return np.pi * np.sqrt(2 * Bm / B2s)
    B = interp1d(s, b, kind="quadratic", assume_sorted=True)
    sm1 = brentq(lambda x: B(x) - Bm, s[0], s[1])
    sm2 = brentq(lambda x: B(x) - Bm, s[-2], s[-1])
    return quad(lambda x: 1 / np.sqrt(1 - B(x) / Bm), sm1, sm2, epsrel=1e-4)[0]


def bounceperiod(tpos, field, Bm, v):
    """2 S_b / v (rapt/flutils.py:232-252)."""
    return (2 / v) * halfbouncepath(tpos, field, Bm)


def eye(tpos, field, Bm, usedipole=False):
    """Second invariant I = int sqrt(1 - B/Bm) ds (rapt/flutils.py:65-151)."""
    if usedipole:
        return dipoleI(np.sqrt(np.dot(tpos[1:3], tpos[1:3])), Bm)
    fl = Fieldline(tpos, field, Bmax=Bm)
    fl.trace()
    s, b = fl.gets(), fl.getB()
    Bmin = np.min(b)
    if Bmin > Bm or abs(Bmin - Bm) / Bm < 1e-12:
        return 0
    inside = np.where(b < Bm)[0]
    keep = slice(max(inside[0] - 1, 0), inside[-1] + 2)
    s, b = s[keep].copy(), b[keep].copy()
    eqpa = np.degrees(np.arcsin(np.sqrt(Bmin / Bm)))
    if eqpa < 70:
        sm1 = (Bm - b[0]) * (s[1] - s[0]) / (b[1] - b[0]) + s[0]
        sm2 = (Bm - b[-2]) * (s[-1] - s[-2]) / (b[-1] - b[-2]) + s[-2]
        s[0], s[-1] = sm1, sm2
        b[0], b[-1] = Bm, Bm
        val = simpson(np.sqrt(1 - b[1:-1] / Bm), x=s[1:-1])
        val += (2 / 3) * (s[-1] - s[-2]) * np.sqrt((Bm - b[-2]) / Bm)
        val += (2 / 3) * (s[1] - s[0]) * np.sqrt((Bm - b[1]) / Bm)
        return val
    B = interp1d(s, b, kind="quadratic", assume_sorted=True)
    sm1 = brentq(lambda x: B(x) - Bm, s[0], s[1])
    sm2 = s[-2] if B(s[-2]) == Bm else brentq(lambda x: B(x) - Bm, s[-2], s[-1])
    return quad(lambda x: np.sqrt(1 - B(x) / Bm), sm1, sm2, epsrel=1e-4)[0]


def gradI(tpos, field, Bm, usedipole=False):
    """Gradient of I across the field by central differences (rapt/flutils.py:153-229)."""
    from . import params
    tpos = np.array(tpos, dtype=float)
    x, y = tpos[1], tpos[2]
    d = params["eyegradientstep"]
    if usedipole:
        return dipole_gradI(x, y, Bm, d)
    r = np.array((x, y, 0)) / max(abs(x), abs(y))
    b = field.unitb(tpos)
    r = r - np.dot(r, b) * b
    r = r / np.sqrt(np.dot(r, r))
    v1 = np.zeros(4)
    v1[1:4] = r
    v2 = np.zeros(4)
    v2[1:4] = np.cross(b, r)
    I = lambda p: eye(p, field, Bm)

    def deriv(u):
        I1, I2 = I(tpos + d * u), I(tpos - d * u)
        if I1 == 0:
            return (I(tpos) - I2) / d
        if I2 == 0:
            return (I1 - I(tpos)) / d
        return (I1 - I2) / (2 * d)
    return deriv(v1) * v1[1:] + deriv(v2) * v2[1:]


# ---- field-line diagnostics (host, as in the reference) -------------------------------
def criticalpoints(tpos, field, Bm):
    """(s, B, B'') at each extremum of |B| along the field line through tpos,
    traced to |B| = Bm (rapt/flutils.py:21-63): the trace points where the
    sign of dB changes seed a local minimisation of the cubic interpolant of
    B(s); B'' is its central second difference with step 0.1 ds (the
    reference's scipy.misc.derivative(n=2), removed from scipy, is that
    three-point formula)."""
    from scipy.optimize import minimize
    fl = Fieldline(tpos, field, Bmax=Bm)
    fl.trace()
    s, b = fl.gets(), fl.getB()
    db = np.diff(b)
    seeds = [s[i] for i in range(1, len(b) - 1) if db[i - 1] * db[i] <= 0]
    B = interp1d(s, b, kind="cubic", assume_sorted=True)
    h = 0.1 * (s[1] - s[0])
    out = []
    for s0 in seeds:
        sm = minimize(B, s0).x[0]
        out.append((sm, float(B(sm)), float((B(sm + h) - 2 * B(sm) + B(sm - h)) / h ** 2)))
    return out


def fl_error(tpos, field, Bm):
    """(max, mean) relative distance of the traced field-line points from the
    dipole line r = r0 cos^2(latitude) through tpos (rapt/flutils.py:318-352;
    the mean divides by the number of points less one, as there)."""
    fl = Fieldline(tpos, field, Bmax=Bm)
    fl.trace()
    r0 = np.sqrt(np.dot(tpos[1:], tpos[1:]))
    pts = fl.curve[:, 1:]
    r = np.sqrt(np.sum(pts * pts, axis=1))
    err = np.abs(1 - r / (r0 * np.cos(np.arcsin(pts[:, 2] / r)) ** 2))
    return (float(np.max(err)) if len(err) else 0.0,
            float(np.sum(err) / (len(fl.curve) - 1)))


def eye_error(tpos, field, Bm):
    """I from the traced field line minus the dipole formula's
    (rapt/flutils.py:354-376)."""
    return eye(tpos, field, Bm, usedipole=False) - eye(tpos, field, Bm, usedipole=True)
