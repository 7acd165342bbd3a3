"""Physics helpers with the signatures of rapt/utils.py (host side, numpy).

These convert between tracer representations and set up initial conditions;
they run once per tracer construction or mode switch, not per step.  The
device kernels carry their own copies of the conversions that happen inside a
batched Adaptive run.
"""
import numpy as np
from scipy.integrate import quad
from scipy.optimize import brentq

from .constants import B0, Re, c, e


def _tp(t, pos):
    return np.concatenate([[t], np.asarray(pos, dtype=float)])


def gamma(mass=0, **kwargs):
    """Relativistic factor from ``velocity=`` or ``momentum=`` (rapt/utils.py:15-27)."""
    if "velocity" in kwargs:
        v = np.asarray(kwargs["velocity"], dtype=float)
        return 1 / np.sqrt(1 - np.dot(v, v) / c ** 2)
    if "momentum" in kwargs:
        if mass == 0:
            raise ValueError("Particle mass not given.")
        p = np.asarray(kwargs["momentum"], dtype=float)
        return np.sqrt(1 + np.dot(p, p) / (mass * c) ** 2)
    raise ValueError("Either velocity of momentum vectors should be given.")


def cyclotron_period(t, pos, vel, field, mass, charge):
    """2 pi gamma m / (|q| B) from a velocity vector (rapt/utils.py:29-66)."""
    vel = np.asarray(vel, dtype=float)
    g = 1.0 / np.sqrt(1 - np.dot(vel, vel) / c ** 2)
    B = field.magB(_tp(t, pos))
    return 2 * np.pi * g * mass / B / abs(charge)


def cyclotron_period2(t, pos, speed, field, mass, charge):
    """2 pi gamma m / (|q| B) from the speed (rapt/utils.py:68-104)."""
    g = 1.0 / np.sqrt(1 - (speed / c) ** 2)
    B = field.magB(_tp(t, pos))
    return 2 * np.pi * g * mass / B / abs(charge)


def cyclotron_radius(t, pos, vel, field, mass, charge):
    """gamma m v_perp / (|q| B) from a velocity vector (rapt/utils.py:106-145)."""
    vel = np.asarray(vel, dtype=float)
    vsq = np.dot(vel, vel)
    g = 1.0 / np.sqrt(1 - vsq / c ** 2)
    B = field.B(_tp(t, pos))
    Bmag = np.sqrt(np.dot(B, B))
    vpar = np.dot(vel, B) / Bmag
    vperp = np.sqrt(vsq - vpar ** 2)
    return g * mass * vperp / (abs(charge) * Bmag)


def cyclotron_radius2(t, pos, vpar, v, field, mass, charge):
    """gamma m v_perp / (|q| B) from parallel and total speed (rapt/utils.py:147-187)."""
    g = 1.0 / np.sqrt(1 - (v / c) ** 2)
    B = field.B(_tp(t, pos))
    Bmag = np.sqrt(np.dot(B, B))
    vperp = np.sqrt((v - vpar) * (v + vpar))
    return g * mass * vperp / (abs(charge) * Bmag)


def magnetic_moment(t, pos, vpar, v, field, mass):
    """gamma^2 m v_perp^2 / (2B) (rapt/utils.py:189-216)."""
    g = 1.0 / np.sqrt(1 - (v / c) ** 2)
    Bmag = field.magB(_tp(t, pos))
    return g ** 2 * mass * (v - vpar) * (v + vpar) / (2 * Bmag)


def speedfromKE(KE, mass, unit="ev"):
    """Speed for a kinetic energy, non-relativistic below 1e-6 mc^2 (rapt/utils.py:218-249)."""
    mc2 = mass * c ** 2
    if unit.lower() == "ev":
        KE = KE * e
    if KE / mc2 < 1e-6:
        return np.sqrt(2 * KE / mass)
    return c * np.sqrt(1 - (mc2 / (mc2 + KE)) ** 2)


def guidingcenter(t, r, v, field, mass, charge, tol=1e-3, maxiter=20, debug=False):
    """Solve R = r - rho(R) by fixed-point iteration (rapt/utils.py:251-326).

    Returns (R, v_parallel, speed), or the list of iterates when ``debug``.
    Raises RuntimeError when ``maxiter`` iterations do not reach ``tol``
    (the reference prints a message and returns None).
    """
    v = np.asarray(v, dtype=float)
    r = np.asarray(r, dtype=float)
    vsq = np.dot(v, v)
    g = 1 / np.sqrt(1 - vsq / c ** 2)

    def rho(q):
        B = field.B(_tp(t, q))
        return g * mass / (charge * np.dot(B, B)) * np.cross(B, v)

    R_old = r - rho(r)
    iterates = [R_old]
    for _ in range(maxiter):
        R = r - rho(R_old)
        iterates.append(R)
        step = R - R_old
        if np.sqrt(np.dot(step, step)) / np.sqrt(np.dot(R, R)) < tol:
            if debug:
                return iterates
            B = field.B(_tp(t, R))
            return R, np.dot(v, B) / np.sqrt(np.dot(B, B)), np.sqrt(vsq)
        R_old = R
    raise RuntimeError("guidingcenter: no convergence to %g in %d iterations" % (tol, maxiter))


def getperp(v):
    """A unit vector perpendicular to v, chosen as in rapt/utils.py:328-374."""
    assert len(v) == 3
    if v[0] == 0.0 and v[1] == 0.0 and v[2] == 0.0:
        raise ValueError("Zero vector")
    if v[0] == 0:
        return np.array([1, 0, 0])
    if v[1] == 0:
        return np.array([0, 1, 0])
    if v[2] == 0:
        return np.array([0, 0, 1])
    cc = -1.0 * (v[0] + v[1]) / v[2]
    return np.array([1, 1, cc]) / np.sqrt(2 + cc ** 2)


def GCtoFP(t, R, vp, speed, field, mass, charge, gyrophase=0):
    """Particle position and velocity for a guiding centre (rapt/utils.py:376-433)."""
    R = np.asarray(R, dtype=float)
    B = field.B(_tp(t, R))
    b = B / np.sqrt(np.dot(B, B))
    pa = np.arccos(vp / speed)
    rc = cyclotron_radius2(t, R, vp, speed, field, mass, charge)
    u = getperp(B)
    u = u / np.sqrt(np.dot(u, u))
    w = np.cross(b, u)
    s = np.sign(charge)
    pos = R + rc * (np.cos(gyrophase) * u + np.sin(gyrophase) * w)
    vel = speed * (np.cos(pa) * b + s * np.sin(pa) * np.sin(gyrophase) * u
                   - s * np.sin(pa) * np.cos(gyrophase) * w)
    return pos, vel


def dipole_drift_period_p(L, v, eqpa):
    """Walt's proton drift period in a dipole (rapt/utils.py:435-460)."""
    b = v / c
    g = 1 / np.sqrt(1 - b ** 2)
    a = np.radians(eqpa)
    return 8.481 / (L * g * b ** 2) * (1 - (1 / 3) * (np.sin(a) ** 0.62))


def dipole_drift_period_e(L, v, eqpa):
    """Walt's electron drift period in a dipole (rapt/utils.py:462-487)."""
    b = v / c
    g = 1 / np.sqrt(1 - b ** 2)
    a = np.radians(eqpa)
    return 1.557e4 / (L * g * b ** 2) * (1 - 1 / 3 * (np.sin(a) ** 0.62))


def dipole_bounce_period(L, v, eqpa):
    """Walt's bounce period in a dipole (rapt/utils.py:489-513)."""
    a = np.radians(eqpa)
    return 0.117 * L / (v / c) * (1 - 0.4635 * (np.sin(a) ** 0.75))


def Bm(tpos, pa, field):
    """Mirror field B/sin^2(pa) at tpos (rapt/utils.py:516-522)."""
    return field.magB(tpos) / np.sin(np.radians(pa)) ** 2


def dipoleBm(L, eqpa):
    """Dipole mirror field for equatorial pitch angle eqpa (rapt/utils.py:524-529)."""
    return B0 / L ** 3 / np.sin(np.radians(eqpa)) ** 2


def dipoleI(R0, Bm):
    """Second invariant of a dipole field line by quadrature (rapt/utils.py:531-557)."""
    C = 1 / ((Bm / B0) ** 2 * (R0 / Re) ** 6)
    ym = brentq(lambda y: y ** 6 - C * (4 - 3 * y), 0, 1)
    th1 = np.arcsin(np.sqrt(ym))
    th2 = np.pi - th1
    K = np.sqrt(C)
    f = lambda x: np.sqrt((1 - K * np.sqrt(1 + 3 * x ** 2) / (1 - x ** 2) ** 3) * (1 + 3 * x ** 2))
    return R0 * quad(f, np.cos(th2), np.cos(th1))[0]


def dipole_gradI(x, y, Bm, h):
    """Radial central difference of dipoleI (rapt/utils.py:559-581)."""
    r0 = np.sqrt(x ** 2 + y ** 2)
    f = lambda dx: dipoleI(r0 + dx, Bm)
    return (f(h) - f(-h)) / (2 * h) * np.array((x, y, 0)) / r0


def drift_error(b):
    """Relative deviation of a BounceCenter trajectory's distance from the
    origin from the first row's, row by row -- the error against the circular
    drift path of a dipole (rapt/utils.py:583-606)."""
    xyz = np.asarray(b.trajectory)[:, 1:]
    r0 = np.sqrt(np.dot(xyz[0], xyz[0]))
    return (np.sqrt(np.sum(xyz * xyz, axis=1)) - r0) / r0
