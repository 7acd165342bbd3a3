"""Guiding-centre tracer with the API of rapt/GuidingCenter.py.

``advance`` runs on the GPU: DOPRI5 with scipy's call semantics and one of
the three equations of motion (Tao-Chan-Brizard, Brizard-Chan,
Northrop-Teller; rapt/GuidingCenter.py:329-395), one row per output step
``params['GCtimestep']`` or ``bounceperiod()/bounceresolution``
(rapt/GuidingCenter.py:442-458).  Deliberate differences from the reference,
in code paths that crash or return wrong values there: empty-argument
detection (rapt/GuidingCenter.py:123), ValueError from ``init`` (:188),
``getp`` (:507) and ``getv`` (:510-515) return momentum and speed.
"""
import pickle

import numpy as np

from . import utils as ru
from .constants import c

_EOMS = ("TaoChanBrizardEOM", "BrizardChanEOM", "NorthropTellerEOM")
# field models whose half-bounce path the device computes (rapt_halfbounce)
_BOUNCE_MODELS = (1, 2, 5, 8, 9, 10)


class GuidingCenter:
    """Guiding centre of a charged particle.  ``trajectory`` rows are
    (t, x, y, z, ppar); ``mu`` is the (constant) magnetic moment."""

    def __init__(self, pos=[], v=0, pa=None, ppar=None, t0=0, mass=None, charge=None,
                 field=None):
        self.pos = pos
        self.v = v
        self.t0 = t0
        self.tcur = t0
        self.mass = mass
        self.charge = charge
        self.field = field
        self.trajectory = np.zeros((1, 5))
        self.check_adiabaticity = False
        if np.size(pos) > 0 and v != 0:
            g = 1 / np.sqrt(1 - (v / c) ** 2)
            if pa is not None:
                vpar = 0 if pa == 90 else v * np.cos(pa * np.pi / 180)
                ppar = g * mass * vpar
            self.mu = ru.magnetic_moment(self.tcur, self.pos, ppar / (mass * g), self.v,
                                         self.field, self.mass)
            self.trajectory[0, 0] = t0
            self.trajectory[0, 1:4] = pos[:]
            self.trajectory[0, 4] = ppar

    def init(self, p):
        """Initialise from the last state of a GuidingCenter or a Particle."""
        from .Particle import Particle
        if isinstance(p, GuidingCenter):
            B = p.field.magB(p.trajectory[-1, :4])
            g = np.sqrt(1 + 2 * p.mu * B / (p.mass * c * c) + (p.trajectory[-1, 4] / p.mass / c) ** 2)
            if g - 1 < 1e-6:
                v = np.sqrt(2 * p.mu * B / p.mass + (p.trajectory[-1, 4] / p.mass) ** 2)
            else:
                v = c * np.sqrt(1 - 1 / g ** 2)
            self.__init__(pos=p.trajectory[-1, 1:4], v=v, ppar=p.trajectory[-1, 4],
                          t0=p.trajectory[-1, 0], mass=p.mass, charge=p.charge, field=p.field)
            self.check_adiabaticity = p.check_adiabaticity
        elif isinstance(p, Particle):
            mom = p.trajectory[-1, 4:]
            gm = np.sqrt(p.mass ** 2 + np.dot(mom, mom) / c ** 2)
            pos, vp, v = ru.guidingcenter(p.trajectory[-1, 0], p.trajectory[-1, 1:4], mom / gm,
                                          p.field, p.mass, p.charge)
            g = 1 / np.sqrt(1 - (v / c) ** 2)
            self.__init__(pos=pos, v=v, ppar=p.mass * g * vp, t0=p.trajectory[-1, 0],
                          mass=p.mass, charge=p.charge, field=p.field)
            self.check_adiabaticity = p.check_adiabaticity
        else:
            raise ValueError("Particle or GuidingCenter objects required.")

    def save(self, filename):
        with open(filename, "wb") as f:
            pickle.dump(self, f)

    def load(self, filename):
        """Load a pickled GuidingCenter written by ``save`` (trusted files only)."""
        with open(filename, "rb") as f:
            p = pickle.load(f)
        self.__dict__.update(p.__dict__)

    def setke(self, ke, unit="ev"):
        """Set the kinetic energy keeping the pitch angle (rapt/GuidingCenter.py:225-255)."""
        mc = self.mass * c
        t, x, y, z, ppar = self.trajectory[-1]
        B = self.field.magB(self.trajectory[-1, :4])
        g2m1 = 2 * self.mu * B / (mc * c) + (ppar / mc) ** 2
        if np.sqrt(g2m1 + 1) - 1 < 1e-6:
            ptot = np.sqrt(2 * self.mass * self.mu * B + ppar ** 2)
        else:
            ptot = np.sqrt(g2m1) * mc
        pa_old = np.degrees(np.arccos(ppar / ptot))
        v_new = ru.speedfromKE(ke, self.mass, unit)
        self.__init__(pos=[x, y, z], v=v_new, pa=pa_old, t0=t, mass=self.mass,
                      charge=self.charge, field=self.field)

    def setpa(self, pa):
        """Set the pitch angle (degrees) keeping the energy (rapt/GuidingCenter.py:257-285)."""
        mc = self.mass * c
        t, x, y, z, ppar = self.trajectory[-1]
        B = self.field.magB(self.trajectory[-1, :4])
        gsq = 1 + 2 * self.mu * B / (mc * c) + (ppar / mc) ** 2
        if np.sqrt(gsq) - 1 < 1e-6:
            v = np.sqrt(2 * self.mass * self.mu * B + ppar ** 2) / self.mass
        else:
            v = c * np.sqrt(1 - 1 / gsq)
        self.__init__(pos=[x, y, z], v=v, pa=pa, t0=t, mass=self.mass, charge=self.charge,
                      field=self.field)

    def isadiabatic(self):
        """rho_c / L < epss (and tau_c / T < epst for non-static fields)."""
        from . import params as P
        tpos = self.trajectory[-1, :4]
        ok = self.cycrad() / self.field.lengthscale(tpos) < P["epss"]
        if self.field.static or not ok:
            return ok
        with np.errstate(divide="ignore"):
            return self.cycper() / self.field.timescale(tpos) < P["epst"]

    # ---- tracing -----------------------------------------------------------------
    def advance(self, delta, eom="TaoChanBrizardEOM"):
        """Advance by ``delta`` seconds on the GPU; raises NonAdiabatic when
        ``check_adiabaticity`` is set and the motion stops being adiabatic."""
        GuidingCenter._advance_many([self], delta, eom=eom)

    def _dt(self):
        from . import params as P
        if P["GCtimestep"] != 0:
            return P["GCtimestep"]
        return self.bounceperiod() / P["bounceresolution"]

    @staticmethod
    def _advance_many(gs, delta, eom="TaoChanBrizardEOM", device_dt=False, **kw):
        """Advance GuidingCenters that share one field.  Those that check
        adiabaticity run in a launch of their own, so that a switch point stops
        only them.  device_dt: output steps bounceperiod()/bounceresolution from
        the device half-bounce kernel (within the reference's 1e-4 quadrature
        tolerance) instead of one host quadrature per tracer."""
        from . import NonAdiabatic
        if eom not in _EOMS:
            raise ValueError("unknown eom %r" % (eom,))
        raise_na = False
        for check in (True, False):
            grp = [g for g in gs if bool(g.check_adiabaticity) == check]
            if grp:
                raise_na |= GuidingCenter._advance_group(grp, delta, eom, check, device_dt, **kw)
        if raise_na:
            raise NonAdiabatic

    @staticmethod
    def _batch_dt(gs, device_dt):
        from . import batch
        from . import params as P
        if P["GCtimestep"] != 0:
            return np.full(len(gs), float(P["GCtimestep"]))
        if device_dt and gs[0].field._model in _BOUNCE_MODELS:
            last = np.stack([g.trajectory[-1] for g in gs])
            dt, _ = batch.gc_bounce_dt(gs[0].field, last, np.array([g.mass for g in gs], float),
                                       np.array([g.mu for g in gs], float), params=P)
            return dt
        return np.array([g._dt() for g in gs], dtype=float)

    @staticmethod
    def _advance_group(gs, delta, eom, check, device_dt, **kw):
        from . import batch
        from . import params as P
        last = np.stack([g.trajectory[-1] for g in gs])
        dt = GuidingCenter._batch_dt(gs, device_dt)
        res = batch.advance_gcs(gs[0].field, last, np.array([g.mass for g in gs], float),
                                np.array([g.charge for g in gs], float),
                                np.array([g.mu for g in gs], float),
                                np.array([g.v for g in gs], float), dt, delta, params=P,
                                eom=eom, check=check, **kw)
        raise_na = False
        for i, g in enumerate(gs):
            rows = res.rows[i]
            if len(rows):
                g.trajectory = np.vstack((g.trajectory, rows))
            g.tcur = float(res.tcur[i])
            st = res.status[i]
            if st == 2:
                if g.check_adiabaticity:
                    raise_na = True
            elif st < 0:
                import warnings
                from ._native import STATUS_TEXT
                warnings.warn("GuidingCenter.advance: %s" % STATUS_TEXT.get(int(st), st),
                              stacklevel=4)
        return raise_na

    # ---- accessors (rapt/GuidingCenter.py:460-624) ---------------------------------------
    def gett(self):
        return self.trajectory[:, 0]

    def getx(self):
        return self.trajectory[:, 1]

    def gety(self):
        return self.trajectory[:, 2]

    def getz(self):
        return self.trajectory[:, 3]

    def getpp(self):
        return self.trajectory[:, 4]

    def getr(self):
        return np.sqrt(self.getx() ** 2 + self.gety() ** 2 + self.getz() ** 2)

    def gettheta(self):
        return np.arctan2(self.gety(), self.getx())

    def getphi(self):
        return np.arccos(self.getz() / self.getr())

    def getB(self):
        from . import _native as N
        return N.context().field_eval(self.field.descriptor(), "magB",
                                      self.trajectory[:, :4])[:, 0]

    def getgamma(self):
        mc = self.mass * c
        pp = self.trajectory[:, 4]
        return np.sqrt(1 + 2 * self.mu * self.getB() / (mc * c) + (pp / mc) ** 2)

    def getp(self):
        g = self.getgamma()
        pp = self.trajectory[:, 4]
        nr = np.sqrt(2 * self.mass * self.mu * self.getB() + pp ** 2)
        return np.where(g - 1 < 1e-6, nr, self.mass * c * np.sqrt(g ** 2 - 1))

    def getv(self):
        return self.getp() / self.getgamma() / self.mass

    def getBm(self):
        mc = self.mass * c
        g = self.getgamma()
        B = self.getB()
        pp = self.trajectory[:, 4]
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = B / (1 - (pp / mc) ** 2 / ((g - 1) * (g + 1)))
        return np.where(g - 1 < 1e-6, B + 0.5 * pp ** 2 / (self.mu * self.mass), rel)

    def getke(self):
        g = self.getgamma()
        pp = self.trajectory[:, 4]
        nr = self.mu * self.getB() + 0.5 * pp ** 2 / self.mass
        return np.where(g - 1 < 1e-6, nr, (g - 1) * self.mass * c * c)

    def cycrad(self):
        t, r, pp = self.trajectory[-1, 0], self.trajectory[-1, 1:4], self.trajectory[-1, 4]
        Bmag = self.field.magB(self.trajectory[-1, :4])
        g = np.sqrt(1 + 2 * self.mu * Bmag / (self.mass * c * c) + (pp / self.mass / c) ** 2)
        if g - 1 < 1e-6:
            vp = pp / self.mass
            v = np.sqrt(2 * self.mu * Bmag / self.mass + vp ** 2)
        else:
            vp = pp / self.mass / g
            v = c * np.sqrt(1 - 1 / g ** 2)
        return ru.cyclotron_radius2(t, r, vp, v, self.field, self.mass, self.charge)

    def cycper(self):
        # keeps the reference's unnormalised pp**2 term (rapt/GuidingCenter.py:535)
        t, r, pp = self.trajectory[-1, 0], self.trajectory[-1, 1:4], self.trajectory[-1, 4]
        Bmag = self.field.magB(self.trajectory[-1, :4])
        g = np.sqrt(1 + 2 * self.mu * Bmag / (self.mass * c * c) + pp ** 2)
        if g - 1 < 1e-6:
            vp = pp / self.mass
            v = np.sqrt(2 * self.mu * Bmag / self.mass + vp ** 2)
        else:
            v = c * np.sqrt(1 - 1 / g ** 2)
        return ru.cyclotron_period2(t, r, v, self.field, self.mass, self.charge)

    # ---- equations of motion (rapt/GuidingCenter.py:329-395), on the device ------------
    def _eom(self, eom, t, Y):
        from . import _native as N
        from . import params as P
        ty = np.r_[t, np.asarray(Y, dtype=float)[:4]][None, :]
        return N.context().rhs_eval(self.field.descriptor(), N.MODE_GC, ty,
                                    [self.mass, self.charge, self.mu, self.v], eom=eom,
                                    enforce_equatorial=P["enforce equatorial"])[0]

    def _TaoChanBrizardEOM(self, t, Y):
        """d(x, y, z, ppar)/dt in the Tao-Chan-Brizard equations (E fields and
        time-dependent fields allowed)."""
        return self._eom("TaoChanBrizardEOM", t, Y)

    def _BrizardChanEOM(self, t, Y):
        """d(x, y, z, ppar)/dt in the Brizard-Chan equations (static B)."""
        return self._eom("BrizardChanEOM", t, Y)

    def _NorthropTellerEOM(self, t, Y):
        """d(x, y, z, ppar)/dt in the Northrop-Teller equations."""
        return self._eom("NorthropTellerEOM", t, Y)

    def bounceperiod(self):
        """Bounce period at the current position (rapt/GuidingCenter.py:593-606)."""
        from . import flutils
        tpos, ppar = self.trajectory[-1, 0:4], self.trajectory[-1, 4]
        Bmag = self.field.magB(tpos)
        g = np.sqrt(1 + 2 * self.mu * Bmag / (self.mass * c * c) + (ppar / (self.mass * c)) ** 2)
        if g - 1 < 1e-6:
            p = np.sqrt(2 * self.mass * self.mu * Bmag + ppar ** 2)
            v = p / self.mass
            Bmirror = p ** 2 / (2 * self.mass * self.mu)
        else:
            p = self.mass * c * np.sqrt((g + 1) * (g - 1))
            Bmirror = p ** 2 / ((p - ppar) * (p + ppar)) * Bmag
            v = p / self.mass / g
        return flutils.bounceperiod(tpos, self.field, Bmirror, v)

    def geteye(self, step=1):
        """(time, second invariant) pairs along the trajectory, every `step` rows
        (rapt/GuidingCenter.py:608-624); I on the device (rapt_eye)."""
        from . import _native as N
        from . import params
        Bm = self.getBm()
        rows = self.trajectory[::step]
        res, st = N.context().eye(self.field.descriptor(), rows[:, :4], Bm[::step],
                                  params["fieldlineresolution"])
        bad = np.nonzero(st != 1)[0]
        if bad.size:
            raise RuntimeError("geteye: field-line integral failed at row %d (status %d)"
                               % (bad[0] * step, st[bad[0]]))
        return np.array(list(zip(rows[:, 0], res)))
