"""Full-Lorentz tracer with the API of rapt/Particle.py.

``advance`` runs on the GPU (librapt_hip.so): DOP853 with scipy's call
semantics, one output row per cyclotron_period/cyclotronresolution, exactly
as rapt/Particle.py:270-309 drives scipy.  Construction, conversions and
accessors are host-side numpy as in the reference.  Deliberate differences
from the reference, all in code paths that crash or return wrong values
there: empty-argument detection does not compare arrays with ``[]``
(rapt/Particle.py:105), ``init`` raises ValueError for a wrong type
(:166), ``getv`` returns the speed (:427-431) and ``getB`` returns its array
(:456-461).
"""
import pickle

import numpy as np

from . import utils as ru
from .constants import c


def _given(x):
    return x is not None and np.size(x) > 0


class Particle:
    """A relativistic charged particle in given E and B fields.

    Parameters: pos (m), vel (m/s), t0 (s), mass (kg), charge (C), field.
    ``trajectory`` rows are (t, x, y, z, px, py, pz).
    """

    def __init__(self, pos=[], vel=[], t0=0, mass=None, charge=None, field=None):
        self.pos = np.array(pos)
        self.vel = np.array(vel)
        self.t0 = t0
        self.tcur = t0
        self.mass = mass
        self.charge = charge
        self.field = field
        self.trajectory = np.zeros((1, 7))
        self.check_adiabaticity = False
        if _given(pos) and _given(vel) and self.mass is not None:
            g = 1 / np.sqrt(1 - np.dot(self.vel, self.vel) / c ** 2)
            mom = self.mass * g * self.vel
            self.trajectory = np.concatenate(([self.tcur], self.pos, mom)).reshape(1, 7)

    # ---- conversions (rapt/Particle.py:111-228) --------------------------------
    def init(self, p, gyrophase=0):
        """Initialise from the last state of a Particle or GuidingCenter."""
        from .GuidingCenter import GuidingCenter
        if isinstance(p, Particle):
            mom = p.trajectory[-1, 4:]
            gm = np.sqrt(p.mass ** 2 + np.dot(mom, mom) / c ** 2)
            self.__init__(pos=p.trajectory[-1, 1:4], vel=p.trajectory[-1, 4:] / gm,
                          t0=p.trajectory[-1, 0], mass=p.mass, charge=p.charge, field=p.field)
        elif isinstance(p, GuidingCenter):
            B = p.field.magB(p.trajectory[-1, :4])
            gammasq = 1 + 2 * p.mu * B / (p.mass * c * c) + (p.trajectory[-1, 4] / p.mass / c) ** 2
            if np.sqrt(gammasq) - 1 < 1e-6:
                v = np.sqrt(2 * p.mu * B / p.mass + (p.trajectory[-1, 4] / p.mass) ** 2)
            else:
                v = c * np.sqrt(1 - 1 / gammasq)
            vpar = p.trajectory[-1, 4] / p.mass / np.sqrt(gammasq)
            pos, vel = ru.GCtoFP(self.tcur, p.trajectory[-1, 1:4], vpar, v, p.field, p.mass,
                                 p.charge, gyrophase)
            self.__init__(pos=pos, vel=vel, t0=p.trajectory[-1, 0], mass=p.mass,
                          charge=p.charge, field=p.field)
        else:
            raise ValueError("Particle or GuidingCenter objects required.")

    def setke(self, ke, unit="ev"):
        """Rescale the speed to kinetic energy ``ke`` (eV by default)."""
        assert ke > 0
        s = ru.speedfromKE(ke, self.mass, unit)
        mom = self.trajectory[-1, 4:]
        gm = np.sqrt(self.mass ** 2 + np.dot(mom, mom) / c ** 2)
        v = mom / gm
        v = v * (s / np.sqrt(np.dot(v, v)))
        self.__init__(self.pos, v, self.t0, self.mass, self.charge, self.field)

    def setpa(self, pa):
        """Set the pitch angle to ``pa`` degrees at constant energy."""
        tpos = self.trajectory[-1, 0:4]
        mom = self.trajectory[-1, 4:]
        gm = np.sqrt(self.mass ** 2 + np.dot(mom, mom) / c ** 2)
        v = mom / gm
        s = np.sqrt(np.dot(v, v))
        b = self.field.unitb(tpos)
        spar = np.dot(v, b)
        if abs(spar - s) < 1e-12:
            p = ru.getperp(b)
        else:
            vperp = v - spar * b
            p = vperp / np.sqrt(np.dot(vperp, vperp))
        w = s * np.sin(pa * np.pi / 180) * p + s * np.cos(pa * np.pi / 180) * b
        self.__init__(self.pos, w, self.t0, self.mass, self.charge, self.field)

    # ---- tracing (rapt/Particle.py:230-309) -------------------------------------
    def advance(self, delta):
        """Advance by ``delta`` seconds on the GPU; raises Adiabatic when
        ``check_adiabaticity`` is set and the motion becomes adiabatic."""
        Particle._advance_many([self], delta)

    @staticmethod
    def _advance_many(ps, delta, **kw):
        """Advance Particles that share one field.  Those that check
        adiabaticity run in a launch of their own, so that a switch point stops
        only them and the others reach t0 + delta as the reference traces them."""
        from . import Adiabatic
        raise_adiabatic = False
        for check in (True, False):
            grp = [p for p in ps if bool(p.check_adiabaticity) == check]
            if grp:
                raise_adiabatic |= Particle._advance_group(grp, delta, check, **kw)
        if raise_adiabatic:
            raise Adiabatic

    @staticmethod
    def _advance_group(ps, delta, check, **kw):
        from . import batch
        from . import params as P
        last = np.stack([p.trajectory[-1] for p in ps])
        mass = np.array([p.mass for p in ps], dtype=float)
        charge = np.array([p.charge for p in ps], dtype=float)
        res = batch.advance_particles(ps[0].field, last, mass, charge, delta, params=P,
                                      check=check, **kw)
        raise_adiabatic = False
        for i, p in enumerate(ps):
            rows = res.rows[i]
            if len(rows):
                p.trajectory = np.vstack((p.trajectory, rows))
            p.tcur = float(res.tcur[i])
            st = res.status[i]
            if st == 2:
                if p.check_adiabaticity:
                    raise_adiabatic = True
            elif st < 0:
                import warnings
                from ._native import STATUS_TEXT
                warnings.warn("Particle.advance: %s" % STATUS_TEXT.get(int(st), st),
                              stacklevel=4)
        return raise_adiabatic

    # ---- persistence --------------------------------------------------------------
    def save(self, filename):
        with open(filename, "wb") as f:
            pickle.dump(self, f)

    def load(self, filename):
        """Load a pickled Particle written by ``save`` (trusted files only)."""
        with open(filename, "rb") as f:
            p = pickle.load(f)
        self.__dict__.update(p.__dict__)

    # ---- adiabaticity (rapt/Particle.py:345-384) ------------------------------------
    def isadiabatic(self):
        from . import params as P
        tpos = self.trajectory[-1, :4]
        ok = self.cycrad() / self.field.lengthscale(tpos) < P["epss"]
        if self.field.static or not ok:
            return ok
        with np.errstate(divide="ignore"):
            return self.cycper() / self.field.timescale(tpos) < P["epst"]

    # ---- accessors (rapt/Particle.py:386-494) ------------------------------------------
    def gett(self):
        return self.trajectory[:, 0]

    def getx(self):
        return self.trajectory[:, 1]

    def gety(self):
        return self.trajectory[:, 2]

    def getz(self):
        return self.trajectory[:, 3]

    def getpx(self):
        return self.trajectory[:, 4]

    def getpy(self):
        return self.trajectory[:, 5]

    def getpz(self):
        return self.trajectory[:, 6]

    def getp(self):
        return np.sqrt(self.getpx() ** 2 + self.getpy() ** 2 + self.getpz() ** 2)

    def getgamma(self):
        psq = np.sum(self.trajectory[:, 4:7] ** 2, axis=1)
        return np.sqrt(1 + psq / (self.mass * c) ** 2)

    def getvx(self):
        return self.getpx() / self.getgamma() / self.mass

    def getvy(self):
        return self.getpy() / self.getgamma() / self.mass

    def getvz(self):
        return self.getpz() / self.getgamma() / self.mass

    def getv(self):
        return self.getp() / self.getgamma() / self.mass

    def getr(self):
        return np.sqrt(self.getx() ** 2 + self.gety() ** 2 + self.getz() ** 2)

    def gettheta(self):
        return np.arctan2(self.gety(), self.getx())

    def getphi(self):
        return np.arccos(self.getz() / self.getr())

    def getke(self):
        g = self.getgamma()
        t = self.trajectory
        ke_nr = 0.5 * (t[:, 4] ** 2 + t[:, 5] ** 2 + t[:, 6] ** 2) / self.mass
        mc2 = self.mass * c * c
        return np.where(g - 1 < 1e-6, ke_nr, (g - 1) * mc2)

    def getB(self):
        from . import _native as N
        return N.context().field_eval(self.field.descriptor(), "magB",
                                      self.trajectory[:, :4])[:, 0]

    def _gc_rows(self):
        """Guiding centre, v_par, speed and mu of every row, on the device
        (rapt_particle_gc: utils.guidingcenter and utils.magnetic_moment per row)."""
        from . import _native as N
        out, st = N.context().particle_gc(self.field.descriptor(), self.trajectory, self.mass,
                                          self.charge)
        bad = np.nonzero(st == N.GC_NOCONV)[0]
        if bad.size:
            raise RuntimeError("guidingcenter: no convergence to 1e-3 in 20 iterations at "
                               "row %d" % bad[0])
        return out

    def guidingcenter(self):
        """(X, Y, Z, v_par, speed) of every row (rapt/Particle.py:463-472)."""
        return self._gc_rows()[:, :5]

    def mu(self):
        """First invariant of every row (rapt/Particle.py:473-482)."""
        return self._gc_rows()[:, 5]

    def cycrad(self):
        t, r, mom = self.trajectory[-1, 0], self.trajectory[-1, 1:4], self.trajectory[-1, 4:]
        gm = np.sqrt(self.mass ** 2 + np.dot(mom, mom) / c ** 2)
        return ru.cyclotron_radius(t, r, mom / gm, self.field, self.mass, self.charge)

    def cycper(self):
        t, r, mom = self.trajectory[-1, 0], self.trajectory[-1, 1:4], self.trajectory[-1, 4:]
        gm = np.sqrt(self.mass ** 2 + np.dot(mom, mom) / c ** 2)
        return ru.cyclotron_period(t, r, mom / gm, self.field, self.mass, self.charge)
