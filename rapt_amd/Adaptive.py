"""Adaptive Particle/GuidingCenter tracer with the API of rapt/Adaptive.py.

The switching loop of rapt/Adaptive.py:187-222 runs inside the GPU kernel:
every lane tests adiabaticity after each output row, converts between
full-orbit and guiding-centre state (utils.guidingcenter / utils.GCtoFP, as
Particle.init and GuidingCenter.init do) and restarts the integrator.  The
returned rows are split back into the ``trajlist`` of Particle and
GuidingCenter segments.  As in the reference, guiding-centre segments use
``params['GCtimestep']``, which must be non-zero here.

Construction decides the first mode on the host, as Adaptive.__init__ does
(rapt/Adaptive.py:70-104), so ``trajlist`` and the accessors work before
any advance.  The first ``advance`` makes that decision again on the device
and replaces ``trajlist`` with the device's segments; later calls continue
the last tracer of ``trajlist`` (a second call may be made at any time, with
the reference's loop variable t restarting at 0 and becoming the new tracer's
absolute tcur after a switch).  Deliberate differences: a failed integration
segment ends the run with a warning instead of restarting it, and ``tcur``
follows the trajectory (the reference leaves it at t0).
"""
import pickle

import numpy as np

from .constants import c
from .GuidingCenter import GuidingCenter
from .Particle import Particle


class Adaptive:
    def __init__(self, pos=None, vel=None, t0=0, mass=None, charge=None, field=None):
        self.pos = np.array(pos)
        self.vel = np.array(vel)
        self.tcur = t0
        self.mass = mass
        self.charge = charge
        self.field = field
        self.p = Particle(pos, vel, t0, mass, charge, field)
        self.p.check_adiabaticity = True
        self.nswitch = 0
        self._start()

    def _start(self):
        """Adaptive.__init__ / setke / setpa (rapt/Adaptive.py:96-104, 140-186):
        a GuidingCenter when the particle is adiabatic, else the particle."""
        self._fresh = True
        self.trajlist = [self.p]
        if self.field is None or self.mass is None or np.size(self.pos) == 0:
            return
        if self.p.isadiabatic():
            g = GuidingCenter()
            g.init(self.p)
            g.check_adiabaticity = True
            self.trajlist = [g]

    def setke(self, ke, unit="ev"):
        """Rescale the speed to kinetic energy ``ke``; restarts the tracer
        (rapt/Adaptive.py:140-164)."""
        self.p.setke(ke, unit)
        self.p.check_adiabaticity = True
        self._restart_from_p()

    def setpa(self, pa):
        """Set the pitch angle (degrees) at constant energy; restarts the tracer
        (rapt/Adaptive.py:166-186)."""
        self.p.setpa(pa)
        self.p.check_adiabaticity = True
        self._restart_from_p()

    def _restart_from_p(self):
        row = self.p.trajectory[-1]
        mom = row[4:7]
        gm = np.sqrt(self.mass ** 2 + np.dot(mom, mom) / c ** 2)
        self.pos = row[1:4].copy()
        self.vel = mom / gm
        self.tcur = row[0]
        self._start()

    @staticmethod
    def _segments(rows):
        segs = []
        i = 0
        while i < len(rows):
            mode = rows[i, 0]
            j = i + 1
            while j < len(rows) and rows[j, 0] == mode:
                j += 1
            segs.append(rows[i:j])
            i = j
        return segs

    def _tracer(self, seg, p=None):
        if seg[0, 0] == 0:
            if p is None:
                p = Particle()
            p.mass, p.charge, p.field = self.mass, self.charge, self.field
            p.trajectory = seg[:, 1:8].copy()
            p.tcur = p.trajectory[-1, 0]
            p.check_adiabaticity = True
            return p
        g = GuidingCenter()
        g.mass, g.charge, g.field = self.mass, self.charge, self.field
        g.trajectory = seg[:, 1:6].copy()
        g.mu = seg[0, 6]
        g.v = self._gc_speed(g.trajectory[-1], g.mu)
        g.tcur = g.trajectory[-1, 0]
        g.check_adiabaticity = True
        return g

    def _gc_speed(self, row5, mu):
        """Speed of a guiding centre from its row, as GuidingCenter.init
        computes it (rapt/GuidingCenter.py:156-166)."""
        B = self.field.magB(row5[:4])
        g = np.sqrt(1 + 2 * mu * B / (self.mass * c * c) + (row5[4] / self.mass / c) ** 2)
        if g - 1 < 1e-6:
            return float(np.sqrt(2 * mu * B / self.mass + (row5[4] / self.mass) ** 2))
        return float(c * np.sqrt(1 - 1 / g ** 2))

    def advance(self, delta):
        """Advance by ``delta`` seconds (switching on the device).  Can be
        called many times; later calls continue the last tracer."""
        from . import batch
        if self._fresh:
            res = batch.advance_adaptive(self.field, self.pos.reshape(1, 3),
                                         self.vel.reshape(1, 3), self.p.t0, self.mass,
                                         self.charge, delta)
            rows = res.rows[0]
            segs = self._segments(rows)
            # the first Particle segment is self.p itself, as in the reference
            self.trajlist = [self._tracer(s, self.p if (k == 0 and s[0, 0] == 0) else None)
                             for k, s in enumerate(segs)]
            self._fresh = False
        else:
            cur = self.trajlist[-1]
            last = cur.trajectory[-1]
            if isinstance(cur, GuidingCenter):
                mode, mu, v = 1, cur.mu, self._gc_speed(last, cur.mu)
            else:
                mode, mu, v = 0, 0.0, 0.0
            res = batch.continue_adaptive(self.field, mode, last[None, :], self.mass,
                                          self.charge, mu, v, delta)
            rows = res.rows[0]
            segs = self._segments(rows)
            if segs and segs[0][0, 0] == mode:
                width = 7 if mode == 0 else 5
                cur.trajectory = np.vstack((cur.trajectory, segs[0][:, 1:1 + width]))
                cur.tcur = cur.trajectory[-1, 0]
                segs = segs[1:]
            self.trajlist.extend(self._tracer(s) for s in segs)
        self.nswitch += int(np.sum(np.diff(rows[:, 0]) != 0)) if len(rows) else 0
        self.tcur = float(res.tcur[0])
        st = res.status[0]
        if st < 0:
            import warnings
            from ._native import STATUS_TEXT
            warnings.warn("Adaptive.advance: %s" % STATUS_TEXT.get(int(st), st), stacklevel=2)

    def save(self, filename):
        with open(filename, "wb") as f:
            pickle.dump(self, f)

    def load(self, filename):
        """Load a pickled Adaptive written by ``save`` (trusted files only)."""
        with open(filename, "rb") as f:
            p = pickle.load(f)
        self.__dict__.update(p.__dict__)

    def _cat(self, name):
        if not self.trajlist:
            return np.array([])
        return np.concatenate([getattr(p, name)() for p in self.trajlist])

    def gett(self):
        return self._cat("gett")

    def getx(self):
        return self._cat("getx")

    def gety(self):
        return self._cat("gety")

    def getz(self):
        return self._cat("getz")

    def getr(self):
        return self._cat("getr")

    def getphi(self):
        return self._cat("getphi")

    def gettheta(self):
        return self._cat("gettheta")

    def getke(self):
        return self._cat("getke")
