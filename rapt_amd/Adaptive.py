"""Adaptive Particle/GuidingCenter tracer with the API of rapt/Adaptive.py.

The whole switching loop of rapt/Adaptive.py:187-222 runs inside the GPU
kernel: every lane tests adiabaticity after each output row, converts between
full-orbit and guiding-centre state (utils.guidingcenter / utils.GCtoFP, as
Particle.init and GuidingCenter.init do) and restarts the integrator.  The
returned rows are split back into the ``trajlist`` of Particle and
GuidingCenter segments.  As in the reference, guiding-centre segments use
``params['GCtimestep']``, which must be non-zero here.  A failed integration
segment ends the run with a warning instead of restarting it.
"""
import pickle

import numpy as np

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
        self.trajlist = None  # built on the first advance (mode decided on the device)
        self.nswitch = 0

    def _segments(self, rows):
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

    def _build(self, rows):
        out = []
        for seg in self._segments(rows):
            if seg[0, 0] == 0:
                p = Particle()
                p.mass, p.charge, p.field = self.mass, self.charge, self.field
                p.trajectory = seg[:, 1:8].copy()
                p.tcur = p.trajectory[-1, 0]
                p.check_adiabaticity = True
                out.append(p)
            else:
                g = GuidingCenter()
                g.mass, g.charge, g.field = self.mass, self.charge, self.field
                g.trajectory = seg[:, 1:6].copy()
                g.mu = seg[0, 6]
                g.tcur = g.trajectory[-1, 0]
                g.check_adiabaticity = True
                out.append(g)
        return out

    def advance(self, delta):
        """Advance by ``delta`` seconds (switching on the device)."""
        from . import batch
        if self.trajlist is not None:
            raise NotImplementedError("Adaptive.advance continues only from the initial state")
        res = batch.advance_adaptive(self.field, self.pos.reshape(1, 3), self.vel.reshape(1, 3),
                                     self.p.t0, self.mass, self.charge, delta)
        rows = res.rows[0]
        self._rows = rows
        self.nswitch = int(np.sum(np.diff(rows[:, 0]) != 0)) if len(rows) else 0
        self.trajlist = self._build(rows)
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
