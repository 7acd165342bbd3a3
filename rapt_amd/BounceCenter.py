"""Bounce-averaged drift tracer, API of rapt/BounceCenter.py.

Construction and accessors follow rapt/BounceCenter.py:74-316.  Its drift
right-hand side needs 5-7 field-line integrals per evaluation
(rapt/BounceCenter.py:235-244); that device kernel is not built yet, so
``advance`` raises NotImplementedError rather than running on the CPU.
"""
import pickle

import numpy as np

from . import utils as ru


class BounceCenter:
    def __init__(self, pos=[], v=None, t0=0, pa=None, mass=None, charge=None, field=None):
        self.pos = pos
        self.v = v
        self.t0 = t0
        self.tcur = t0
        self.mass = mass
        self.charge = charge
        self.field = field
        self.isequatorial = False
        if field is not None and not field.static:
            raise RuntimeError("BounceCenter does not work with nonstatic fields or electric "
                               "fields.")
        self.pa = pa
        if np.size(pos) > 0 and v is not None and pa is not None:
            self.trajectory = np.concatenate(([t0], pos)).reshape(1, 4)
            # the reference passes pa (degrees) to np.cos directly
            # (rapt/BounceCenter.py:114); kept as written
            self.mu = ru.magnetic_moment(t0, pos, self.v * np.cos(self.pa), self.v, field, mass)
            assert self.mu > 0

    def advance(self, delta):
        raise NotImplementedError("BounceCenter.advance: the GPU bounce-averaged drift kernel "
                                  "is not built yet")

    def save(self, filename):
        with open(filename, "wb") as f:
            pickle.dump(self, f)

    def load(self, filename):
        """Load a pickled BounceCenter written by ``save`` (trusted files only)."""
        with open(filename, "rb") as f:
            p = pickle.load(f)
        self.__dict__.update(p.__dict__)

    def gett(self):
        return self.trajectory[:, 0]

    def getx(self):
        return self.trajectory[:, 1]

    def gety(self):
        return self.trajectory[:, 2]

    def getz(self):
        return self.trajectory[:, 3]

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
