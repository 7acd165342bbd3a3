"""Field-line tracer with the interface of rapt/fieldline.py:12-133 (host side).

Integrates dr/ds = +-b with rapt_amd.rkf in chunks of length ds = 1/(curvature
* fieldlineresolution), forward with tolerance 1e-3 and backward with 1e-4 as
the reference does (rapt/fieldline.py:66,88), until |B| leaves [Bmin, Bmax].
Used to set up the guiding-centre output step (bounce period); tracing itself
runs on the GPU.
"""
import numpy as np
from scipy.integrate import odeint

from .rkf import rkf


class Fieldline:
    def __init__(self, tpos, field, ds=0, stopcond=None, Bmin=None, Bmax=None, maxchunks=100000):
        from . import params
        self.time = tpos[0]
        self.initpt = np.concatenate(([0], tpos[1:]))
        self.curve = np.zeros((1, 4))
        self.curve[0, :] = self.initpt
        self.field = field
        self.Bmin = Bmin
        self.Bmax = Bmax
        self.solver = params["flsolver"]
        self.maxchunks = maxchunks
        self.ds = self._estimate_step() if ds == 0 else ds
        self.stopcond = stopcond if stopcond is not None else (lambda b, c: False)

    def _estimate_step(self):
        from . import params
        k = self.field.curvature(np.concatenate(([self.time], self.curve[-1, 1:])))
        return 1 / k / params["fieldlineresolution"]

    def _half(self, sign, tol):
        def deriv(Y, t=0):
            out = np.zeros(4)
            out[0] = sign
            out[1:] = sign * self.field.unitb(np.concatenate([[self.time], Y[1:]]))
            return out
        line = np.zeros((1, 4))
        line[0] = self.initpt
        nchunk = 0
        while not self.stopcond(line[-1, :], self.field):
            nchunk += 1
            if nchunk > self.maxchunks:
                raise RuntimeError("field line did not reach the stopping field in %d chunks"
                                   % self.maxchunks)
            if self.solver == "odeint":
                r = odeint(deriv, line[-1, :], [0, self.ds], rtol=1e-6, atol=1e-6)
            else:
                try:
                    _, r = rkf(deriv, 0, self.ds, line[-1, :], tol=tol, hmax=self.ds, hmin=1e-6)
                except RuntimeError:
                    break
            line = np.concatenate((line, r[1:, :]))
            B = self.field.magB(np.concatenate([[self.time], line[-1, 1:]]))
            if self.Bmin is not None and B < self.Bmin:
                break
            if self.Bmax is not None and B > self.Bmax:
                break
        return line

    def trace(self):
        forward = self._half(1.0, 1e-3)
        backward = self._half(-1.0, 1e-4)
        self.curve = np.concatenate((backward[:0:-1, :], forward))

    def reset(self):
        self.curve = np.zeros((1, 4))

    def gets(self):
        return self.curve[:, 0]

    def getx(self):
        return self.curve[:, 1]

    def gety(self):
        return self.curve[:, 2]

    def getz(self):
        return self.curve[:, 3]

    def getB(self):
        return np.array([self.field.magB(np.concatenate(([self.time], row[1:])))
                         for row in self.curve])

    def getr(self):
        return np.sqrt(self.getx() ** 2 + self.gety() ** 2 + self.getz() ** 2)

    def gettheta(self):
        return np.arctan2(self.gety(), self.getx())

    def getphi(self):
        return np.arccos(self.getz() / self.getr())
