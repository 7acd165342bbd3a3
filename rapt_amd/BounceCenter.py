"""Bounce-averaged drift tracer, API of rapt/BounceCenter.py.

Construction and accessors follow rapt/BounceCenter.py:74-316.  ``advance``
runs on the GPU (rapt_bounce_center_advance): dopri5 on the drift
gamma m v^2 / (q S_b B^2) grad I x B, whose S_b and grad I are field-line
integrals evaluated on the device for every stage
(rapt/BounceCenter.py:206-251).  The output step is params['BCtimestep']
times the bounce period at the current point (host restatement of
flutils.bounceperiod, as the reference computes it once per call).

With ``isequatorial = True`` the drift is the gradient plus inertial drift
of a centre with no parallel motion, ``_v_grad + _v_inert``
(rapt/BounceCenter.py:154-204, 237-238; rapt_bounce_center_equatorial).  The
reference passes those two the 3-vector ``Y[:4]`` = (x, y, z), which every
field's ``t, x, y, z = tpos`` rejects, so its branch raises ValueError on
the first step; here they take (t, x, y, z), the tpos its deriv builds.
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
        """Advance for delta seconds (rapt/BounceCenter.py:206-251).  Rows are
        appended at the times np.arange(tcur, tcur + delta, dt), each holding
        the state one output step later, as the reference records them."""
        from . import _native as N
        from . import flutils
        from . import params
        from .constants import c
        v = self.v
        gamma = 1.0 / np.sqrt(1 - (v / c) ** 2)
        Bm = self.mass * gamma ** 2 * v ** 2 / (2 * self.mu)
        bp = flutils.bounceperiod(self.trajectory[-1, :4], self.field, Bm, v)
        dt = params["BCtimestep"] * bp
        nrow = len(np.arange(self.tcur, self.tcur + delta, dt))
        rtol, atol = params["solvertolerances"]
        if self.isequatorial:
            rows, _, st, _ = N.context().bounce_center_equatorial(
                self.field.descriptor(), self.trajectory[-1, :4],
                [self.mu, self.mass, self.charge, gamma], dt, self.tcur, nrow, rtol, atol)
        else:
            rows, _, st, _ = N.context().bounce_center_advance(
                self.field.descriptor(), self.trajectory[-1, :4], Bm,
                gamma * self.mass * v * v / self.charge, dt, self.tcur, nrow, rtol, atol,
                params["eyegradientstep"], params["fieldlineresolution"])
        if st[0] != 1:
            import warnings
            warnings.warn("BounceCenter.advance stopped early (status %d)" % st[0])
        r = rows[0]
        if st[0] != 1:
            r = r[:np.count_nonzero(np.any(r != 0, axis=1))]
        self.trajectory = np.vstack((self.trajectory, r))
        self.tcur = self.trajectory[-1, 0]

    def _v_grad(self, tpos):
        """Gradient drift mu (B x grad|B|) / (q gamma |B|^2) at tpos = (t, x, y, z)
        (rapt/BounceCenter.py:154-172), from the device's field evaluation."""
        from . import _native as N
        from .constants import c
        gamma = 1.0 / np.sqrt(1 - (self.v / c) ** 2)
        ctx, d = N.context(), self.field.descriptor()
        tp = np.asarray(tpos, dtype=np.float64)[None, :4]
        Bvec = ctx.field_eval(d, "B", tp)[0]
        gB = ctx.field_eval(d, "gradB", tp)[0]
        return self.mu * np.cross(Bvec, gB) / (self.charge * gamma * np.dot(Bvec, Bvec))

    def _v_inert(self, tpos):
        """Inertial drift (v_g . grad) v_g at tpos = (t, x, y, z)
        (rapt/BounceCenter.py:174-204)."""
        from . import _native as N
        tpos = np.asarray(tpos, dtype=np.float64)[:4]
        vg = self._v_grad(tpos)
        vgdir = np.zeros(4)
        vgmag = np.sqrt(np.dot(vg, vg))
        vgdir[0] = tpos[0]
        vgdir[1:] = vg / vgmag
        d = self.field.gradientstepsize
        v1 = self._v_grad(tpos - d * vgdir)
        v2 = self._v_grad(tpos + d * vgdir)
        Bvec = N.context().field_eval(self.field.descriptor(), "B", tpos[None, :])[0]
        return self.mass * vgmag / (self.charge * np.dot(Bvec, Bvec) * 2 * d) * \
            np.cross(Bvec, (v2 - v1))

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
