"""Field models: the plugin API of rapt/fields.py, with a device descriptor.

Each model keeps the reference's host methods (``B``, ``E``, ``unitb``,
``magB``, ``gradB``, ``jacobianB``, ``curvature``, ``curlb``, ``dBdt``,
``dbdt``, ``lengthscale``, ``timescale``; all take ``tpos = [t, x, y, z]``)
so user code that probes a field keeps working.  Tracing never calls them:
``advance()`` hands the model's ``descriptor()`` (model id + constants) to the
HIP library, whose kernels evaluate the same formulas on the device.  A user
subclass that overrides ``B``/``E`` in Python (the reference's way to define
a new field) is traced: ``descriptor()`` runs the overrides once on symbolic
coordinates and hands the device the resulting register program
(``fieldexpr``; model ``FIELD_EXPR``).  Overrides that branch on field values
or leave numpy's elementwise math cannot be traced and raise
NotImplementedError.

Reference: rapt/fields.py:5-814.  ``Grid`` (rapt/fields.py:513-814) keeps
every parsed snapshot resident in device memory and is interpolated on the
device (trilinear in space, linear in time), see ``Grid``.
"""
import itertools

import numpy as np

from . import fieldexpr
from .constants import B0, Re

# model ids shared with include/rapt_hip.h (enum rapt_field_model)
FIELD_ZERO = 0
FIELD_EARTH_DIPOLE = 1
FIELD_DOUBLE_DIPOLE = 2
FIELD_UNIFORM_BZ = 3
FIELD_UNIFORM_CROSSED_EB = 4
FIELD_VAR_EARTH_DIPOLE = 5
FIELD_PARABOLIC = 6
FIELD_CURRENT_SHEET = 7
FIELD_DIPOLE_UNIFORM = 8
FIELD_GRID = 9
FIELD_EXPR = 10


class _Field:
    """Base class (rapt/fields.py:5-280).  Subclasses override ``B``/``E``."""

    # central-difference curl stencil (rapt/fields.py:33-36)
    _M1 = np.array([[0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0, -1, 0, 0, 1, 0],
                    [0, 0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0, 0],
                    [0, 1, 0, 0, -1, 0, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0]])
    _model = None

    def __init__(self):
        self.gradientstepsize = 1e-6
        self.timederivstepsize = 1e-3
        self.static = True

    # ---- device descriptor -------------------------------------------------
    def _params(self):
        return ()

    def descriptor(self):
        """Model id and constants for the HIP kernels.

        A built-in model gives its id and constants.  A subclass whose ``B``
        or ``E`` is Python code of its own (the reference's plugin contract,
        rapt/fields.py:5-74) is traced into a register program at every call,
        so attributes changed since the last call take effect as they would
        in the reference: model ``FIELD_EXPR`` with the program under
        ``"expr"``.  Raises NotImplementedError when the overrides cannot be
        traced.
        """
        common = dict(static=bool(self.static), gstep=float(self.gradientstepsize),
                      tstep=float(self.timederivstepsize))
        builtin = (self._model is not None and
                   type(self).B is _MODEL_B.get(self._model, type(self).B) and
                   type(self).E is _MODEL_E.get(self._model, type(self).E))
        if not builtin:
            try:
                prog = fieldexpr.trace(self)
            except NotImplementedError as e:
                raise NotImplementedError("%s cannot run on the device: %s"
                                          % (type(self).__name__, e)) from None
            return dict(model=FIELD_EXPR, params=(), expr=prog, **common)
        return dict(model=self._model, params=tuple(float(v) for v in self._params()),
                    **common)

    # ---- host API (rapt/fields.py:43-280) -----------------------------------
    def B(self, tpos):
        return np.zeros(3)

    def E(self, tpos):
        return np.zeros(3)

    def unitb(self, tpos):
        Bvec = self.B(tpos)
        return Bvec / np.sqrt(np.dot(Bvec, Bvec))

    def magB(self, tpos):
        Bvec = self.B(tpos)
        return np.sqrt(np.dot(Bvec, Bvec))

    def gradB(self, tpos):
        d = self.gradientstepsize
        tpos = np.asarray(tpos, dtype=float)
        return np.array([
            (self.magB(tpos + (0, d, 0, 0)) - self.magB(tpos - (0, d, 0, 0))) / (2 * d),
            (self.magB(tpos + (0, 0, d, 0)) - self.magB(tpos - (0, 0, d, 0))) / (2 * d),
            (self.magB(tpos + (0, 0, 0, d)) - self.magB(tpos - (0, 0, 0, d))) / (2 * d)])

    def jacobianB(self, tpos):
        d = self.gradientstepsize
        tpos = np.asarray(tpos, dtype=float)
        out = np.zeros((3, 3))
        out[:, 0] = (self.B(tpos + (0, d, 0, 0)) - self.B(tpos - (0, d, 0, 0))) / (2 * d)
        out[:, 1] = (self.B(tpos + (0, 0, d, 0)) - self.B(tpos - (0, 0, d, 0))) / (2 * d)
        out[:, 2] = (self.B(tpos + (0, 0, 0, d)) - self.B(tpos - (0, 0, 0, d))) / (2 * d)
        return out

    def curvature(self, tpos):
        # rapt/fields.py:155-174, including its elementwise use of np.dot(gB, |B|)
        Bvec = self.B(tpos)
        Bm = np.sqrt(np.dot(Bvec, Bvec))
        gB = self.gradB(tpos)
        gBperp = gB - (gB * Bm / Bm ** 2) * Bvec
        return np.sqrt(np.dot(gBperp, gBperp)) / Bm

    def curlb(self, tpos):
        d = self.gradientstepsize
        tpos = np.asarray(tpos, dtype=float)
        beta = np.concatenate((
            self.unitb(tpos + (0, d, 0, 0)), self.unitb(tpos - (0, d, 0, 0)),
            self.unitb(tpos + (0, 0, d, 0)), self.unitb(tpos - (0, 0, d, 0)),
            self.unitb(tpos + (0, 0, 0, d)), self.unitb(tpos - (0, 0, 0, d))))
        return np.dot(self._M1, beta) / (2 * d)

    def dBdt(self, tpos):
        if self.static:
            return 0
        d = self.timederivstepsize
        tpos = np.asarray(tpos, dtype=float)
        return (self.magB(tpos + [d, 0, 0, 0]) - self.magB(tpos - [d, 0, 0, 0])) / d / 2

    def dbdt(self, tpos):
        if self.static:
            return 0
        d = self.timederivstepsize
        tpos = np.asarray(tpos, dtype=float)
        return (self.unitb(tpos + [d, 0, 0, 0]) - self.unitb(tpos - [d, 0, 0, 0])) / d / 2

    def lengthscale(self, tpos):
        return self.magB(tpos) / np.max(abs(self.jacobianB(tpos)))

    def timescale(self, tpos):
        if self.static:
            return None
        with np.errstate(divide="ignore"):
            return self.magB(tpos) / abs(self.dBdt(tpos))


class EarthDipole(_Field):
    """Tilt-free static Earth dipole (rapt/fields.py:282-317)."""
    _model = FIELD_EARTH_DIPOLE

    def __init__(self, B0=B0):
        _Field.__init__(self)
        self.gradientstepsize = Re * 1e-6
        self._coeff = -3 * B0 * Re ** 3

    def _params(self):
        return (self._coeff,)

    def B(self, tpos):
        t, x, y, z = tpos
        r2 = x * x + y * y + z * z
        return self._coeff / pow(r2, 2.5) * np.array([x * z, y * z, (z * z - r2 / 3)])


class DoubleDipole(_Field):
    """Earth dipole plus an image dipole at x = distance (rapt/fields.py:319-362)."""
    _model = FIELD_DOUBLE_DIPOLE

    def __init__(self, B0=B0, distance=20 * Re, imagestrength=1):
        _Field.__init__(self)
        self.gradientstepsize = Re / 1000
        self._dd = distance
        assert imagestrength >= 1
        self._k = imagestrength
        self._coeff = -B0 * Re ** 3

    def _params(self):
        return (self._coeff, self._dd, self._k)

    def B(self, tpos):
        t, x, y, z = tpos
        B1 = np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / pow(x * x + y * y + z * z, 5.0 / 2.0)
        x -= self._dd
        B2 = self._k * np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / pow(x * x + y * y + z * z, 5.0 / 2.0)
        return self._coeff * (B1 + B2)


class UniformBz(_Field):
    """B = (0, 0, Bz) (rapt/fields.py:364-390)."""
    _model = FIELD_UNIFORM_BZ

    def __init__(self, Bz=1):
        _Field.__init__(self)
        self.Bz = Bz

    def _params(self):
        return (self.Bz,)

    def B(self, tpos):
        return np.array((0, 0, self.Bz))


class UniformCrossedEB(UniformBz):
    """E = (0, Ey, 0), B = (0, 0, Bz) (rapt/fields.py:392-427)."""
    _model = FIELD_UNIFORM_CROSSED_EB

    def __init__(self, Ey=1, Bz=1):
        UniformBz.__init__(self)
        self.static = False
        self.Ey = Ey
        self.Bz = Bz

    def _params(self):
        return (self.Bz, self.Ey)

    def E(self, tpos):
        return np.array((0, self.Ey, 0))


class VarEarthDipole(_Field):
    """Sinusoidally varying dipole (rapt/fields.py:429-470)."""
    _model = FIELD_VAR_EARTH_DIPOLE

    def __init__(self, amp=0.1, period=10):
        _Field.__init__(self)
        self.gradientstepsize = Re / 1000
        self.static = False
        self._amp = amp
        self._period = period

    def _params(self):
        return (-B0 * Re ** 3, self._amp, self._period)

    def B(self, tpos):
        t, x, y, z = tpos
        return -B0 * Re ** 3 * (1 + self._amp * np.sin(2 * np.pi * t / self._period)) * np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / pow(x * x + y * y + z * z, 5.0 / 2.0)


class Parabolic(_Field):
    """Parabolic tail field (rapt/fields.py:472-511), kept as written: for
    |z| > 1 the x-component is sign(z) times the module constant B0."""
    _model = FIELD_PARABOLIC

    def __init__(self, B0=10.0, Bn=1.0, d=0.2):
        _Field.__init__(self)
        self.B0 = B0
        self.Bn = Bn
        self.d = d

    def _params(self):
        return (self.B0, self.Bn, self.d, B0)

    def B(self, tpos):
        z = tpos[3]
        if abs(z) <= 1.0:
            return np.array([self.B0 * z / self.d, 0, self.Bn])
        return np.array([np.sign(z) * B0, 0, self.Bn])


class CurrentSheet(_Field):
    """Harris-type current sheet with a normal component and a uniform E.

    B = (B0 tanh(z/lam), 0, Bn), E = (Ex, Ey, Ez).  ``static`` is False
    whenever E is non-zero, following the reference's convention
    (rapt/fields.py:41).  A user-analytic model in the sense of the
    reference's override contract (rapt/fields.py:5-74).
    """
    _model = FIELD_CURRENT_SHEET

    def __init__(self, B0=20e-9, lam=0.5 * Re, Bn=1e-9, E=(0.0, 0.3e-3, 0.0),
                 gradientstepsize=None):
        _Field.__init__(self)
        self.B0 = B0
        self.lam = lam
        self.Bn = Bn
        self._E = np.array(E, dtype=float)
        self.static = not np.any(self._E != 0)
        # The base default (1e-6 m) is ill-conditioned for a length scale of
        # order lam; use lam/1000 unless told otherwise.
        self.gradientstepsize = lam / 1000 if gradientstepsize is None else gradientstepsize

    def _params(self):
        return (self.B0, self.lam, self.Bn, self._E[0], self._E[1], self._E[2])

    def B(self, tpos):
        z = tpos[3]
        return np.array([self.B0 * np.tanh(z / self.lam), 0, self.Bn])

    def E(self, tpos):
        return self._E.copy()


class DipolePlusUniform(EarthDipole):
    """EarthDipole plus a uniform field Bu (static)."""
    _model = FIELD_DIPOLE_UNIFORM

    def __init__(self, Bu=(0.0, 0.0, -10e-9), B0=B0):
        EarthDipole.__init__(self, B0)
        self._Bu = np.array(Bu, dtype=float)

    def _params(self):
        return (self._coeff, self._Bu[0], self._Bu[1], self._Bu[2])

    def B(self, tpos):
        return EarthDipole.B(self, tpos) + self._Bu


_grid_ids = itertools.count(1)


class Grid(_Field):
    """Fields sampled on a Cartesian grid (rapt/fields.py:513-814).

    Subclass it and override ``parsefile(filename)`` to return a dict with
    "time" (float), "x", "y", "z" (1-D, strictly increasing; uniform spacing
    not required) and "Bx", "By", "Bz", "Ex", "Ey", "Ez" (3-D arrays indexed
    [x, y, z]), all SI, as the reference's contract states.

    Differences from the reference, by design:

    - every file is parsed at construction and all snapshots stay resident
      on the GPU (288 GB of HBM holds long MHD runs); the reference keeps a
      sliding window of three snapshots and forgets earlier times.  Within
      any window the interpolation is identical: linear in t between the
      bracketing snapshots, trilinear in space (scipy's
      RegularGridInterpolator, method "linear");
    - a single file gives a time-independent 3-D grid (the reference's
      constructor sets ``time_indep`` instead of ``_time_indep`` and then
      calls its 3-D interpolator with four arguments);
    - outside the grid the host methods raise ValueError as the reference
      does; on the device B and E are NaN there and a traced particle stops
      with a non-finite or step-size status.

    ``B``/``E`` may not be overridden for tracing (the device interpolates the
    parsed data only).  ``static`` stays True unless the subclass sets it, as
    in the reference.
    """
    _model = FIELD_GRID

    def __init__(self, filelist):
        assert len(filelist) > 0
        _Field.__init__(self)
        self.gradientstepsize = 1e-3 * Re
        self.files = list(filelist)
        gs = [self.parsefile(f) for f in self.files]
        g0 = gs[0]
        self._x, self._y, self._z = (np.asarray(g0[k], dtype=float) for k in ("x", "y", "z"))
        shape = (len(self._x), len(self._y), len(self._z))
        for g in gs:
            for k in ("Bx", "By", "Bz", "Ex", "Ey", "Ez"):
                if np.shape(g[k]) != shape:
                    raise ValueError("%s has shape %s, grid is %s" % (k, np.shape(g[k]), shape))
        self._t = np.array([float(g["time"]) for g in gs])
        self._B = np.ascontiguousarray(
            np.stack([np.stack([g["Bx"], g["By"], g["Bz"]], axis=-1) for g in gs]), dtype=float)
        self._E = np.ascontiguousarray(
            np.stack([np.stack([g["Ex"], g["Ey"], g["Ez"]], axis=-1) for g in gs]), dtype=float)
        self._time_indep = len(gs) == 1
        self._grid_id = next(_grid_ids)
        self._interp = None

    def parsefile(self, filename):
        """Parse one snapshot (override when subclassing)."""
        return dict()

    def grid_arrays(self):
        """(x, y, z, t, B [nt][nx][ny][nz][3], E likewise) as uploaded to the device."""
        return self._x, self._y, self._z, self._t, self._B, self._E

    def _interpolators(self):
        if self._interp is None:
            from scipy.interpolate import RegularGridInterpolator as RGI
            ax = (self._x, self._y, self._z)
            if self._time_indep:
                self._interp = (RGI(ax, self._B[0]), RGI(ax, self._E[0]))
            else:
                ax = (self._t,) + ax
                self._interp = (RGI(ax, self._B), RGI(ax, self._E))
        return self._interp

    def Bgrid(self, tpos):
        tpos = np.asarray(tpos, dtype=float)
        q = tpos[1:] if self._time_indep else tpos
        return self._interpolators()[0](q)[0]

    def Egrid(self, tpos):
        tpos = np.asarray(tpos, dtype=float)
        q = tpos[1:] if self._time_indep else tpos
        return self._interpolators()[1](q)[0]

    def B(self, tpos):
        return self.Bgrid(tpos)

    def E(self, tpos):
        return self.Egrid(tpos)

    def descriptor(self):
        d = _Field.descriptor(self)
        d["grid"] = self
        return d


_MODEL_B = {FIELD_EARTH_DIPOLE: EarthDipole.B, FIELD_DOUBLE_DIPOLE: DoubleDipole.B,
            FIELD_UNIFORM_BZ: UniformBz.B, FIELD_UNIFORM_CROSSED_EB: UniformBz.B,
            FIELD_VAR_EARTH_DIPOLE: VarEarthDipole.B, FIELD_PARABOLIC: Parabolic.B,
            FIELD_CURRENT_SHEET: CurrentSheet.B, FIELD_DIPOLE_UNIFORM: DipolePlusUniform.B,
            FIELD_GRID: Grid.B}
_MODEL_E = {FIELD_EARTH_DIPOLE: _Field.E, FIELD_DOUBLE_DIPOLE: _Field.E,
            FIELD_UNIFORM_BZ: _Field.E, FIELD_UNIFORM_CROSSED_EB: UniformCrossedEB.E,
            FIELD_VAR_EARTH_DIPOLE: _Field.E, FIELD_PARABOLIC: _Field.E,
            FIELD_CURRENT_SHEET: CurrentSheet.E, FIELD_DIPOLE_UNIFORM: _Field.E,
            FIELD_GRID: Grid.E}
