"""User-written fields, defined the way the reference asks for new fields:
subclass ``_Field``, override ``B`` and/or ``E`` in Python, set ``static`` and
``gradientstepsize`` (rapt/fields.py:5-74).

``classes(fields_module)`` builds the same class bodies on either field
module, so the fixture generator runs them under the reference and the tests
run the identical code under rapt_amd (whose ``descriptor()`` traces them
onto the device).
"""
import numpy as np

RE = 6378137  # rapt/__init__.py:10 (an int there)


def classes(fm):
    class SheetField(fm._Field):
        """B = (B0 tanh(z/lam), 0, Bn), E uniform."""

        def __init__(self, B0=20e-9, lam=0.5 * RE, Bn=1e-9, E=(0.0, 0.3e-3, 0.0)):
            fm._Field.__init__(self)
            self.B0, self.lam, self.Bn = B0, lam, Bn
            self._E = np.array(E, dtype=float)
            self.static = not np.any(self._E != 0)
            self.gradientstepsize = lam / 1000

        def B(self, tpos):
            z = tpos[3]
            return np.array([self.B0 * np.tanh(z / self.lam), 0, self.Bn])

        def E(self, tpos):
            return self._E.copy()

    class DipolePlusUniformField(fm.EarthDipole):
        """The Earth's dipole plus a uniform field, calling the parent's B."""

        def __init__(self, Bu=(0.0, 0.0, -10e-9)):
            fm.EarthDipole.__init__(self)
            self._Bu = np.array(Bu, dtype=float)

        def B(self, tpos):
            return fm.EarthDipole.B(self, tpos) + self._Bu

    class ChargedDipole(fm._Field):
        """A magnetic dipole at the origin (B0 at unit distance in the x-y
        plane) carrying a point charge Q: time-dependent in the reference's
        sense (static = False) because E is non-zero."""

        def __init__(self, B0=1, Q=1e-6):
            fm._Field.__init__(self)
            self.B0, self.Q = B0, Q
            self._kc = 8.9875517873681764e9  # Coulomb constant
            self.static = False

        def B(self, tpos):
            t, x, y, z = tpos
            r2 = x * x + y * y + z * z
            return self.B0 * np.array([3 * x * z, 3 * y * z, 2 * z * z - x * x - y * y]) / \
                pow(r2, 2.5)

        def E(self, tpos):
            t, x, y, z = tpos
            return self._kc * self.Q * np.array([x, y, z]) / pow(x * x + y * y + z * z, 1.5)

    class TwinDipole(fm._Field):
        """Two aligned dipoles, the image one at x = d with relative strength
        k; the position variable is shifted in place between the two terms."""

        def __init__(self, B0=3.07e-5, d=20 * RE, k=1):
            fm._Field.__init__(self)
            self.gradientstepsize = RE / 1000
            self._d, self._k = d, k
            self._c = -B0 * RE ** 3

        def B(self, tpos):
            t, x, y, z = tpos
            B1 = np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / \
                pow(x * x + y * y + z * z, 5.0 / 2.0)
            x -= self._d
            B2 = self._k * np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / \
                pow(x * x + y * y + z * z, 5.0 / 2.0)
            return self._c * (B1 + B2)

    class BreathingSheet(fm._Field):
        """A sheet whose thickness oscillates in time, with a Gaussian normal
        component and an inductive E: numpy's elementwise functions of t and
        the coordinates."""

        def __init__(self, B0=20e-9, lam=0.5 * RE, Bn=2e-9, w=0.3, amp=0.2, E0=1e-4):
            fm._Field.__init__(self)
            self.B0, self.lam, self.Bn, self.w, self.amp, self.E0 = B0, lam, Bn, w, amp, E0
            self.static = False
            self.gradientstepsize = lam / 1000

        def B(self, tpos):
            t, x, y, z = tpos
            L = self.lam * (1 + self.amp * np.sin(self.w * t))
            return np.array([self.B0 * np.tanh(z / L), 0 * x,
                             self.Bn * np.exp(-(x / (10 * self.lam)) ** 2)])

        def E(self, tpos):
            t, x, y, z = tpos
            return np.array([0 * t, self.E0 * np.cos(self.w * t) / np.cosh(z / self.lam), 0 * t])

    return dict(SheetField=SheetField, DipolePlusUniformField=DipolePlusUniformField,
                ChargedDipole=ChargedDipole, TwinDipole=TwinDipole,
                BreathingSheet=BreathingSheet)
