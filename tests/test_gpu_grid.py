"""Gridded fields (Grid, rapt/fields.py:513-814) on the device.

- Interpolation against scipy's RegularGridInterpolator (the reference's own
  interpolator, method "linear") on a dipole sampled on non-uniform and
  uniform grids with three snapshots: B and E at random (t, x, y, z) agree to
  1e-12 relative; grid nodes and the upper corner return the node values;
  points outside the grid are NaN on the device and ValueError on the host.
- Multilinear fields are reproduced exactly by the interpolation, so a
  gridded uniform crossed E x B field gives the analytic UniformCrossedEB
  model's Particle trajectories bit for bit, and a gridded uniform Bz the
  analytic model's guiding-centre trajectories.
"""
import numpy as np
import pytest

import rapt_amd as rapt
from rapt_amd import _native as N
from rapt_amd import constants as K
from rapt_amd import fields as F
from rapt_amd import batch

pytestmark = pytest.mark.gpu


class _SampledDipole(F.Grid):
    """Earth dipole times (1 + 0.1 t) sampled on the given axes; E = 1e-3 x
    the dipole rotated, to exercise E interpolation."""

    def __init__(self, axes, times):
        self._axes = axes
        F.Grid.__init__(self, [str(t) for t in times])

    def parsefile(self, filename):
        t = float(filename)
        x, y, z = self._axes
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        d = F.EarthDipole()
        r2 = X * X + Y * Y + Z * Z
        s = d._coeff / r2 ** 2.5 * (1 + 0.1 * t)
        Bx, By, Bz = s * X * Z, s * Y * Z, s * (Z * Z - r2 / 3)
        return dict(time=t, x=x, y=y, z=z, Bx=Bx, By=By, Bz=Bz, Ex=1e-3 * By, Ey=1e-3 * Bz,
                    Ez=1e-3 * Bx)


def _axes(uniform):
    if uniform:
        a = np.linspace(2.0, 6.0, 21) * K.Re
        return a, a.copy(), np.linspace(-2.0, 2.0, 17) * K.Re
    a = 2.0 * K.Re * (3.0 ** (np.arange(21) / 20.0))        # geometric 2..6 Re
    return a, np.sort(np.r_[a[::2], 4.1 * K.Re]), np.sinh(np.linspace(-1.4, 1.4, 17)) * K.Re


@pytest.mark.parametrize("uniform", [True, False])
def test_grid_interpolation_matches_scipy(uniform):
    g = _SampledDipole(_axes(uniform), [0.0, 1.0, 3.0])
    x, y, z = g._x, g._y, g._z
    rng = np.random.default_rng(11)
    n = 400
    tp = np.c_[rng.uniform(0, 3, n), rng.uniform(x[0], x[-1], n), rng.uniform(y[0], y[-1], n),
               rng.uniform(z[0], z[-1], n)]
    # grid nodes, the last corner, and snapshot times
    tp[:4] = [[1.0, x[3], y[5], z[7]], [3.0, x[-1], y[-1], z[-1]], [0.0, x[0], y[0], z[0]],
              [2.0, x[10], y[10], z[8]]]
    ctx = N.context()
    Bd = ctx.field_eval(g.descriptor(), "B", tp)
    Ed = ctx.field_eval(g.descriptor(), "E", tp)
    Bh = np.array([g.B(p) for p in tp])
    Eh = np.array([g.E(p) for p in tp])
    scale = np.linalg.norm(Bh, axis=1)[:, None]
    errB = np.max(np.abs(Bd - Bh) / scale)
    errE = np.max(np.abs(Ed - Eh) / (1e-3 * scale))
    assert errB < 1e-12 and errE < 1e-12, (errB, errE)
    np.testing.assert_allclose(Bd[0], g._B[1, 3, 5, 7], rtol=1e-15)
    np.testing.assert_allclose(Bd[1], g._B[2, -1, -1, -1], rtol=1e-15)


def test_grid_outside_is_nan_on_device_and_valueerror_on_host():
    g = _SampledDipole(_axes(True), [0.0, 1.0])
    x = g._x
    tp = np.array([[0.5, x[0] - 1.0, 3 * K.Re, 0.0], [1.5, 4 * K.Re, 3 * K.Re, 0.0],
                   [0.5, 4 * K.Re, 3 * K.Re, 0.1 * K.Re]])
    B = N.context().field_eval(g.descriptor(), "B", tp)
    assert np.all(np.isnan(B[:2])) and np.all(np.isfinite(B[2]))
    with pytest.raises(ValueError):
        g.B(tp[0])


class _Uniform(F.Grid):
    def __init__(self, Bz, Ey, times=(0.0,)):
        self._f = (Bz, Ey)
        F.Grid.__init__(self, [str(t) for t in times])

    def parsefile(self, filename):
        x = np.linspace(-40.0, 40.0, 9) * K.Re
        z = np.linspace(-10.0, 10.0, 5) * K.Re
        shape = (len(x), len(x), len(z))
        one = np.ones(shape)
        return dict(time=float(filename), x=x, y=x.copy(), z=z, Bx=0 * one, By=0 * one,
                    Bz=self._f[0] * one, Ex=0 * one, Ey=self._f[1] * one, Ez=0 * one)


def test_gridded_uniform_crossed_field_gives_analytic_trajectories():
    Bz, Ey = 1e-7, 1e-3
    ana = F.UniformCrossedEB(Ey=Ey, Bz=Bz)
    grd = _Uniform(Bz, Ey, times=(0.0, 50.0, 100.0))
    grd.static = ana.static
    rng = np.random.default_rng(3)
    n = 64
    v = np.array([rapt.utils.speedfromKE(k, K.m_pr) for k in rng.uniform(1e3, 1e4, n)])
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    g = 1 / np.sqrt(1 - (v / K.c) ** 2)
    last = np.zeros((n, 7))
    last[:, 1:3] = rng.uniform(-5, 5, (n, 2)) * K.Re
    last[:, 4:7] = (K.m_pr * g * v)[:, None] * u
    a = batch.advance_particles(ana, last, K.m_pr, K.e, 20.0, params=rapt.params)
    b = batch.advance_particles(grd, last, K.m_pr, K.e, 20.0, params=rapt.params)
    np.testing.assert_array_equal(a.status, b.status)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("lpp", [1, 4])
def test_gridded_uniform_bz_guiding_centres(lpp):
    Bz = -2e-7
    ana = F.UniformBz(Bz)
    grd = _Uniform(Bz, 0.0)
    grd.gradientstepsize = ana.gradientstepsize
    n = 32
    rng = np.random.default_rng(5)
    last = np.zeros((n, 5))
    last[:, 1:3] = rng.uniform(-5, 5, (n, 2)) * K.Re
    last[:, 4] = rng.uniform(-1, 1, n) * 1e-24
    mu = rng.uniform(1e-16, 1e-15, n)
    v = np.sqrt(2 * mu * abs(Bz) / K.m_el + (last[:, 4] / K.m_el) ** 2)
    prm = rapt.params.copy()
    a = batch.advance_gcs(ana, last, K.m_el, -K.e, mu, v, 0.5, 5.0, params=prm,
                          lanes_per_particle=lpp)
    b = batch.advance_gcs(grd, last, K.m_el, -K.e, mu, v, 0.5, 5.0, params=prm,
                          lanes_per_particle=lpp)
    np.testing.assert_array_equal(a.status, b.status)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(x, y)


def test_cell_and_point_layouts_agree(monkeypatch):
    """The cell-major device copy (the default) and the point-major one
    (RAPT_GRID_POINTS=1) give bit-identical field values and trajectories."""
    g1 = _SampledDipole(_axes(False), [0.0, 1.0, 3.0])
    g2 = _SampledDipole(_axes(False), [0.0, 1.0, 3.0])
    rng = np.random.default_rng(12)
    x, y, z = g1._x, g1._y, g1._z
    n = 300
    tp = np.c_[rng.uniform(0, 3, n), rng.uniform(x[0], x[-1], n), rng.uniform(y[0], y[-1], n),
               rng.uniform(z[0], z[-1], n)]
    ctx = N.context()
    cells = [ctx.field_eval(g1.descriptor(), w, tp) for w in ("B", "E", "curlb")]
    monkeypatch.setenv("RAPT_GRID_POINTS", "1")
    points = [ctx.field_eval(g2.descriptor(), w, tp) for w in ("B", "E", "curlb")]
    for a, b in zip(cells, points):
        np.testing.assert_array_equal(a, b)
