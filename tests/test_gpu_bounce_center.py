"""BounceCenter.advance on the GPU against the dipole closed form.

In a pure dipole the bounce-averaged drift of an equatorial guiding line is
an azimuthal rotation at constant radius with the angular velocity omega_d of
the closed form in test_gpu_labels (S_b and I integrated along
r = L Re sin^2 theta, dI/dr by the same central difference).  After k output
steps the state is at phi0 + omega_d (k+1) dt on the same circle; the rows
carry the reference's times np.arange(tcur, tcur + delta, dt), one step behind
the state (rapt/BounceCenter.py:247-250).
Tolerances: radius 1e-6 relative, z 1e-6 Re, phase 1e-6 relative (measured:
radius to 1e-10, phase to 1.5e-8).
"""
import numpy as np
import pytest

import rapt_amd as rapt
from rapt_amd import constants as K
from rapt_amd import fields as F

from test_gpu_labels import _dipole

pytestmark = pytest.mark.gpu


def test_bounce_center_drifts_at_the_dipole_rate():
    L, phi0 = 4.0, 0.3
    Kev = 1e6
    v = K.c * np.sqrt(1 - 1 / (1 + Kev * K.e / (K.m_pr * K.c ** 2)) ** 2)
    pos = [L * K.Re * np.cos(phi0), L * K.Re * np.sin(phi0), 0.0]
    f = F.EarthDipole()
    saved = rapt.params["BCtimestep"]
    rapt.params["BCtimestep"] = 10.0   # ~0.55 rad of drift per output step
    try:
        bc = rapt.BounceCenter(pos, v, 0.0, 1.0, K.m_pr, K.e, f)
        g = 1 / np.sqrt(1 - (v / K.c) ** 2)
        Bm = K.m_pr * g ** 2 * v ** 2 / (2 * bc.mu)
        Beq = K.B0 / L ** 3
        alpha = np.degrees(np.arcsin(np.sqrt(Beq / Bm)))
        Tb, wd, _ = _dipole(L, alpha, np.log10(Kev), K.m_pr, K.e)
        dt = 10.0 * Tb
        import time
        t0 = time.perf_counter()
        bc.advance(9.5 * dt)   # arange(0, 9.5 dt, dt): 10 rows
        print("advance: %.2f s" % (time.perf_counter() - t0))
    finally:
        rapt.params["BCtimestep"] = saved
    tr = bc.trajectory
    assert tr.shape == (11, 4)
    np.testing.assert_allclose(tr[1:, 0], np.arange(10) * dt, rtol=1e-4, atol=0)
    r = np.hypot(tr[1:, 1], tr[1:, 2])
    assert np.max(np.abs(r / (L * K.Re) - 1)) < 1e-6
    assert np.max(np.abs(tr[1:, 3])) < 1e-6 * K.Re
    phi = np.unwrap(np.arctan2(tr[:, 2], tr[:, 1]))
    want = phi0 + wd * dt * np.arange(1, 11)
    rel = np.abs(phi[1:] - phi0 - (want - phi0)) / np.abs(want - phi0)
    print("phase rel err", rel.max(), "drift rad/s", wd)
    assert np.max(rel) < 1e-6
    assert np.all(np.diff(phi) < 0)   # protons drift westward


def test_bounce_center_rejects_time_dependent_field():
    with pytest.raises(RuntimeError):
        rapt.BounceCenter([4 * K.Re, 0, 0], 1e7, 0.0, 1.0, K.m_pr, K.e,
                          F.VarEarthDipole(amp=0.1, period=5.0))
