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


EQ_FIELDS = {"dip_p": F.EarthDipole, "dd_e": F.DoubleDipole}


@pytest.mark.parametrize("name", sorted(EQ_FIELDS))
def test_equatorial_branch_vs_reference(golden, name):
    """isequatorial (rapt/BounceCenter.py:237-238): gradient plus inertial
    drift through the C-ABI at the reference's output step, against the
    reference's rows (its loop given the tpos its deriv builds).  Positions
    within 1e-8 of |r|: the inertial term differences the gradient drift
    over +-d and carries its rounding noise (1e-3 of a term 1e-4 of v_g)."""
    from rapt_amd import _native as N
    g = golden("equatorial")
    _, _, _, v, pa, m, q, mu = g[name + "/in"]
    dt, delta, Bm, gamma = g[name + "/dt"]
    ref = g[name + "/traj"]
    rtol, atol = rapt.params["solvertolerances"]
    rows, end, st, acc = N.context().bounce_center_equatorial(
        EQ_FIELDS[name]().descriptor(), ref[0], [mu, m, q, gamma], dt, 0.0, len(ref) - 1,
        rtol, atol)
    assert st[0] == 1
    got = rows[0]
    np.testing.assert_array_equal(got[:, 0], ref[1:, 0])
    err = np.max(np.linalg.norm(got[:, 1:] - ref[1:, 1:], axis=1) /
                 np.linalg.norm(ref[1:, 1:], axis=1))
    print(name, "max position error / |r|", err)
    assert err < 1e-8


def test_equatorial_object_api(golden):
    """BounceCenter with isequatorial set, advance() through the object API:
    the output step comes from the device's bounce period (within 1e-4 of the
    reference's), the rows follow the reference's."""
    g = golden("equatorial")
    pos = g["dip_p/in"][:3]
    _, _, _, v, pa, m, q, mu = g["dip_p/in"]
    dt, delta, Bm, gamma = g["dip_p/dt"]
    ref = g["dip_p/traj"]
    bc = rapt.BounceCenter(list(pos), v, 0, pa, m, q, F.EarthDipole())
    assert bc.mu == pytest.approx(mu, rel=1e-12)
    tp = g["dip_p/tpos"][0]
    vg, vi = g["dip_p/v_grad"][0], g["dip_p/v_inert"][0]
    np.testing.assert_allclose(bc._v_grad(tp), vg, rtol=0, atol=1e-8 * np.linalg.norm(vg))
    np.testing.assert_allclose(bc._v_inert(tp), vi, rtol=0, atol=5e-3 * np.linalg.norm(vi))
    bc.isequatorial = True
    bc.advance(delta - 0.5 * dt)
    tr = bc.trajectory
    assert tr.shape == ref.shape
    np.testing.assert_allclose(tr[:, 0], ref[:, 0], rtol=1e-4, atol=0)
    err = np.max(np.linalg.norm(tr[:, 1:] - ref[:, 1:], axis=1) /
                 np.linalg.norm(ref[:, 1:], axis=1))
    assert err < 1e-6
