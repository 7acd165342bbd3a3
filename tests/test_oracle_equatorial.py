"""Pins of the oracle's equatorial BounceCenter drift (oracle/rapt_oracle.c,
orc_bc_equatorial), the branch of deriv taken when isequatorial is set
(rapt/BounceCenter.py:154-204, 237-238):

- the reference's own _v_grad and _v_inert at 4-vector positions and its
  advance() loop with the tpos deriv builds (tests/golden/equatorial.npz):
  identical, value by value and row by row;
- the dipole closed forms at the equator: |v_grad| = 3 mu / (q gamma r),
  westward for a positive charge, and v_inert = -m v_g^2 / (q B r) phi-hat,
  the centripetal correction of a circular drift.
"""
import numpy as np
import pytest

import rapt_amd.fields as F
from rapt_amd import constants as K

CASES = {"dip_p": F.EarthDipole, "dd_e": F.DoubleDipole}


@pytest.mark.parametrize("name", sorted(CASES))
def test_drifts_match_reference(oracle, golden, name):
    g = golden("equatorial")
    _, _, _, v, pa, m, q, mu = g[name + "/in"]
    gamma = g[name + "/dt"][3]
    fo = oracle.field_from_desc(CASES[name]().descriptor())
    for tp, vg, vi in zip(g[name + "/tpos"], g[name + "/v_grad"], g[name + "/v_inert"]):
        a, b = oracle.bc_equatorial_drifts(fo, tp, mu, m, q, gamma)
        np.testing.assert_array_equal(a, vg)
        np.testing.assert_array_equal(b, vi)


@pytest.mark.parametrize("name", sorted(CASES))
def test_advance_matches_reference(oracle, golden, name):
    g = golden("equatorial")
    _, _, _, v, pa, m, q, mu = g[name + "/in"]
    dt, delta, Bm, gamma = g[name + "/dt"]
    ref = g[name + "/traj"]
    fo = oracle.field_from_desc(CASES[name]().descriptor())
    rows, end, info = oracle.bc_equatorial(fo, oracle.gc_params(), ref[0], mu, m, q, gamma, dt,
                                           0.0, len(ref) - 1)
    assert info["status"] == 1
    np.testing.assert_array_equal(rows, ref[1:])
    np.testing.assert_array_equal(end, g[name + "/end"])


@pytest.mark.parametrize("r_re,charge", [(3.0, K.e), (5.5, K.e), (4.0, -K.e)])
def test_dipole_closed_forms(oracle, r_re, charge):
    f = F.EarthDipole()
    fo = oracle.field_from_desc(f.descriptor())
    phi = 0.7
    r = r_re * K.Re
    tp = np.array([0.0, r * np.cos(phi), r * np.sin(phi), 0.0])
    m, mu, gamma = K.m_pr, 3e-7, 1.02
    vg, vi = oracle.bc_equatorial_drifts(fo, tp, mu, m, charge, gamma)
    B = K.B0 / r_re ** 3
    phat = np.array([-np.sin(phi), np.cos(phi), 0.0])
    V = 3 * mu / (abs(charge) * gamma * r)
    sgn = -np.sign(charge)  # westward for positive charges
    np.testing.assert_allclose(vg, sgn * V * phat, rtol=0, atol=1e-8 * V)
    Vi = m * V ** 2 / (abs(charge) * B * r)
    # the difference of v_g over +-d carries the rounding noise of grad|B|
    # (~1e-10) divided by 2d/r (~4e-7): up to 1.5e-3 of the inertial term
    # (measured), 5e-3 allowed; a wrong sign or factor 2 is far outside
    np.testing.assert_allclose(vi, sgn * Vi * phat, rtol=0, atol=5e-3 * Vi)
