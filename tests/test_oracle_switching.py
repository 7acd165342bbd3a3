"""Pins of the oracle's switching path: the adiabaticity tests, the
Particle <-> GuidingCenter conversions and the Adaptive loop
(oracle/rapt_oracle.c orc_particle_isadiabatic, orc_gc_isadiabatic,
orc_adaptive_to_gc, orc_adaptive_to_particle, orc_adaptive_run,
orc_adaptive_continue).

- decisions of the reference's own tracer objects (tests/golden/make_golden.py
  make_utils, make_switching): Particle.isadiabatic in the current sheet and
  in a time-dependent dipole (rapt/Particle.py:345-384), GuidingCenter
  .isadiabatic in a static and a time-dependent dipole, with the cycper of
  rapt/GuidingCenter.py:531-541 (rapt/GuidingCenter.py:287-327), under
  four (epss, epst) settings;
- the switch rows of Adaptive runs in the tanh current sheet straddle epss
  under the closed-form rho/L of that field;
- GuidingCenter -> Particle -> GuidingCenter returns the guiding centre to
  the fixed-point tolerance and keeps the speed (rapt/utils.py:251-433);
- an Adaptive run in UniformCrossedEB is a guiding centre from the start and
  drifts at Ey/Bz while streaming along B at its parallel speed;
- a second Adaptive.advance continues the last tracer as Adaptive.advance
  does with t reset to 0 (rapt/Adaptive.py:187-222).
"""
import numpy as np
import pytest

import rapt_amd.fields as F
from rapt_amd import constants as K
from rapt_amd import workloads as W

from _sheet import C, segments as _segments, sheet_ratio as _sheet_ratio


def test_particle_isadiabatic_matches_reference_sheet(oracle, golden):
    U = golden("utils")
    dec = U["sheet_particle_adiabatic"]
    f = oracle.field_from_desc(F.CurrentSheet().descriptor())
    prm = oracle.particle_params()
    for i, (pos, vel) in enumerate(zip(U["sheet_pos"], U["sheet_vel"])):
        g = 1 / np.sqrt(1 - vel @ vel / C ** 2)
        row = np.r_[0.0, pos, K.m_pr * g * vel]
        assert oracle.particle_isadiabatic(f, prm, row, K.m_pr, K.e) == bool(dec[i, 0]), i
        cr = oracle.cyclotron_radius(f, 0.0, pos, vel, K.m_pr, K.e)
        assert cr == pytest.approx(dec[i, 1], rel=1e-13)
        ls = oracle.field_eval(f, "lengthscale", row[:4])[0, 0]
        assert ls == pytest.approx(dec[i, 2], rel=2e-8)  # numpy tanh vs libm, times lam/d


@pytest.mark.parametrize("case", ["gc_dd", "gc_var"])
def test_gc_isadiabatic_matches_reference(oracle, golden, case):
    S = golden("switching")
    field = F.DoubleDipole() if case == "gc_dd" else F.VarEarthDipole(amp=0.2, period=7.0)
    f = oracle.field_from_desc(field.descriptor())
    rows, mqmu, dec = S[case + "/row"], S[case + "/mqmu"], S[case + "/dec"]
    flips = 0
    for k, (es, et) in enumerate(S["settings"]):
        prm = oracle.gc_params(epss=es, epst=et)
        got = np.array([oracle.gc_isadiabatic(f, prm, rows[i], *mqmu[i]) for i in range(len(rows))])
        np.testing.assert_array_equal(got, dec[:, k].astype(bool), err_msg="setting %d" % k)
        flips += np.sum(dec[:, k] != dec[:, 0])
    ls = oracle.field_eval(f, "lengthscale", rows[:, :4])[:, 0]
    np.testing.assert_array_equal(ls, dec[:, 6])
    if case == "gc_var":
        np.testing.assert_array_equal(oracle.field_eval(f, "timescale", rows[:, :4])[:, 0],
                                      dec[:, 7])
        # the epst settings change decisions: the cycper branch is exercised
        assert np.any(dec[:, 2] != dec[:, 0]) and np.any(dec[:, 3] != dec[:, 0])
    assert flips > 0


def test_gc_isadiabatic_cycper_relativistic_matches_reference(oracle, golden):
    """Relativistic electrons at small pitch angles against epst thresholds
    across their cycper / timescale: the decisions follow the reference's
    cycper, whose pp**2 is not divided by (m c)^2 (rapt/GuidingCenter.py:535)."""
    S = golden("switching")
    f = oracle.field_from_desc(F.VarEarthDipole(amp=0.2, period=7.0).descriptor())
    rows, mqmu, dec = S["gc_var_e/row"], S["gc_var_e/mqmu"], S["gc_var_e/dec"]
    for k, (es, et) in enumerate(S["settings_e"]):
        prm = oracle.gc_params(epss=es, epst=et)
        got = np.array([oracle.gc_isadiabatic(f, prm, rows[i], *mqmu[i]) for i in range(len(rows))])
        np.testing.assert_array_equal(got, dec[:, k].astype(bool), err_msg="setting %d" % k)
    assert 0 < dec[:, :5].sum() < dec[:, :5].size


def test_particle_isadiabatic_time_dependent_matches_reference(oracle, golden):
    S = golden("switching")
    f = oracle.field_from_desc(F.VarEarthDipole(amp=0.2, period=7.0).descriptor())
    rows, mq, dec = S["p_var/row"], S["p_var/mq"], S["p_var/dec"]
    for k, (es, et) in enumerate(S["settings"]):
        prm = oracle.particle_params(epss=es, epst=et)
        got = np.array([oracle.particle_isadiabatic(f, prm, rows[i], *mq[i])
                        for i in range(len(rows))])
        np.testing.assert_array_equal(got, dec[:, k].astype(bool), err_msg="setting %d" % k)
    assert np.any(dec[:, 3] != dec[:, 0])  # the epst branch decides some
    for i in range(len(rows)):
        p = rows[i, 4:7]
        m, q = mq[i]
        v = p / np.sqrt(m ** 2 + p @ p / C ** 2)
        assert oracle.cyclotron_radius(f, rows[i, 0], rows[i, 1:4], v, m, q) == dec[i, 4]
        assert oracle.cyclotron_period(f, rows[i, 0], rows[i, 1:4], v, m, q) == dec[i, 5]


def _sheet_runs(oracle, n, delta):
    wl = W.current_sheet_protons(n=n, seed=7)
    fo = oracle.field_from_desc(F.CurrentSheet().descriptor())
    prm = oracle.particle_params(gctimestep=0.5)
    runs = []
    for i in range(n):
        rows, info = oracle.adaptive_run(fo, prm, wl["pos"][i], wl["vel"][i], 0.0, K.m_pr, K.e,
                                         delta)
        assert info["status"] == 1, (i, info)
        g = 1 / np.sqrt(1 - wl["vel"][i] @ wl["vel"][i] / C ** 2)
        runs.append((rows, np.r_[0.0, wl["pos"][i], K.m_pr * g * wl["vel"][i]]))
    return runs


def test_adaptive_switch_rows_straddle_epss(oracle):
    """Every row at which Adaptive tests adiabaticity (each row after a
    segment's start row) lies on the side of epss its segment requires, and
    the last row of a segment that ends in a switch on the other side; the
    start (Adaptive.__init__ tests the particle) decides the first mode --
    with rho/L in closed form, not the oracle's finite differences."""
    epss, tiny = 5e-2, 1e-6
    nsw = 0
    for rows, p0 in _sheet_runs(oracle, 96, 60.0):
        r0 = _sheet_ratio(p0, 0, 0.0, K.m_pr, K.e)
        if rows[0, 0] == 1:
            assert r0 < epss * (1 + tiny)
        else:
            assert r0 >= epss * (1 - tiny)
            np.testing.assert_array_equal(rows[0, 1:8], p0)
        segs = _segments(rows)
        for j, seg in enumerate(segs):
            mode = int(seg[0, 0])
            mu = seg[0, 6] if mode == 1 else 0.0
            r = np.array([_sheet_ratio(row[1:], mode, mu, K.m_pr, K.e) for row in seg])
            tested = r[1:]
            switched = j < len(segs) - 1
            body = tested[:-1] if switched else tested
            if mode == 0:
                assert np.all(body >= epss * (1 - tiny)), (j, body.min())
                if switched:
                    assert tested[-1] < epss * (1 + tiny)
            else:
                assert np.all(body < epss * (1 + tiny)), (j, body.max())
                if switched:
                    assert tested[-1] >= epss * (1 - tiny)
            nsw += switched
        # t becomes the new tracer's absolute tcur after a switch, so the run
        # ends one output step past t0 + delta (a Particle step is at most
        # T_c / 10 = 6.6 s in the sheet's weakest field, Bn)
        assert 60.0 <= rows[-1, 1] < 60.0 + 6.6, rows[-1, 1]
    assert nsw >= 50  # both directions are exercised many times


def test_gc_particle_round_trip(oracle):
    """GCtoFP at gyrophase 0 then the guiding-centre iteration returns the
    guiding centre within the iteration's 1e-3 tolerance, keeps the speed,
    and keeps the parallel momentum and moment to first order in rho/L."""
    fo = oracle.field_from_desc(F.EarthDipole().descriptor())
    wl = W.mixed_dipole_uniform(n=64, seed=9)
    worst = 0.0
    for i in range(64):
        m, q = wl["mass"][i], wl["charge"][i]
        g = 1 / np.sqrt(1 - wl["vel"][i] @ wl["vel"][i] / C ** 2)
        prow = np.r_[0.0, wl["pos"][i] * [1.0, 1.0, 1.0] + [0, 0, 0.4 * K.Re],
                     m * g * wl["vel"][i]]
        out = oracle.adaptive_to_gc(fo, prow, m, q)
        assert out is not None
        g5, mu, v = out
        p7 = oracle.adaptive_to_particle(fo, g5, m, q, mu)
        p = p7[4:7]
        gm = np.sqrt(m ** 2 + p @ p / C ** 2)
        assert np.sqrt(p @ p) / gm == pytest.approx(v, rel=1e-12)
        back = oracle.adaptive_to_gc(fo, p7, m, q)
        assert back is not None
        g5b, mub, vb = back
        d = np.linalg.norm(g5b[1:4] - g5[1:4]) / np.linalg.norm(g5[1:4])
        worst = max(worst, d)
        assert d < 1e-3, (i, d)
        assert vb == pytest.approx(v, rel=1e-12)
        assert mub == pytest.approx(mu, rel=2e-2)
        assert g5b[4] == pytest.approx(g5[4], rel=2e-2, abs=1e-3 * m * g * v)
    # and the gyro-offset is not applied twice or with the wrong sign: the
    # particle sits one cyclotron radius from its guiding centre
    assert worst > 0.0


def test_adaptive_crossed_eb_drifts_at_ey_over_bz(oracle):
    Ey, Bz = 1e-3, 1e-7
    f = F.UniformCrossedEB(Ey=Ey, Bz=Bz)
    fo = oracle.field_from_desc(f.descriptor())
    prm = oracle.particle_params(gctimestep=0.01)
    v0 = np.array([3e4, -1e4, 2e4])
    rows, info = oracle.adaptive_run(fo, prm, [1e5, 2e5, 0.0], v0, 0.0, K.m_pr, K.e, 0.5)
    assert info["status"] == 1 and info["nswitch"] == 0
    assert np.all(rows[:, 0] == 1)  # a guiding centre from the first row
    t, X, Y, Z, pp = rows[:, 1], rows[:, 2], rows[:, 3], rows[:, 4], rows[:, 5]
    mu = rows[0, 6]
    g = np.sqrt(1 + 2 * mu * Bz / (K.m_pr * C * C) + (pp / (K.m_pr * C)) ** 2)
    vpar = pp / (g * K.m_pr)
    np.testing.assert_allclose(X - X[0], Ey / Bz * (t - t[0]), rtol=1e-9, atol=1e-9 * abs(X[0]))
    np.testing.assert_allclose(Y, Y[0], rtol=0, atol=1e-9 * abs(Y[0]))
    np.testing.assert_allclose(Z - Z[0], vpar[0] * (t - t[0]), rtol=1e-9, atol=1e-6)
    assert t[-1] >= 0.5


def test_adaptive_continue_follows_last_tracer(oracle):
    """A second advance continues the last tracer: its rows before the first
    switch are those of the tracer's own advance from its last row (the
    Particle/GuidingCenter drivers, pinned bit-exactly elsewhere), with no
    start row; after a switch the new tracer's start row follows."""
    fo = oracle.field_from_desc(F.CurrentSheet().descriptor())
    # an output step long enough that the guiding-centre steps are limited by
    # the error estimate, not by the output step
    gdt = 4.0
    prm = oracle.particle_params(gctimestep=gdt)
    wl = W.current_sheet_protons(n=24, seed=7)
    checked = 0
    modes = set()
    for i in range(24):
        rows, info = oracle.adaptive_run(fo, prm, wl["pos"][i], wl["vel"][i], 0.0, K.m_pr, K.e,
                                         30.0)
        last = rows[-1]
        assert last[1] >= 30.0
        mode = int(last[0])
        mu = 0.0
        if mode == 1:
            seg = _segments(rows)[-1]
            mu = seg[0, 6]
            pp, B = last[5], oracle.field_eval(fo, "magB", last[1:5])[0, 0]
            g = np.sqrt(1 + 2 * mu * B / (K.m_pr * C * C) + (pp / (K.m_pr * C)) ** 2)
            v = C * np.sqrt(1 - 1 / g ** 2) if g - 1 >= 1e-6 else \
                np.sqrt(2 * mu * B / K.m_pr + (pp / K.m_pr) ** 2)
            state = last[1:6]
        else:
            v = 0.0
            state = last[1:8]
        more, info2 = oracle.adaptive_continue(fo, prm, mode, state, mu, v, K.m_pr, K.e, 20.0)
        assert info2["status"] == 1
        if mode == 0:
            own, inf = oracle.particle_advance(fo, oracle.particle_params(
                gctimestep=gdt, check_adiabaticity=True), state, K.m_pr, K.e, 20.0)
            k = len(own)
            np.testing.assert_array_equal(more[:k, 1:8], own)
        else:
            own, inf = oracle.gc_advance(fo, oracle.gc_params(gctimestep=gdt,
                                                              check_adiabaticity=True),
                                         0, state, K.m_pr, K.e, mu, v, gdt, 20.0)
            k = len(own)
            np.testing.assert_array_equal(more[:k, 1:6], own)
        assert np.all(more[:k, 0] == mode)
        if inf["switched"]:
            # the new tracer's start row; its tcur (past 30 s) is not below
            # this advance's delta of 20 s, so the loop ends there
            assert more[k, 0] != mode and len(more) == k + 1
        else:
            assert len(more) == k
        modes.add((mode, bool(inf["switched"])))
        checked += 1
    assert checked == 24 and len(modes) >= 3
