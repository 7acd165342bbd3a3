"""CPU tests of the object API around the GPU tracers (rapt/Particle.py,
rapt/GuidingCenter.py, rapt/Adaptive.py, rapt/BounceCenter.py):

- ``import rapt`` gives the same objects under the reference's module names;
- host accessors on the reference's own golden trajectories equal the
  reference's accessor outputs (tests/golden/make_golden.py make_accessors);
- pickle save/load round trips;
- Adaptive's host decision of the first mode, setke and setpa;
- batched advances launch tracers that check adiabaticity apart from those
  that do not (the launches themselves are replaced by a recorder here).
"""
import numpy as np
import pytest

import rapt_amd as rapt
import rapt_amd.fields as F
from rapt_amd import constants as K
from rapt_amd import workloads as W


def test_rapt_alias_package():
    import rapt as R
    import rapt.fields as rf
    from rapt.Adaptive import Adaptive
    from rapt.GuidingCenter import GuidingCenter
    from rapt.Particle import Particle
    from rapt.rkf import rkf
    from rapt import utils, flutils, fieldline  # noqa: F401
    assert R.params is rapt.params
    assert Particle is rapt.Particle and R.Particle is rapt.Particle
    assert GuidingCenter is rapt.GuidingCenter and Adaptive is rapt.Adaptive
    assert rf.DoubleDipole is F.DoubleDipole and R.fields is F
    assert R.Adiabatic is rapt.Adiabatic and R.NonAdiabatic is rapt.NonAdiabatic
    assert rkf is rapt.rkf.rkf
    assert (R.e, R.m_pr, R.m_el, R.c, R.B0, R.Re, R.re) == \
        (K.e, K.m_pr, K.m_el, K.c, K.B0, K.Re, K.Re)


def _particle_from(traj):
    p = rapt.Particle()
    p.mass, p.charge, p.field = K.m_pr, K.e, F.EarthDipole()
    p.trajectory = traj.copy()
    return p


def _gc_from(traj, mu):
    g = rapt.GuidingCenter()
    g.mass, g.charge, g.field = K.m_el, -K.e, F.DoubleDipole()
    g.mu = float(mu)
    g.trajectory = traj.copy()
    return g


def test_particle_host_accessors_match_reference(golden):
    A, P = golden("accessors"), golden("particles")
    off = np.r_[0, np.cumsum(P["c1/len"])]
    for k in range(3):
        p = _particle_from(P["c1/traj"][off[k]:off[k + 1]])
        for nm in ("gett", "getx", "gety", "getz", "getpx", "getpy", "getpz", "getp",
                   "getgamma", "getvx", "getvy", "getvz", "getr", "gettheta", "getphi",
                   "getke"):
            np.testing.assert_array_equal(getattr(p, nm)(), A["p%d/%s" % (k, nm)], err_msg=nm)
        # the reference's getv is sqrt(1 - gamma^2) c (NaN); here the speed
        v = np.sqrt(A["p%d/getvx" % k] ** 2 + A["p%d/getvy" % k] ** 2 + A["p%d/getvz" % k] ** 2)
        np.testing.assert_allclose(p.getv(), v, rtol=1e-14)


def test_gc_host_accessors_match_reference(golden):
    A, G = golden("accessors"), golden("gcs")
    off = np.r_[0, np.cumsum(G["c2/len"])]
    for k in range(3):
        g = _gc_from(G["c2/traj"][off[k]:off[k + 1]], G["c2/mu"][k])
        for nm in ("gett", "getx", "gety", "getz", "getpp", "getr", "gettheta", "getphi"):
            np.testing.assert_array_equal(getattr(g, nm)(), A["g%d/%s" % (k, nm)], err_msg=nm)


@pytest.mark.parametrize("kind", ["Particle", "GuidingCenter", "Adaptive", "BounceCenter"])
def test_save_load_round_trip(tmp_path, kind):
    f = F.EarthDipole()
    if kind == "Particle":
        t = rapt.Particle([5 * K.Re, 0, 0], [1e5, 2e6, 3e5], 0.5, K.m_pr, K.e, f)
        t.trajectory = np.vstack([t.trajectory, t.trajectory * 1.001])
    elif kind == "GuidingCenter":
        t = rapt.GuidingCenter([5 * K.Re, 0, 0], 2e7, 35.0, None, 0.0, K.m_el, -K.e, f)
        t.check_adiabaticity = True
    elif kind == "Adaptive":
        t = rapt.Adaptive([5 * K.Re, 0, 0], [1e5, 2e6, 3e5], 0.0, K.m_pr, K.e, f)
    else:
        t = rapt.BounceCenter([5 * K.Re, 0, 0], 2e7, 0.0, 40.0, K.m_pr, K.e, f)
    fn = str(tmp_path / (kind + ".pkl"))
    t.save(fn)
    u = type(t).__new__(type(t))
    u.load(fn)
    assert set(u.__dict__) == set(t.__dict__)
    for k, v in t.__dict__.items():
        w = u.__dict__[k]
        if isinstance(v, np.ndarray):
            np.testing.assert_array_equal(w, v, err_msg=k)
        elif isinstance(v, (int, float, bool, str)) or v is None:
            assert w == v, k
    if kind == "Adaptive":
        assert [type(x) for x in u.trajlist] == [type(x) for x in t.trajlist]
        np.testing.assert_array_equal(u.trajlist[0].trajectory, t.trajlist[0].trajectory)
    assert type(u.field) is type(t.field)
    assert u.field.descriptor() == t.field.descriptor()


def test_adaptive_host_start_setke_setpa():
    """Adaptive.__init__ starts as a GuidingCenter exactly when the particle is
    adiabatic (rapt/Adaptive.py:96-104); setke keeps the direction and sets
    the speed, setpa the pitch angle, and both restart the tracer list
    (rapt/Adaptive.py:140-186)."""
    f = F.CurrentSheet()
    wl = W.current_sheet_protons(n=24, seed=5)
    kinds = set()
    for i in range(24):
        a = rapt.Adaptive(wl["pos"][i], wl["vel"][i], 0.0, K.m_pr, K.e, f)
        ad = a.p.isadiabatic()
        assert isinstance(a.trajlist[0], rapt.GuidingCenter if ad else rapt.Particle)
        kinds.add(ad)
        assert len(a.getx()) == 1
        a.setke(3e3)
        v = a.vel
        assert np.sqrt(v @ v) == pytest.approx(rapt.utils.speedfromKE(3e3, K.m_pr), rel=1e-12)
        u0 = wl["vel"][i] / np.linalg.norm(wl["vel"][i])
        np.testing.assert_allclose(v / np.linalg.norm(v), u0, rtol=0, atol=1e-12)
        assert isinstance(a.trajlist[0], rapt.GuidingCenter if a.p.isadiabatic()
                          else rapt.Particle)
        a.setpa(30.0)
        b = f.unitb(a.p.trajectory[-1, :4])
        cosa = (a.vel @ b) / np.linalg.norm(a.vel)
        assert cosa == pytest.approx(np.cos(np.radians(30.0)), rel=1e-12)
        assert len(a.trajlist) == 1 and a._fresh
    assert kinds == {True, False}


def test_batched_advance_launches_checking_tracers_apart(monkeypatch):
    """A tracer that does not check adiabaticity is never launched with the
    check on, so it is not stopped at another tracer's switch point."""
    from rapt_amd import batch
    calls = []

    class Res:
        def __init__(self, n, check):
            self.rows = [np.zeros((2, 7)) + (1 if check else 2) for _ in range(n)]
            self.tcur = np.full(n, 1.0)
            self.status = np.full(n, 2 if check else 1, np.int32)

    def fake(field, last, mass, charge, delta, params=None, check=False, **kw):
        calls.append((len(last), check))
        return Res(len(last), check)

    monkeypatch.setattr(batch, "advance_particles", fake)
    f = F.EarthDipole()
    ps = [rapt.Particle([5 * K.Re, 0, 0], [1e5, 2e6, 3e5], 0.0, K.m_pr, K.e, f)
          for _ in range(3)]
    ps[1].check_adiabaticity = True
    with pytest.raises(rapt.Adiabatic):
        batch.advance(ps, 1.0)
    assert sorted(calls) == [(1, True), (2, False)]
    assert np.all(ps[0].trajectory[1:] == 2) and np.all(ps[2].trajectory[1:] == 2)
    assert np.all(ps[1].trajectory[1:] == 1)
