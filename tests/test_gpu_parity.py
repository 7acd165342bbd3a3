"""GPU parity: librapt_hip.so (through rapt_amd's API and C-ABI) against the
reference's golden trajectories and the CPU oracle.

Tolerances (written per test):
  - fixed step count (DOP853 with first_step = max_step = h): positions and
    momenta within 1e-10 relative of the oracle;
  - adaptive runs: positions at the output times within 1e-7 relative;
  - energy / first-invariant drift no larger than the oracle's (+ one part
    in 1e-9 of slack for rounding).
"""
import numpy as np
import pytest

import rapt_amd as rapt
import rapt_amd.fields as F
from rapt_amd import _native as N
from rapt_amd import batch
from rapt_amd import constants as K
from rapt_amd import workloads as W

pytestmark = pytest.mark.gpu

PARAMS = dict(rapt.params)


def P(**kw):
    d = dict(PARAMS)
    d.update(kw)
    return d


def rel_err(a, b, axis=0):
    scale = np.max(np.abs(b), axis=axis)
    scale = np.where(scale == 0, 1.0, scale)
    return np.max(np.abs(a - b) / scale, axis=axis)


def pos_err(a, b):
    """max |r_a - r_b| / |r_b| over rows (positions in columns 1..3)."""
    d = np.linalg.norm(a[:, 1:4] - b[:, 1:4], axis=1)
    return np.max(d / np.linalg.norm(b[:, 1:4], axis=1))


def mom_err(a, b):
    d = np.linalg.norm(a[:, 4:7] - b[:, 4:7], axis=1)
    return np.max(d / np.linalg.norm(b[:, 4:7], axis=1))


FIELDS = {
    "EarthDipole": lambda: F.EarthDipole(),
    "DoubleDipole": lambda: F.DoubleDipole(),
    "UniformBz": lambda: F.UniformBz(Bz=2e-7),
    "UniformCrossedEB": lambda: F.UniformCrossedEB(Ey=1e-3, Bz=1e-7),
    "VarEarthDipole": lambda: F.VarEarthDipole(amp=0.2, period=7.0),
    "Parabolic": lambda: F.Parabolic(B0=10.0, Bn=1.0, d=0.2),
    "CurrentSheet": lambda: F.CurrentSheet(),
    "DipolePlusUniform": lambda: F.DipolePlusUniform(),
}


@pytest.mark.parametrize("name", sorted(FIELDS))
def test_field_eval_device_vs_reference(golden, name):
    """Device field models against the reference's values (rapt/fields.py).
    Values to a few ulp; finite-difference derivatives to the ulp level
    amplified by r/gradientstepsize (EarthDipole: ~4e6)."""
    g = golden("fields")
    f = FIELDS[name]()
    ctx = N.context()
    tp = g[name + "/tpos"]
    for w, tol in (("B", 1e-14), ("E", 0), ("unitb", 1e-14), ("magB", 1e-14),
                   ("gradB", 2e-8), ("curlb", 2e-8), ("jacobianB", 2e-8),
                   ("lengthscale", 2e-8), ("dbdt", 2e-8)):
        ref = g[name + "/" + w]
        got = ctx.field_eval(f.descriptor(), w, tp)
        if ref.shape != got.shape:
            assert np.all(ref == 0) and np.all(got == 0), w
            continue
        if name == "CurrentSheet" and w in ("gradB", "curlb", "jacobianB", "lengthscale"):
            tol = 1e-6  # tanh differs by an ulp; d = lam/1000 amplifies it
        if w == "dbdt":
            # b(t) of these models has a fixed direction: the reference's value is
            # rounding noise of order ulp / timederivstepsize
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11, err_msg=w)
            continue
        finite = np.isfinite(ref)
        np.testing.assert_array_equal(np.isfinite(got), finite, err_msg=w)
        scale = np.max(np.abs(ref[finite])) if finite.any() else 1.0
        np.testing.assert_allclose(got[finite], ref[finite], rtol=0, atol=tol * scale,
                                   err_msg="%s.%s" % (name, w))


def _golden_particles(golden):
    Pz = golden("particles")
    lens = Pz["c1/len"]
    off = np.r_[0, np.cumsum(lens)]
    return [Pz["c1/traj"][off[i]:off[i + 1]] for i in range(len(lens))], Pz["c1/tcur"]


def test_particle_config1_vs_reference(golden):
    """Six dipole protons, Particle.advance(10) at cyclotronresolution 20:
    the GPU rows at the reference's output times within 1e-7."""
    refs, tcurs = _golden_particles(golden)
    last = np.stack([r[0] for r in refs])
    res = batch.advance_particles(F.EarthDipole(), last, K.m_pr, K.e, 10.0,
                                  params=P(cyclotronresolution=20))
    assert np.all(res.status == N.ST_DONE)
    for i, ref in enumerate(refs):
        got = np.vstack([ref[:1], res.rows[i]])
        assert got.shape == ref.shape
        # output step from the cyclotron period: same formula, rounded differently
        # (a few ulp in dt, accumulated over ~1000 rows)
        np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-12, atol=0)
        assert pos_err(got, ref) < 1e-7
        assert mom_err(got, ref) < 1e-7
        assert res.tcur[i] == pytest.approx(tcurs[i], rel=1e-12)


def test_particle_api_single(golden):
    """rapt_amd.Particle(...).advance routes to the GPU kernel."""
    refs, _ = _golden_particles(golden)
    ref = refs[0]
    rapt.params["cyclotronresolution"] = 20
    try:
        mom = ref[0, 4:]
        gm = np.sqrt(K.m_pr ** 2 + np.dot(mom, mom) / K.c ** 2)
        p = rapt.Particle(ref[0, 1:4], mom / gm, 0.0, K.m_pr, K.e, F.EarthDipole())
        p.advance(10.0)
    finally:
        rapt.params["cyclotronresolution"] = PARAMS["cyclotronresolution"]
    assert p.trajectory.shape == ref.shape
    assert pos_err(p.trajectory, ref) < 1e-7


def test_particle_fixed_step_vs_oracle(oracle):
    """Fixed step count (first_step = max_step = h, non-binding tolerances):
    GPU and oracle agree to 1e-10 in positions and momenta, 64 protons of the
    dipole workload over the full 10 s."""
    wl = W.dipole_protons_1mev(n=64)
    m, q = K.m_pr, K.e
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / K.c ** 2)
    last = np.zeros((64, 7))
    last[:, 1:4] = wl["pos"]
    last[:, 4:7] = (m * g)[:, None] * wl["vel"]
    f = F.EarthDipole()
    prm = P(cyclotronresolution=20, solvertolerances=(1e6, 1e6))
    fo = oracle.field_from_desc(f.descriptor())
    # one fixed h per particle: a quarter of its output step
    Bm = N.context().field_eval(f.descriptor(), "magB", last[:, :4])[:, 0]
    hstep = 2 * np.pi * g * m / Bm / q / 20 / 4
    for i in range(0, 64, 8):
        res = batch.advance_particles(f, last[i:i + 1], m, q, 10.0, params=prm,
                                      first_step=hstep[i], max_step=hstep[i])
        rows_o, info = oracle.particle_advance(
            fo, oracle.particle_params(rtol=1e6, atol=1e6, cyclotronresolution=20,
                                       first_step=hstep[i], max_step=hstep[i]),
            last[i], m, q, 10.0)
        got = res.rows[0]
        assert got.shape == rows_o.shape
        assert res.naccpt[0] == info["naccpt"]
        assert pos_err(got, rows_o) < 1e-10
        assert mom_err(got, rows_o) < 1e-10


def test_particle_energy_drift_not_worse(oracle):
    """Static dipole: |p| is conserved; the GPU's drift over 10 s is no
    larger than the oracle's (the reference's arithmetic) plus 1e-9."""
    wl = W.dipole_protons_1mev(n=32)
    m, q = K.m_pr, K.e
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / K.c ** 2)
    last = np.zeros((32, 7))
    last[:, 1:4] = wl["pos"]
    last[:, 4:7] = (m * g)[:, None] * wl["vel"]
    f = F.EarthDipole()
    res = batch.advance_particles(f, last, m, q, 10.0, params=P(cyclotronresolution=20))
    fo = oracle.field_from_desc(f.descriptor())
    for i in range(32):
        rows_o, _ = oracle.particle_advance(fo, oracle.particle_params(cyclotronresolution=20),
                                            last[i], m, q, 10.0)
        p0 = np.linalg.norm(last[i, 4:7])
        dg = np.max(np.abs(np.linalg.norm(res.rows[i][:, 4:7], axis=1) / p0 - 1))
        do = np.max(np.abs(np.linalg.norm(rows_o[:, 4:7], axis=1) / p0 - 1))
        assert dg <= do + 1e-9, (i, dg, do)
        assert pos_err(res.rows[i], rows_o) < 1e-7


def _gc_golden(golden):
    G = golden("gcs")
    lens = G["c2/len"]
    off = np.r_[0, np.cumsum(lens)]
    return G, [G["c2/traj"][off[i]:off[i + 1]] for i in range(len(lens))]


def test_gc_config2_vs_reference(golden):
    """Six DoubleDipole electrons, GuidingCenter.advance(3), reference output
    step: positions at the reference's output times within 1e-7."""
    G, refs = _gc_golden(golden)
    last = np.stack([r[0] for r in refs])
    res = batch.advance_gcs(F.DoubleDipole(), last, K.m_el, -K.e, G["c2/mu"], G["c2/v"],
                            G["c2/dt"], 3.0, params=P(solvertolerances=(1e-8, 1e-8)))
    assert np.all(res.status == N.ST_DONE)
    for i, ref in enumerate(refs):
        got = np.vstack([ref[:1], res.rows[i]])
        assert got.shape == ref.shape
        np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-12, atol=0)
        assert pos_err(got, ref) < 1e-7
        assert rel_err(got[:, 4], ref[:, 4]) < 1e-6
        assert abs(res.tcur[i] - G["c2/tcur"][i]) <= 1e-12 * G["c2/tcur"][i]


@pytest.mark.parametrize("name,field,eom", [
    ("bc", lambda: F.DoubleDipole(), "BrizardChanEOM"),
    ("nt", lambda: F.DoubleDipole(), "NorthropTellerEOM"),
    ("earth", lambda: F.EarthDipole(), "TaoChanBrizardEOM"),
    ("dipunif", lambda: F.DipolePlusUniform(), "TaoChanBrizardEOM"),
    ("sheet", lambda: F.CurrentSheet(), "TaoChanBrizardEOM"),
])
def test_gc_variants_vs_reference(golden, name, field, eom):
    G = golden("gcs")
    ref = G[name + "/traj"]
    v, pa, mu, dt, delta, tcur = G[name + "/in"]
    tol = (1e-8, 1e-8) if name in ("bc", "nt") else (1.49012e-8, 1.49012e-8)
    m, q = (K.m_pr, K.e) if name == "sheet" else (K.m_el, -K.e)
    res = batch.advance_gcs(field(), ref[:1], m, q, mu, v, dt, delta,
                            params=P(solvertolerances=tol), eom=eom)
    got = np.vstack([ref[:1], res.rows[0]])
    assert got.shape == ref.shape
    assert pos_err(got, ref) < 1e-7


def test_gc_api_bounceperiod(golden):
    """rapt_amd.GuidingCenter.advance with the default output step
    (bounceperiod/bounceresolution) reproduces the reference."""
    G, refs = _gc_golden(golden)
    rapt.params["solvertolerances"] = (1e-8, 1e-8)
    try:
        g = rapt.GuidingCenter(list(G["c2/pos"][0]), float(G["c2/v"][0]), float(G["c2/pa"][0]),
                               None, 0, K.m_el, -K.e, F.DoubleDipole())
        g.advance(3.0)
    finally:
        rapt.params["solvertolerances"] = PARAMS["solvertolerances"]
    assert g.trajectory.shape == refs[0].shape
    assert pos_err(g.trajectory, refs[0]) < 1e-7


def test_resume_after_row_capacity_is_exact():
    """A run cut into many launches by a tiny row buffer continues exactly:
    identical rows to one uninterrupted launch."""
    wl = W.doubledipole_electrons_gc(n=16)
    f = F.DoubleDipole()
    g = 1 / np.sqrt(1 - (wl["v"] / K.c) ** 2)
    ppar = g * K.m_el * wl["v"] * np.cos(np.radians(wl["pa"]))
    B = N.context().field_eval(f.descriptor(), "magB", np.c_[np.zeros(16), wl["pos"]])[:, 0]
    mu = g ** 2 * K.m_el * (wl["v"] * np.sin(np.radians(wl["pa"]))) ** 2 / (2 * B)
    last = np.c_[np.zeros(16), wl["pos"], ppar]
    prm = P(solvertolerances=(1e-8, 1e-8))
    one = batch.advance_gcs(f, last, K.m_el, -K.e, mu, wl["v"], 0.05, 2.0, params=prm)
    cut = batch.advance_gcs(f, last, K.m_el, -K.e, mu, wl["v"], 0.05, 2.0, params=prm,
                            row_cap=np.full(16, 7))
    assert cut.launches >= 3
    for a, b in zip(one.rows, cut.rows):
        np.testing.assert_array_equal(a, b)


def _gc_batch(field, n, seed_pa=None):
    wl = W.doubledipole_electrons_gc(n=n)
    g = 1 / np.sqrt(1 - (wl["v"] / K.c) ** 2)
    pa = np.radians(wl["pa"])
    ppar = g * K.m_el * wl["v"] * np.cos(pa)
    B = N.context().field_eval(field.descriptor(), "magB", np.c_[np.zeros(n), wl["pos"]])[:, 0]
    mu = g ** 2 * K.m_el * (wl["v"] * np.sin(pa)) ** 2 / (2 * B)
    return np.c_[np.zeros(n), wl["pos"], ppar], mu, wl["v"]


@pytest.mark.parametrize("name,field,eom", [
    ("tcb", lambda: F.DoubleDipole(), "TaoChanBrizardEOM"),
    ("bc", lambda: F.DoubleDipole(), "BrizardChanEOM"),
    ("nt", lambda: F.DoubleDipole(), "NorthropTellerEOM"),
    ("vardipole", lambda: F.VarEarthDipole(amp=0.2, period=7.0), "TaoChanBrizardEOM"),
    ("crossedEB", lambda: F.UniformCrossedEB(Ey=1e-3, Bz=1e-7), "TaoChanBrizardEOM"),
    ("dipunif", lambda: F.DipolePlusUniform(), "TaoChanBrizardEOM"),
])
@pytest.mark.parametrize("lpp", [4, 2, 3, 8, 5])
def test_gc_grouped_lanes_match_single_lane(name, field, eom, lpp):
    """Four lanes per guiding centre (the default), two lanes, and the mixed
    layout (3: the front of the work order on quads, the rest on pairs)
    against one lane per guiding centre: bit-identical rows, step counts and
    statuses.  The library is built without FMA contraction and
    the kernels form every stencil point, sum and norm in the same order, so
    any difference is a defect."""
    f = field()
    n = 40
    last, mu, v = _gc_batch(f, n)
    prm = P(solvertolerances=(1e-8, 1e-8))
    a = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 1.0, params=prm, eom=eom,
                          lanes_per_particle=1)
    b = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 1.0, params=prm, eom=eom,
                          lanes_per_particle=lpp)
    np.testing.assert_array_equal(a.status, b.status)
    np.testing.assert_array_equal(a.naccpt, b.naccpt)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(y, x)
    np.testing.assert_array_equal(b.tcur, a.tcur)
    np.testing.assert_array_equal(b.state, a.state)


@pytest.mark.parametrize("lpp", [4, 2, 3, 8, 5])
def test_gc_grouped_refill_and_adiabaticity_check(lpp):
    """Two- and four-lane kernels with fewer resident groups than guiding
    centres (work refill) and the adiabaticity check on: bit-identical rows
    and statuses to the one-lane kernel."""
    f = F.DoubleDipole()
    n = 150
    last, mu, v = _gc_batch(f, n)
    prm = P(solvertolerances=(1e-8, 1e-8), epss=0.02)
    a = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 1.0, params=prm, check=True,
                          lanes_per_particle=1)
    b = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 1.0, params=prm, check=True,
                          lanes_per_particle=lpp, lanes=64)
    np.testing.assert_array_equal(a.status, b.status)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(y, x)


@pytest.mark.parametrize("lpp", [1, 2, 3, 4, 5, 8])
def test_work_order_does_not_change_results(lpp):
    """The library's work order (shortest output step first, the default)
    and input order give bit-identical per-particle results."""
    f = F.DoubleDipole()
    n = 96
    last, mu, v = _gc_batch(f, n)
    dt = np.random.default_rng(3).uniform(0.02, 0.2, n)
    prm = P(solvertolerances=(1e-8, 1e-8))
    a = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, dt, 0.6, params=prm,
                          lanes_per_particle=lpp, lanes=32, work_order=0)
    b = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, dt, 0.6, params=prm,
                          lanes_per_particle=lpp, lanes=32, work_order=1)
    np.testing.assert_array_equal(a.status, b.status)
    np.testing.assert_array_equal(a.state, b.state)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(x, y)


def test_fp32_field_particle_tracks_fp64():
    """FP32-field Particle mode (field and force in float, state and error
    control in double, I-controller) against FP64 with the same controller
    (beta = 0) over 1 s of 1 MeV protons in EarthDipole: positions within
    1e-5 relative and |p| drift below 1e-5.  Float rounding of the force
    (~6e-8 per evaluation) accumulates over the run."""
    wl = W.dipole_protons_1mev(n=64)
    m, q = K.m_pr, K.e
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / K.c ** 2)
    last = np.zeros((64, 7))
    last[:, 1:4] = wl["pos"]
    last[:, 4:7] = (m * g)[:, None] * wl["vel"]
    a = batch.advance_particles(F.EarthDipole(), last, m, q, 1.0, params=P(), beta=0.0)
    b = batch.advance_particles(F.EarthDipole(), last, m, q, 1.0, params=P(),
                                precision="fp32_field")
    assert np.all(b.status == N.ST_DONE), np.unique(b.status, return_counts=True)
    worst_pos = worst_e = 0.0
    for i, (x, y) in enumerate(zip(a.rows, b.rows)):
        assert x.shape == y.shape
        worst_pos = max(worst_pos, pos_err(y, x))
        p0 = np.linalg.norm(last[i, 4:7])
        worst_e = max(worst_e, np.max(np.abs(np.linalg.norm(y[:, 4:7], axis=1) - p0)) / p0)
    print("fp32 particle: pos %.3g, |p| drift %.3g" % (worst_pos, worst_e))
    assert worst_pos < 1e-5
    assert worst_e < 1e-5


def test_fp32_field_gc_tracks_fp64():
    """FP32-field guiding centres (field and finite-difference drift terms in
    float with step |r|/1024, algebra and state in double, I-controller)
    against FP64 with the same controller over 1 s of DoubleDipole electrons:
    positions within 1e-4 relative."""
    f = F.DoubleDipole()
    n = 64
    last, mu, v = _gc_batch(f, n)
    prm = P(solvertolerances=(1e-8, 1e-8))
    a = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 1.0, params=prm, beta=-1.0)
    b = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 1.0, params=prm,
                          precision="fp32_field")
    assert np.all(b.status == N.ST_DONE), np.unique(b.status, return_counts=True)
    worst = max(pos_err(y, x) for x, y in zip(a.rows, b.rows))
    print("fp32 gc: pos %.3g" % worst)
    assert worst < 1e-4


def test_precision_rejects_unknown_value():
    d = N.make_desc(N.MODE_GC, P(), 1.0)
    d.precision = 7
    ctx = N.context()
    ctx.set_field(F.DoubleDipole().descriptor())
    b = N.Batch(ctx, 1)
    try:
        b.upload(np.zeros((N.NSLOT, 1)), np.zeros(1, np.int32))
        b.set_rows(5, np.array([4], np.int64))
        with pytest.raises(RuntimeError, match="precision"):
            b.run(d)
    finally:
        b.close()


def test_autotuner_picks_a_candidate_and_keeps_results():
    """The autotuner returns one of its candidate shapes, and a run with the
    tuned shape gives the same rows as the default launch."""
    from rapt_amd import autotune as AT
    f = F.DoubleDipole()
    n = 256
    last, mu, v = _gc_batch(f, n)
    prm = P(solvertolerances=(1e-8, 1e-8))
    S = batch.gc_state(last, K.m_el, -K.e, mu, v, 0.05)
    cap = batch._rows_needed(0.5, np.full(n, 0.05))
    desc = N.make_desc(N.MODE_GC, prm, 0.5)
    shape, table = AT.tune(f, S, np.zeros(n, np.int32), cap, desc, force=True)
    assert shape in [t for t, _ in table]
    assert shape in AT.candidates(desc.mode, desc.stepper, desc.precision, n)
    a = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 0.5, params=prm)
    b = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, 0.05, 0.5, params=prm, **shape)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("lpp", [4, 3, 2, 1, 5])
def test_output_epochs_reproduce_one_launch(lpp):
    """A run in hipGraph-replayed output epochs of 7 rows per particle gives
    each particle's rows of one uninterrupted launch, epoch by epoch (every
    graph replay checked against the single launch)."""
    f = F.DoubleDipole()
    n = 64
    last, mu, v = _gc_batch(f, n)
    dt = np.random.default_rng(5).uniform(0.03, 0.08, n)
    prm = P(solvertolerances=(1e-8, 1e-8))
    one = batch.advance_gcs(f, last, K.m_el, -K.e, mu, v, dt, 1.0, params=prm,
                            lanes_per_particle=lpp)
    S = batch.gc_state(last, K.m_el, -K.e, mu, v, dt)
    desc = N.make_desc(N.MODE_GC, prm, 1.0, lanes_per_particle=lpp)
    ctx = N.context()
    ctx.set_field(f.descriptor())
    b = N.Batch(ctx, n)
    got = [[] for _ in range(n)]
    epochs = 0
    try:
        b.upload(S, np.zeros(n, np.int32))
        for rows, nrows, status, left, cnt in b.epochs(desc, 7):
            epochs += 1
            for i in range(n):
                k0 = sum(len(r) for r in got[i])
                seg = rows[i, :nrows[i]]
                np.testing.assert_array_equal(seg, one.rows[i][k0:k0 + nrows[i]])
                got[i].append(seg)
    finally:
        b.close()
    assert epochs >= 3
    for i in range(n):
        np.testing.assert_array_equal(np.vstack(got[i]), one.rows[i])
    np.testing.assert_array_equal(status, one.status)


@pytest.mark.parametrize("lpp", [4, 3, 2, 1, 5])
def test_cost_work_order_does_not_change_results(lpp):
    """A work order set by cost (Batch.set_cost: here a trial launch's step
    counts and a random permutation, with refill) gives the rows, states and
    statuses of the default order, bit for bit."""
    f = F.DoubleDipole()
    n = 96
    last, mu, v = _gc_batch(f, n)
    dt = np.random.default_rng(7).uniform(0.02, 0.2, n)
    prm = P(solvertolerances=(1e-8, 1e-8))
    S = batch.gc_state(last, K.m_el, -K.e, mu, v, dt)
    cap = batch._rows_needed(0.6, dt)
    desc = N.make_desc(N.MODE_GC, prm, 0.6, lanes_per_particle=lpp, lanes=48)
    ctx = N.context()
    ctx.set_field(f.descriptor())
    b = N.Batch(ctx, n)
    try:
        b.upload(S, np.zeros(n, np.int32))
        b.set_rows(5, cap)
        b.run(desc)
        ref = b.download()
        trial = N.TraceDesc.from_buffer_copy(desc)
        trial.delta = 0.06
        b.reset()
        b.run(trial)
        costs = [b.download(rows=False)["naccpt"].astype(float),
                 np.random.default_rng(8).permutation(n).astype(float)]
        for cost in costs:
            b.set_cost(cost)
            b.reset()
            b.run(desc)
            got = b.download()
            for k in ("status", "nrows", "naccpt", "state", "rows"):
                np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    finally:
        b.close()


def test_work_refill_is_exact():
    """Fewer resident lanes than particles (work refill) gives the same rows."""
    wl = W.dipole_protons_1mev(n=200)
    m, q = K.m_pr, K.e
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / K.c ** 2)
    last = np.zeros((200, 7))
    last[:, 1:4] = wl["pos"]
    last[:, 4:7] = (m * g)[:, None] * wl["vel"]
    a = batch.advance_particles(F.EarthDipole(), last, m, q, 1.0, params=P())
    b = batch.advance_particles(F.EarthDipole(), last, m, q, 1.0, params=P(), lanes=64)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(x, y)


def test_adaptive_current_sheet_vs_oracle(oracle):
    """Adaptive runs in the current sheet: same initial mode and the same
    rows up to the first switch (1e-7); switch counts agree for most
    particles (orbits after a switch are chaotic, as the reference notes)."""
    wl = W.current_sheet_protons(n=16)
    f = F.CurrentSheet()
    prm = P(GCtimestep=0.5)
    res = batch.advance_adaptive(f, wl["pos"], wl["vel"], 0.0, K.m_pr, K.e, 20.0, params=prm)
    fo = oracle.field_from_desc(f.descriptor())
    po = oracle.particle_params(gctimestep=0.5)
    same_switch = 0
    for i in range(16):
        rows_o, info = oracle.adaptive_run(fo, po, wl["pos"][i], wl["vel"][i], 0.0, K.m_pr, K.e,
                                           20.0)
        got = res.rows[i]
        assert got[0, 0] == rows_o[0, 0]  # initial mode
        # rows up to and including the first switch row
        ko = np.argmax(np.r_[rows_o[1:, 0] != rows_o[0, 0], True])
        kg = np.argmax(np.r_[got[1:, 0] != got[0, 0], True])
        k = min(ko, kg)
        a, b = got[:k + 1, 1:5], rows_o[:k + 1, 1:5]
        np.testing.assert_allclose(a[:, 0], b[:, 0], rtol=1e-12)
        d = np.linalg.norm(a[:, 1:4] - b[:, 1:4], axis=1) / np.linalg.norm(b[:, 1:4], axis=1)
        assert np.max(d) < 1e-7
        same_switch += int(np.sum(np.diff(got[:, 0]) != 0) == info["nswitch"])
    assert same_switch >= 12


def test_abi_errors_are_reported():
    """Bad descriptors come back as error codes with a message, not a crash."""
    ctx = N.context()
    bad = dict(model=42, params=(), static=True, gstep=1.0, tstep=1e-3)
    with pytest.raises(RuntimeError, match="unknown field model"):
        ctx.set_field(bad)
    ctx._field_key = None
    with pytest.raises(RuntimeError, match="GCtimestep"):
        batch.advance_adaptive(F.CurrentSheet(), [[K.Re, 0, 0]], [[1e5, 0, 0]], 0.0, K.m_pr, K.e,
                               1.0, params=P(GCtimestep=0))


def test_halfbounce_device_vs_reference(golden):
    """Device half-bounce path S_b (field-line trace, spline, mirror points,
    integral) against flutils.halfbouncepath of the reference, 48 cases in
    the DoubleDipole and EarthDipole fields.  The reference integrates with
    QUADPACK at epsrel 1e-4, which is also the agreement bar here."""
    rows = golden("bounce")["halfbouncepath"]
    ctx = N.context()
    for fid in (0, 1):
        sel = rows[rows[:, 0] == fid]
        f = F.DoubleDipole() if fid == 0 else F.EarthDipole()
        L, ph, Bm, sb = sel[:, 1], sel[:, 3], sel[:, 4], sel[:, 5]
        tpos = np.c_[np.zeros(len(L)), L * K.Re * np.cos(ph), L * K.Re * np.sin(ph),
                     np.zeros(len(L))]
        got, st = ctx.halfbounce(f.descriptor(), tpos, Bm, rapt.params["fieldlineresolution"])
        assert np.all(st == 1), st
        np.testing.assert_allclose(got, sb, rtol=1e-4)
        assert np.median(np.abs(got / sb - 1)) < 5e-6
