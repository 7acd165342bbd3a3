"""User-written fields on the device: a ``_Field`` subclass overriding
``B``/``E`` in Python (rapt/fields.py:5-74), traced into a register program
that the kernels interpret in FP64.  Checked against the reference's values
and trajectories for the same class bodies (tests/golden/userfields.npz) and
against the oracle, which interprets the same program with the C library's
math and matches the reference bit for bit on these cases.

Tolerances: fields within 4e-15 relative of the reference (the device's pow,
exp, tanh, sin, cos and cosh differ from the C library's by an ulp or two);
finite-difference derivatives within 1e-8 of the oracle (they amplify those
ulps by scale/step); trajectories at the output times within 1e-7 relative in
position, the bar of every adaptive comparison.
"""
import numpy as np
import pytest

import rapt_amd as rapt
from rapt_amd import _native as N
from rapt_amd import batch
from rapt_amd import constants as K
from rapt_amd import fields as F

pytestmark = pytest.mark.gpu

PARAMS = dict(rapt.params)


def P(**kw):
    d = dict(PARAMS)
    d.update(kw)
    return d


def pos_err(a, b):
    d = np.linalg.norm(a[:, 1:4] - b[:, 1:4], axis=1)
    return np.max(d / np.linalg.norm(b[:, 1:4], axis=1))


FIELDS = {"ChargedDipole": lambda U: U["ChargedDipole"](),
          "TwinDipole": lambda U: U["TwinDipole"](),
          "BreathingSheet": lambda U: U["BreathingSheet"]()}


@pytest.mark.parametrize("name", sorted(FIELDS))
def test_device_field_values(oracle, golden, user_fields, name):
    g = golden("userfields")
    f = FIELDS[name](user_fields)
    d = f.descriptor()
    tp = g[name + "/tpos"]
    ctx = N.context()
    for w in ("B", "E"):
        ref = g[name + "/" + w]
        got = ctx.field_eval(d, w, tp)
        np.testing.assert_allclose(got, ref, rtol=4e-15, atol=4e-15 * np.max(np.abs(ref)),
                                   err_msg=w)
    fo = oracle.field_from_desc(d)
    for w in ("magB", "gradB", "curvature", "lengthscale", "dBdt", "jacobianB", "curlb"):
        if w == "dBdt" and f.static:
            continue
        ref = oracle.field_eval(fo, w, tp)
        got = ctx.field_eval(d, w, tp)
        np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-8 * np.max(np.abs(ref)),
                                   err_msg=w)


def test_particle_object_in_charged_dipole(golden, user_fields):
    """The reference's new-field example through the object API: a 1 eV
    proton at 5 m in a charged dipole, setke(1), advance(4e-4)."""
    g = golden("userfields")
    old = dict(rapt.params)
    try:
        rapt.params.clear()
        rapt.params.update(PARAMS)
        p = rapt.Particle([5, 0, 0], [0, 1, 0], 0, rapt.m_pr, rapt.e,
                          user_fields["ChargedDipole"](Q=1e-6))
        p.setke(1)
        p.advance(4e-4)
    finally:
        rapt.params.clear()
        rapt.params.update(old)
    ref = g["cd_particle/traj"]
    assert p.trajectory.shape == ref.shape
    assert pos_err(p.trajectory, ref) < 1e-7
    assert p.tcur == g["cd_particle/tcur"][0]


@pytest.mark.parametrize("name,cls", [("twin_particle", "TwinDipole"),
                                      ("breathing_particle", "BreathingSheet")])
def test_particle_user_fields_vs_reference(golden, user_fields, name, cls):
    g = golden("userfields")
    ref = g[name + "/traj"]
    m, q, delta = g[name + "/in"][6:9]
    res = batch.advance_particles(user_fields[cls](), ref[:1], m, q, delta, params=P())
    assert res.status[0] == N.ST_DONE
    got = np.vstack([ref[:1], res.rows[0]])
    assert got.shape == ref.shape
    assert pos_err(got, ref) < 1e-7


@pytest.mark.parametrize("name,make", [
    ("cd_gc", lambda U: U["ChargedDipole"](Q=1e-8)),
    ("twin_gc", lambda U: U["TwinDipole"]()),
    ("breathing_gc", lambda U: U["BreathingSheet"]()),
])
def test_gc_user_fields_vs_reference(golden, user_fields, name, make):
    """GuidingCenter.advance in user fields: one with E, one whose B varies in
    time (dB/dt by central differences in t)."""
    g = golden("userfields")
    ref = g[name + "/traj"]
    v, pa, m, q, mu, gdt, delta = g[name + "/in"][3:]
    res = batch.advance_gcs(make(user_fields), ref[:1], m, q, [mu], [v], [gdt], delta,
                            params=P(GCtimestep=gdt))
    assert res.status[0] == N.ST_DONE
    got = np.vstack([ref[:1], res.rows[0]])
    assert got.shape == ref.shape
    assert pos_err(got, ref) < 1e-7


def test_gc_object_user_field(golden, user_fields):
    """GuidingCenter through the object API with GCtimestep set (the default
    output step needs a field-line trace, which user fields do not take)."""
    g = golden("userfields")
    ref = g["breathing_gc/traj"]
    pos, (v, pa, m, q, mu, gdt, delta) = g["breathing_gc/in"][:3], g["breathing_gc/in"][3:]
    old = dict(rapt.params)
    try:
        rapt.params.clear()
        rapt.params.update(P(GCtimestep=gdt))
        gc = rapt.GuidingCenter(list(pos), v, pa, None, 0, m, q,
                                user_fields["BreathingSheet"]())
        gc.advance(delta)
    finally:
        rapt.params.clear()
        rapt.params.update(old)
    assert gc.trajectory.shape == ref.shape
    assert pos_err(gc.trajectory, ref) < 1e-7


def test_charged_dipole_batch_vs_oracle(oracle, user_fields):
    """64 protons (1-100 eV, 3-8 m) in the charged dipole, one launch, each
    against the oracle in the same traced field."""
    rng = np.random.default_rng(17)
    n = 64
    r = rng.uniform(3.0, 8.0, n)
    th = np.arccos(rng.uniform(-0.6, 0.6, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    ke = 10 ** rng.uniform(0, 2, n)
    v = np.sqrt(2 * ke * K.e / K.m_pr)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    last = np.zeros((n, 7))
    last[:, 1:4] = np.c_[r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph),
                         r * np.cos(th)]
    last[:, 4:7] = K.m_pr * v[:, None] * d
    f = user_fields["ChargedDipole"](Q=1e-7)
    res = batch.advance_particles(f, last, K.m_pr, K.e, 1e-4, params=P())
    fo = oracle.field_from_desc(f.descriptor())
    worst = 0.0
    for i in range(n):
        rows_o, info = oracle.particle_advance(fo, oracle.particle_params(), last[i], K.m_pr,
                                               K.e, 1e-4)
        assert res.status[i] == N.ST_DONE and info["status"] == 1, i
        assert res.rows[i].shape == rows_o.shape, i
        worst = max(worst, pos_err(res.rows[i], rows_o))
    assert worst < 1e-7, worst


def test_program_switch_between_launches(user_fields):
    """Two user fields in turn on one context: each launch runs the program
    of the field it was given (a stale program would give the other field's
    numbers)."""
    ctx = N.context()
    tp = np.array([[0.0, 3.0, 1.0, 0.5], [1.0, 4.0, -2.0, 1.0]])
    a, b = user_fields["ChargedDipole"](Q=1e-6), user_fields["ChargedDipole"](Q=3e-6)
    ea = ctx.field_eval(a.descriptor(), "E", tp)
    eb = ctx.field_eval(b.descriptor(), "E", tp)
    ea2 = ctx.field_eval(a.descriptor(), "E", tp)
    np.testing.assert_array_equal(ea, ea2)
    np.testing.assert_allclose(eb, 3 * ea, rtol=1e-15)


def test_user_field_rejections(user_fields):
    """What the device does not take for a user field fails loudly: FP32
    field precision, and BounceCenter in a field that is not static."""
    f = user_fields["TwinDipole"]()
    last = np.zeros((1, 7))
    last[0, 1:4] = [5 * K.Re, 0, 0]
    last[0, 5] = K.m_pr * 1e7
    with pytest.raises(RuntimeError, match="FP64 only"):
        batch.advance_particles(f, last, K.m_pr, K.e, 1.0, params=P(), precision="fp32_field")
    with pytest.raises(Exception, match="static"):
        bc = rapt.BounceCenter([5 * K.Re, 0, 0], 1e7, 0.0, 30.0, K.m_pr, K.e,
                               user_fields["BreathingSheet"]())
        bc.advance(10.0)


def test_user_field_bounce_paths(user_fields):
    """The field-line kernels with a user-written field (interpreted): the
    reference's two-dipole field written by the user against the built-in
    DoubleDipole (the same formula; only rounding differs, so the traces
    agree far inside the reference's own quadrature tolerance): the half-
    bounce path, the second invariant (GuidingCenter.geteye) and
    BounceCenter.advance, and the device output step of a batched
    GuidingCenter advance, which now takes user fields."""
    fu, fb = user_fields["TwinDipole"](), F.DoubleDipole()
    ctx = N.context()
    L = np.array([4.0, 5.0, 6.0, 7.0])
    tp = np.zeros((4, 4))
    tp[:, 1] = L * K.Re
    tp[:, 3] = [0.0, 0.1, -0.2, 0.05]
    tp[:, 3] *= K.Re
    Bm = np.array([3e-6, 1.5e-6, 8e-7, 4e-7])
    res = rapt.params["fieldlineresolution"]
    for fn in (ctx.halfbounce, ctx.eye):
        a, sa = fn(fu.descriptor(), tp, Bm, res)
        b, sb = fn(fb.descriptor(), tp, Bm, res)
        assert np.all(sa == 1) and np.all(sb == 1), (sa, sb)
        np.testing.assert_allclose(a, b, rtol=1e-6)
        # the compiled kernels (the default) against the interpreted ones
        ai, ac = _both_modes(lambda: fn(fu.descriptor(), tp, Bm, res))
        np.testing.assert_array_equal(ai[0], ac[0])
        np.testing.assert_array_equal(ai[1], ac[1])
    # BounceCenter.advance, three output rows
    saved = rapt.params["BCtimestep"]
    rapt.params["BCtimestep"] = 5.0
    try:
        traj = []
        for f in (fu, fb):
            bc = rapt.BounceCenter([5 * K.Re, 0, 0], 1e7, 0.0, 40.0, K.m_pr, K.e, f)
            bc.advance(150.0)  # output step 5 bounce periods, ~58 s
            traj.append(bc.trajectory)
    finally:
        rapt.params["BCtimestep"] = saved
    assert traj[0].shape == traj[1].shape and len(traj[0]) >= 3
    np.testing.assert_allclose(traj[0][:, 1:4], traj[1][:, 1:4], rtol=1e-6, atol=1e-6 * K.Re)

    def bc_run():
        rapt.params["BCtimestep"] = 5.0
        try:
            bc = rapt.BounceCenter([5 * K.Re, 0, 0], 1e7, 0.0, 40.0, K.m_pr, K.e, fu)
            bc.advance(150.0)
        finally:
            rapt.params["BCtimestep"] = saved
        return bc.trajectory
    ti, tc = _both_modes(bc_run)
    np.testing.assert_array_equal(ti, tc)
    np.testing.assert_array_equal(tc, traj[0])
    # batched GuidingCenter advance: device output steps in the user field
    gs = [rapt.GuidingCenter([l * K.Re, 0, 0], 1e7, 45.0, None, 0, K.m_pr, K.e, f)
          for f in (fu, fb) for l in (4.0, 6.0)]
    dt_u = rapt.GuidingCenter._batch_dt(gs[:2], True)
    dt_b = rapt.GuidingCenter._batch_dt(gs[2:], True)
    np.testing.assert_allclose(dt_u, dt_b, rtol=1e-6)
    np.testing.assert_allclose(dt_u, [g._dt() for g in gs[:2]], rtol=1e-4)


def test_device_selection_ops():
    """np.maximum / np.minimum / np.clip / np.sign, np.linalg.norm and np.cross
    in a user field: device values against the host evaluation of the same
    program (and numpy's special-value rules at +-0 and NaN)."""
    from rapt_amd import fieldexpr as FE
    from test_userfields import _SelectField

    class M(rapt.fields._Field):
        def B(self, tpos):
            return np.array([np.maximum(tpos[1], tpos[2]), np.minimum(tpos[1], tpos[2]),
                             np.sign(tpos[3])])
    ctx = N.context()
    for f, pts in ((_SelectField(), None), (M(), np.array(
            [[0.0, np.nan, 1.0, 0.0], [0.0, 1.0, np.nan, -0.0], [0.0, -0.0, 0.0, np.nan],
             [0.0, 0.0, -0.0, -2.0], [0.0, np.inf, -np.inf, 3.0]]))):
        if pts is None:
            rng = np.random.default_rng(22)
            pts = np.c_[rng.uniform(0, 5, 256), rng.uniform(-3, 3, (256, 3)) * K.Re]
            pts[:2, 2] = [0.0, -0.0]
        d = f.descriptor()
        got_b = ctx.field_eval(d, "B", pts)
        got_e = ctx.field_eval(d, "E", pts)
        for i, tp in enumerate(pts):
            b, e = FE.evaluate(d["expr"], tp)
            np.testing.assert_allclose(got_b[i], b, rtol=4e-16, atol=0)
            np.testing.assert_allclose(got_e[i], e, rtol=4e-16, atol=0)
            np.testing.assert_array_equal(np.signbit(got_b[i]), np.signbit(b))


def test_malformed_programs_rejected():
    """rapt_field_expr_set refuses a program that reads a register it did not
    set, leaves a result register unset, or names a constant it lacks; the
    error is raised before anything runs."""
    import ctypes as C
    ctx = N.context()
    ok = np.array([[0, 4, 0, 0], [0, 5, 0, 0], [0, 6, 0, 0]], np.int32)
    k = np.array([1.0])
    cases = [np.array([[3, 7, 1, 9], [21, 4, 7, 0], [0, 5, 0, 0], [0, 6, 0, 0]], np.int32),
             np.array([[0, 4, 0, 0], [0, 5, 0, 0]], np.int32),
             np.array([[0, 4, 3, 0], [0, 5, 0, 0], [0, 6, 0, 0]], np.int32)]
    L = N.lib()
    assert L.rapt_field_expr_set(ctx.h, 3, 0, ok.ctypes.data_as(C.POINTER(C.c_int32)), 1,
                                 k.ctypes.data_as(C.POINTER(C.c_double))) == 0
    for code in cases:
        rc = L.rapt_field_expr_set(ctx.h, len(code), 0,
                                   code.ctypes.data_as(C.POINTER(C.c_int32)), 1,
                                   k.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == -1  # RAPT_E_ARG
    # the context now holds this test's program: make the next field set upload again
    ctx._expr_key = None
    ctx._field_key = None


# ---- compiled user fields (rapt_field_expr_mode) -------------------------------------
def _both_modes(run):
    """run() once with the program interpreted and once compiled at run time,
    on the shared context; returns (interpreted, compiled)."""
    ctx = N.context()
    try:
        ctx.set_expr_mode(False)
        a = run()
        ctx.set_expr_mode(True)
        b = run()
    finally:
        ctx.set_expr_mode(True)
    return a, b


def _assert_same(a, b):
    np.testing.assert_array_equal(a.status, b.status)
    np.testing.assert_array_equal(a.naccpt, b.naccpt)
    for x, y in zip(a.rows, b.rows):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.state, b.state)


def _protons(n, seed, r_lo, r_hi, ke_lo, ke_hi):
    rng = np.random.default_rng(seed)
    r = rng.uniform(r_lo, r_hi, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    th = np.arccos(rng.uniform(-0.5, 0.5, n))
    v = np.sqrt(2 * 10 ** rng.uniform(ke_lo, ke_hi, n) * K.e / K.m_pr)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    last = np.zeros((n, 7))
    last[:, 1:4] = np.c_[r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph),
                         r * np.cos(th)]
    last[:, 4:7] = K.m_pr * v[:, None] * d
    return last


def test_compiled_particle_bit_identical(user_fields):
    """Particle kernel (DOP853) in the charged dipole (E != 0, not static):
    the run-time compiled kernel gives the interpreter's rows bit for bit."""
    last = _protons(64, 23, 3.0, 8.0, 0, 2)
    f = user_fields["ChargedDipole"](Q=1e-7)
    a, b = _both_modes(lambda: batch.advance_particles(f, last, K.m_pr, K.e, 1e-4, params=P()))
    assert np.all(a.status == N.ST_DONE)
    _assert_same(a, b)


@pytest.mark.parametrize("cls", ["TwinDipole", "BreathingSheet"])
def test_compiled_gc_bit_identical(golden, user_fields, cls):
    """GuidingCenter kernels: the static Tao-Chan-Brizard one (TwinDipole)
    and the generic one (BreathingSheet, B varying in time)."""
    g = golden("userfields")
    name = "twin_gc" if cls == "TwinDipole" else "breathing_gc"
    ref = g[name + "/traj"]
    v, pa, m, q, mu, gdt, delta = g[name + "/in"][3:]
    rows0 = np.repeat(ref[:1], 16, axis=0)
    rows0[:, 1:4] *= np.linspace(0.9, 1.1, 16)[:, None]
    f = user_fields[cls]()
    a, b = _both_modes(lambda: batch.advance_gcs(f, rows0, m, q, [mu] * 16, [v] * 16,
                                                 [gdt] * 16, delta, params=P(GCtimestep=gdt)))
    assert np.all(a.status == N.ST_DONE)
    _assert_same(a, b)


def test_compiled_adaptive_bit_identical(user_fields):
    """Adaptive kernel in the user-written current sheet with a uniform E:
    modes, switch rows and trajectories identical in both modes."""
    from rapt_amd import workloads as W
    wl = W.current_sheet_protons(n=32)
    f = user_fields["SheetField"]()
    a, b = _both_modes(lambda: batch.advance_adaptive(f, wl["pos"], wl["vel"], 0.0, K.m_pr,
                                                      K.e, 5.0, params=P(GCtimestep=0.5)))
    assert a.counters["nswitch"] > 0
    assert a.counters["nswitch"] == b.counters["nswitch"]
    _assert_same(a, b)


def test_compiled_field_speed(user_fields):
    """The point of compiling: 4,096 1 MeV protons in the user-written dipole
    plus uniform field, advance(1).  Kernel time of the compiled run at least
    4x below the interpreter's (a program of 26 operations; the built-in
    model's speed is the ceiling)."""
    from rapt_amd import workloads as W
    wl = W.dipole_protons_1mev(n=4096)
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / K.c ** 2)
    last = np.zeros((4096, 7))
    last[:, 1:4] = wl["pos"]
    last[:, 4:7] = (K.m_pr * g)[:, None] * wl["vel"]
    f = user_fields["DipolePlusUniformField"]()
    run = lambda: batch.advance_particles(f, last, K.m_pr, K.e, 1.0, params=P())
    run()  # compile and load outside the timed runs
    a, b = _both_modes(run)
    _assert_same(a, b)
    builtin = batch.advance_particles(rapt.fields.DipolePlusUniform((0.0, 0.0, -10e-9)),
                                      last, K.m_pr, K.e, 1.0, params=P())
    ti, tc, tb = (r.counters["kernel_ms"] for r in (a, b, builtin))
    print("kernel ms: interpreted %.2f compiled %.2f built-in %.2f" % (ti, tc, tb))
    assert tc * 4 < ti, (ti, tc)


@pytest.mark.parametrize("cls", ["TwinDipole", "BreathingSheet"])
def test_compiled_gc_lane_groups_bit_identical(golden, user_fields, cls):
    """A compiled user field takes the lane-group GuidingCenter kernels
    (pairs, quads, octets, the two mixed layouts; the static Tao-Chan-Brizard
    instantiation for TwinDipole, the generic one for BreathingSheet): rows,
    statuses and step counts equal the interpreted one-lane run's bit for
    bit, with refill (fewer resident lanes than particles) included."""
    g = golden("userfields")
    name = "twin_gc" if cls == "TwinDipole" else "breathing_gc"
    ref = g[name + "/traj"]
    v, pa, m, q, mu, gdt, delta = g[name + "/in"][3:]
    n = 40
    rows0 = np.repeat(ref[:1], n, axis=0)
    rows0[:, 1:4] *= np.linspace(0.9, 1.1, n)[:, None]
    f = user_fields[cls]()
    run = lambda **kw: batch.advance_gcs(f, rows0, m, q, [mu] * n, [v] * n, [gdt] * n, delta,
                                         params=P(GCtimestep=gdt), **kw)
    ctx = N.context()
    try:
        ctx.set_expr_mode(False)
        base = run(lanes_per_particle=1)
        ctx.set_expr_mode(True)
        assert np.all(base.status == N.ST_DONE)
        for kw in (dict(lanes_per_particle=2), dict(lanes_per_particle=4),
                   dict(lanes_per_particle=8), dict(lanes_per_particle=4, lanes=64),
                   dict(lanes_per_particle=3, mixed_quads=8),
                   dict(lanes_per_particle=5, mixed_octets=4, mixed_quads=8)):
            _assert_same(base, run(**kw))
    finally:
        ctx.set_expr_mode(True)


def test_compiled_user_field_output_epochs(golden, user_fields):
    """The compiled lane-group kernel captured in the hipGraph of output
    epochs (a module launch inside the capture): every epoch's rows equal
    the single compiled launch's, and those equal the interpreted run's."""
    g = golden("userfields")
    ref = g["twin_gc/traj"]
    v, pa, m, q, mu, gdt, delta = g["twin_gc/in"][3:]
    n = 24
    rows0 = np.repeat(ref[:1], n, axis=0)
    rows0[:, 1:4] *= np.linspace(0.9, 1.1, n)[:, None]
    f = user_fields["TwinDipole"]()
    prm = P(GCtimestep=gdt)
    a, one = _both_modes(lambda: batch.advance_gcs(f, rows0, m, q, [mu] * n, [v] * n,
                                                   [gdt] * n, delta, params=prm))
    _assert_same(a, one)
    S = batch.gc_state(rows0, m, q, [mu] * n, [v] * n, [gdt] * n)
    desc = N.make_desc(N.MODE_GC, prm, delta, lanes_per_particle=4)
    ctx = N.context()
    ctx.set_field(f.descriptor())
    b = N.Batch(ctx, n)
    got = [[] for _ in range(n)]
    epochs = 0
    try:
        b.upload(S, np.zeros(n, np.int32))
        for rows, nrows, status, left, cnt in b.epochs(desc, 5):
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


def test_compiled_user_field_noalloc(user_fields):
    """The allocation-free entry point with a user-written field: it loads no
    module, so it runs the compiled one-lane kernel an earlier run loaded (or
    interprets): the rows equal the batch path's either way."""
    from test_zz_no_alloc import _noalloc
    last = _protons(32, 29, 3.0, 8.0, 0, 2)
    f = user_fields["ChargedDipole"](Q=1e-7)
    a = batch.advance_particles(f, last, K.m_pr, K.e, 1e-4, params=P())
    S = batch.particle_state(last, K.m_pr, K.e)
    dt = batch.particle_dt(f, last, K.m_pr, K.e, P()["cyclotronresolution"])
    cap = batch._rows_needed(1e-4, dt)
    desc = N.make_desc(N.MODE_PARTICLE, P(), 1e-4)
    out = _noalloc(f, N.MODE_PARTICLE, S, np.zeros(32, np.int32), cap, desc)
    np.testing.assert_array_equal(out["status"], a.status)
    for x, y in zip(out["rows"], a.rows):
        np.testing.assert_array_equal(x, y)


def test_compiled_program_and_constant_switches(user_fields):
    """Compiled kernels follow the context's program: two charges of the
    charged dipole (same code, other constants: the same compiled kernel
    reads the new constant table) and another field in turn, each run equal
    to its interpreted run."""
    last = _protons(16, 31, 3.0, 8.0, 0, 2)
    fields = [user_fields["ChargedDipole"](Q=1e-7), user_fields["ChargedDipole"](Q=3e-7),
              user_fields["BreathingSheet"](), user_fields["ChargedDipole"](Q=1e-7)]
    runs = []
    for f in fields:
        a, b = _both_modes(lambda: batch.advance_particles(f, last, K.m_pr, K.e, 5e-5,
                                                           params=P()))
        _assert_same(a, b)
        runs.append(b)
    for x, y in zip(runs[0].rows, runs[3].rows):
        np.testing.assert_array_equal(x, y)
    assert not all(np.array_equal(x, y) for x, y in zip(runs[0].rows, runs[1].rows))


def test_compiled_equatorial_bounce_center(user_fields):
    """BounceCenter with isequatorial set in the user-written two-dipole field:
    the compiled k_bounce_center_eq equals the interpreted one bit for bit and
    the built-in DoubleDipole's run within 1e-6."""
    def run(f):
        bc = rapt.BounceCenter([5 * K.Re, 0.0, 0.0], 1e7, 0.0, 90.0, K.m_pr, K.e, f)
        bc.isequatorial = True
        bc.advance(200.0)
        return bc.trajectory
    ti, tc = _both_modes(lambda: run(user_fields["TwinDipole"]()))
    np.testing.assert_array_equal(ti, tc)
    tb = run(F.DoubleDipole())
    assert tc.shape == tb.shape and len(tc) >= 3
    np.testing.assert_allclose(tc[:, 1:], tb[:, 1:], rtol=0, atol=1e-6 * K.Re)


@pytest.mark.parametrize("name", sorted(FIELDS))
def test_compiled_field_eval_rhs_diagnostics(user_fields, name):
    """Field evaluation (every quantity), the right-hand sides and the per-row
    guiding-centre diagnostics of a user field: the compiled kernels equal
    the interpreted ones bit for bit."""
    f = FIELDS[name](user_fields)
    d = f.descriptor()
    ctx = N.context()
    rng = np.random.default_rng(41)
    n = 96
    tp = np.c_[rng.uniform(0, 5, n), rng.uniform(3, 8, (n, 3)) * (K.Re if name != "ChargedDipole"
                                                                  else 1.0)]
    whats = ["B", "E", "unitb", "magB", "gradB", "jacobianB", "curvature", "curlb",
             "lengthscale"] + ([] if f.static else ["dBdt", "dbdt", "timescale"])
    for w in whats:
        a, b = _both_modes(lambda: ctx.field_eval(d, w, tp))
        np.testing.assert_array_equal(a, b, err_msg=w)
    p = rng.normal(size=(n, 3)) * K.m_pr * 1e4
    gm = np.sqrt(K.m_pr ** 2 + np.sum(p * p, axis=1) / float(K.c) ** 2)
    ty = np.c_[tp, p]
    kk = np.c_[np.full(n, K.m_pr), np.full(n, K.e), gm]
    a, b = _both_modes(lambda: ctx.rhs_eval(d, N.MODE_PARTICLE, ty, kk))
    np.testing.assert_array_equal(a, b)
    ty5 = np.c_[tp, p[:, 0]]
    k4 = np.c_[np.full(n, K.m_pr), np.full(n, K.e), np.full(n, 1e-16), np.full(n, 1e4)]
    a, b = _both_modes(lambda: ctx.rhs_eval(d, N.MODE_GC, ty5, k4))
    np.testing.assert_array_equal(a, b)
    a, b = _both_modes(lambda: ctx.particle_gc(d, ty, K.m_pr, K.e))
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_compiled_every_opcode():
    """A field that uses every operation the tracer records (the library
    functions included): compiled values equal the interpreted ones bit for
    bit, and the host evaluation of the same program to a few ulp."""
    from rapt_amd import fieldexpr as FE

    class AllOps(rapt.fields._Field):
        def __init__(self):
            rapt.fields._Field.__init__(self)
            self.static = False

        def B(self, tpos):
            t, x, y, z = tpos
            u = x / 1e7
            w = y / 1e7
            s = z / 1e7
            a = np.tan(0.3 * u) + np.sinh(0.2 * w) + np.cosh(0.1 * s) + np.arctan(u)
            b = np.arcsin(0.5 * np.tanh(w)) + np.arccos(0.4 * np.tanh(s)) + np.arctan2(w, u)
            c = np.log10(1.0 + u * u) + np.log(2.0 + w * w) + np.hypot(u, s) + abs(s) ** 1.5
            d = np.maximum(u, w) - np.minimum(w, s) + np.sign(s) + np.clip(u, -0.5, 0.5)
            return 1e-9 * np.array([a + np.exp(-t / 10.0), b * np.sin(t) + c, d * np.cos(u) +
                                    np.sqrt(1.0 + s * s) - (-u)])

        def E(self, tpos):
            t, x, y, z = tpos
            return 1e-4 * np.array([np.sin(x / 1e7), 0.0 * y, np.cos(z / 1e7) ** 2])
    f = AllOps()
    d = f.descriptor()
    ops = set(int(o) for o in d["expr"].code[:, 0])
    assert ops == set(range(27)), sorted(set(range(27)) - ops)
    rng = np.random.default_rng(43)
    tp = np.c_[rng.uniform(0, 5, 64), rng.uniform(-2, 2, (64, 3)) * 1e7]
    tp[:2, 3] = [0.0, -0.0]
    ctx = N.context()
    for w in ("B", "E", "gradB", "curlb", "dbdt"):
        a, b = _both_modes(lambda: ctx.field_eval(d, w, tp))
        np.testing.assert_array_equal(a, b, err_msg=w)
    got = ctx.field_eval(d, "B", tp)
    for i in range(len(tp)):
        want = FE.evaluate(d["expr"], tp[i])[0]
        np.testing.assert_allclose(got[i], want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
