"""User-written fields (the reference's way to define a field: subclass
``_Field`` and override ``B``/``E`` in Python, rapt/fields.py:5-74) traced
into register programs (rapt_amd/fieldexpr.py), on the CPU:

- the traced program evaluated on the host and by the oracle's interpreter
  against the user's own Python methods and the reference's fixtures
  (tests/golden/userfields.npz, written by running the same class bodies
  under the reference): bit-identical where the field uses arithmetic and
  pow (which numpy's float64 scalars take from the C library), within a few
  ulps where it calls numpy's own tanh/sin/exp/cosh loops;
- the oracle's Particle and GuidingCenter runs in those fields against the
  reference's trajectories;
- a user restatement of a built-in model against that model;
- what cannot be traced raises NotImplementedError, and attribute changes
  are picked up at the next descriptor().
"""
import numpy as np
import pytest

import rapt_amd.fields as F
from rapt_amd import constants as K
from rapt_amd import fieldexpr as FE

EXACT = ("ChargedDipole", "TwinDipole")  # arithmetic and pow only


def _field(user_fields, name):
    return user_fields[name]()


def test_descriptor_is_a_program(user_fields):
    for name, cls in user_fields.items():
        d = cls().descriptor()
        assert d["model"] == F.FIELD_EXPR and isinstance(d["expr"], FE.Program), name
        prog = d["expr"]
        assert prog.code.shape == (prog.nb + prog.ne, 4) and prog.nb >= 3
        assert np.all(prog.code[:, 1] >= 4) and np.all(prog.code[:, 1] < FE.NREG)
    # a static field without E carries no E program
    assert user_fields["TwinDipole"]().descriptor()["expr"].ne == 0
    assert user_fields["ChargedDipole"]().descriptor()["static"] is False


@pytest.mark.parametrize("name", ["ChargedDipole", "TwinDipole", "BreathingSheet"])
def test_program_matches_python_methods(user_fields, name):
    """The program at 64 random points against the user's own methods."""
    f = _field(user_fields, name)
    prog = f.descriptor()["expr"]
    rng = np.random.default_rng(5)
    scale = 1.0 if name == "ChargedDipole" else K.Re
    pts = np.c_[rng.uniform(0, 20, 64), rng.uniform(-6, 6, (64, 3)) * scale]
    for tp in pts:
        b, e = FE.evaluate(prog, tp)
        hb, he = np.asarray(f.B(tp.copy()), float), np.asarray(f.E(tp.copy()), float)
        if name in EXACT:
            np.testing.assert_array_equal(b, hb)
            np.testing.assert_array_equal(e, he)
        else:
            np.testing.assert_allclose(b, hb, rtol=1e-15, atol=1e-15 * np.max(np.abs(hb)))
            np.testing.assert_allclose(e, he, rtol=1e-15, atol=1e-15 * np.max(np.abs(he)))


@pytest.mark.parametrize("name", ["ChargedDipole", "TwinDipole", "BreathingSheet"])
def test_oracle_interpreter_matches_reference(oracle, golden, user_fields, name):
    """The oracle's interpreter (fields and the _Field derivatives built on
    them) against the reference's values for the same class body."""
    g = golden("userfields")
    fo = oracle.field_from_desc(_field(user_fields, name).descriptor())
    tp = g[name + "/tpos"]
    for w in ("B", "E", "magB", "gradB", "curvature", "dBdt", "lengthscale"):
        key = name + "/" + w
        if key not in g.files:
            continue
        got, ref = oracle.field_eval(fo, w, tp), g[key]
        if name in EXACT:
            np.testing.assert_array_equal(got, ref, err_msg=key)
        else:
            # numpy's tanh/exp loops differ from libm by an ulp; the finite
            # differences of gradB and dBdt amplify it by scale/step
            tol = 1e-15 if w in ("B", "E", "magB") else 1e-10
            np.testing.assert_allclose(got, ref, rtol=tol, atol=tol * np.max(np.abs(ref)),
                                       err_msg=key)


def _rel(got, ref):
    sc = np.max(np.abs(ref), axis=0)
    sc[sc == 0] = 1.0
    return np.max(np.abs(got - ref) / sc)


def test_oracle_particle_in_charged_dipole(oracle, golden, user_fields):
    """The reference's new-field example: a 1 eV proton at 5 m in a charged
    dipole, advance(4e-4): the oracle in the traced field is identical to the
    reference row by row."""
    g = golden("userfields")
    f = user_fields["ChargedDipole"](Q=1e-6)
    row0 = g["cd_particle/row0"]
    rows, info = oracle.particle_advance(oracle.field_from_desc(f.descriptor()),
                                         oracle.particle_params(), row0, K.m_pr, K.e, 4e-4)
    np.testing.assert_array_equal(np.vstack([row0[None], rows]), g["cd_particle/traj"])
    assert info["tcur"] == g["cd_particle/tcur"][0]


@pytest.mark.parametrize("name,cls,tol", [("twin_particle", "TwinDipole", 0),
                                          ("breathing_particle", "BreathingSheet", 1e-13)])
def test_oracle_particle_user_fields(oracle, golden, user_fields, name, cls, tol):
    g = golden("userfields")
    ref = g[name + "/traj"]
    m, q, delta = g[name + "/in"][6:9]
    rows, info = oracle.particle_advance(
        oracle.field_from_desc(user_fields[cls]().descriptor()), oracle.particle_params(),
        ref[0], m, q, delta)
    got = np.vstack([ref[:1], rows])
    assert got.shape == ref.shape
    if tol == 0:
        np.testing.assert_array_equal(got, ref)
    else:
        assert _rel(got, ref) < tol
    assert info["tcur"] == g[name + "/tcur"][0]


@pytest.mark.parametrize("name,make,tol", [
    ("cd_gc", lambda U: U["ChargedDipole"](Q=1e-8), 0),
    ("twin_gc", lambda U: U["TwinDipole"](), 0),
    ("breathing_gc", lambda U: U["BreathingSheet"](), 1e-13),
])
def test_oracle_gc_user_fields(oracle, golden, user_fields, name, make, tol):
    """GuidingCenter.advance in user fields, including one whose B changes in
    time (dB/dt by central differences in t) and one with E."""
    g = golden("userfields")
    ref = g[name + "/traj"]
    v, pa, m, q, mu, gdt, delta = g[name + "/in"][3:]
    rows, info = oracle.gc_advance(oracle.field_from_desc(make(user_fields).descriptor()),
                                   oracle.gc_params(), "TaoChanBrizardEOM", ref[0], m, q, mu, v,
                                   gdt, delta)
    got = np.vstack([ref[:1], rows])
    assert got.shape == ref.shape
    if tol == 0:
        np.testing.assert_array_equal(got, ref)
    else:
        assert _rel(got, ref) < tol
    assert info["tcur"] == g[name + "/tcur"][0]


def test_user_restatement_of_builtin_model(oracle, user_fields):
    """A user's two-dipole field against the built-in DoubleDipole with the
    same constants: the same field to rounding (the built-in model folds
    its coefficient differently)."""
    fu = oracle.field_from_desc(user_fields["TwinDipole"]().descriptor())
    fb = oracle.field_from_desc(F.DoubleDipole().descriptor())
    rng = np.random.default_rng(9)
    pts = np.c_[np.zeros(40), rng.uniform(-8, 8, (40, 3)) * K.Re]
    pts[:, 1] += 2.5 * K.Re * np.sign(pts[:, 1])
    for w in ("B", "gradB", "curvature"):
        a, b = oracle.field_eval(fu, w, pts), oracle.field_eval(fb, w, pts)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b)), err_msg=w)


def test_untraceable_fields_raise():
    class Branchy(F._Field):
        def B(self, tpos):
            return np.array([0.0, 0.0, 1e-7 if tpos[3] > 0 else -1e-7])

    class Floaty(F._Field):
        def B(self, tpos):
            return np.array([0.0, 0.0, float(tpos[1]) * 1e-9])

    class Lookup(F._Field):
        def B(self, tpos):
            return np.array([0.0, 0.0, np.interp(tpos[1], [0, 1], [0, 1])])

    for cls in (Branchy, Floaty, Lookup):
        with pytest.raises(NotImplementedError, match="cannot run on the device"):
            cls().descriptor()


def test_attribute_changes_are_retraced(user_fields):
    f = user_fields["ChargedDipole"](Q=1e-6)
    k1 = f.descriptor()["expr"].key()
    f.Q = 2e-6
    p2 = f.descriptor()["expr"]
    assert p2.key() != k1
    tp = np.array([0.0, 3.0, 1.0, 0.5])
    np.testing.assert_array_equal(FE.evaluate(p2, tp)[1], f.E(tp))


def test_constant_and_zero_fields(user_fields):
    class Uniform(F._Field):
        def B(self, tpos):
            return np.array([0.0, 0.0, 5e-8])
    d = Uniform().descriptor()
    assert d["model"] == F.FIELD_EXPR and d["expr"].ne == 0
    b, e = FE.evaluate(d["expr"], [1.0, 2.0, 3.0, 4.0])
    np.testing.assert_array_equal(b, [0.0, 0.0, 5e-8])
    np.testing.assert_array_equal(e, [0.0, 0.0, 0.0])


class _SelectField(F._Field):
    """A user field written with numpy's selection functions and vector
    helpers: a saturating sheet (np.clip), a one-sided normal component
    (np.maximum / np.minimum), a sign flip across y = 0, |r| from
    np.linalg.norm and an E from np.cross."""

    def __init__(self):
        F._Field.__init__(self)
        self.static = False
        self.gradientstepsize = K.Re / 1000

    def B(self, tpos):
        r = np.asarray(tpos[1:])
        z = tpos[3] / (0.5 * K.Re)
        bx = 20e-9 * np.clip(z, -1.0, 1.0)
        bz = np.maximum(1e-9 * np.sign(tpos[2]), 2e-10) + np.minimum(tpos[1] / K.Re, 0.0) * 1e-10
        return np.array([bx, 3e-10 * np.linalg.norm(r) / K.Re, bz])

    def E(self, tpos):
        return np.cross(np.asarray(tpos[1:]) / K.Re, np.array([0.0, 0.0, 1e-4]))


def test_selection_ops_and_vector_helpers(oracle):
    f = _SelectField()
    prog = f.descriptor()["expr"]
    rng = np.random.default_rng(21)
    pts = np.c_[rng.uniform(0, 5, 64), rng.uniform(-3, 3, (64, 3)) * K.Re]
    pts[:4, 2] = [0.0, -0.0, 1e-300, -1e-300]  # sign at and near zero
    fo = oracle.field_from_desc(f.descriptor())
    ob, oe = oracle.field_eval(fo, "B", pts), oracle.field_eval(fo, "E", pts)
    for i, tp in enumerate(pts):
        b, e = FE.evaluate(prog, tp)
        hb, he = f.B(tp.copy()), f.E(tp.copy())
        # np.linalg.norm of a float array goes through BLAS, which fuses the
        # products into the sum; traced, they are separate: an ulp apart
        np.testing.assert_array_equal(b[[0, 2]], hb[[0, 2]])
        np.testing.assert_allclose(b[1], hb[1], rtol=4e-16, atol=0)
        np.testing.assert_array_equal(e, he)
        np.testing.assert_array_equal(ob[i], b)
        np.testing.assert_array_equal(oe[i], e)


def test_selection_ops_follow_numpy_on_special_values():
    """maximum/minimum propagate a NaN from either side and return the second
    operand on ties (so the sign of a zero result is the second zero's),
    sign(+-0) = +0, as numpy's ufuncs do (the device and the oracle implement
    the same tests)."""
    class M(F._Field):
        def B(self, tpos):
            return np.array([np.maximum(tpos[1], tpos[2]), np.minimum(tpos[1], tpos[2]),
                             np.sign(tpos[3])])
    prog = M().descriptor()["expr"]
    for x, y, z in [(np.nan, 1.0, 0.0), (1.0, np.nan, -0.0), (-0.0, 0.0, np.nan),
                    (0.0, -0.0, -2.0), (np.inf, -np.inf, 3.0)]:
        tp = np.array([0.0, x, y, z])
        got = FE.evaluate(prog, tp)[0]
        want = np.array([np.maximum(x, y), np.minimum(x, y), np.sign(z)])
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(np.signbit(got), np.signbit(want))


def test_user_field_program_compiles_for_gfx950(user_fields):
    """The run-time compiled tracer kernels of a traced program (the
    reference's charged-dipole example: B and E programs) build from the
    headers embedded in the library, on the CPU (no device involved); a bad
    opcode is refused with the library's message."""
    from rapt_amd import _native as N
    prog = user_fields["ChargedDipole"]().descriptor()["expr"]
    # the one-lane Particle kernel and the half-bounce kernel (the others
    # build from the same headers; all of them take a few minutes here)
    msg = N.expr_compile(prog, kernels=(1 << 0) | (1 << 12))
    assert "2 kernels" in msg, msg
    code = prog.code.copy()
    code[0, 0] = 99
    with pytest.raises(RuntimeError, match="rapt_field_expr_compile"):
        N.expr_compile(FE.Program(code, prog.consts, prog.nb, prog.ne), kernels=1)
