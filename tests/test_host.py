"""CPU tests of the host side: the C-ABI library loads and exports every
symbol of include/rapt_hip.h (no GPU call), the product's field/host helpers
match the reference's fixtures, the seeded workloads are stable."""
import ctypes
import os
import re

import numpy as np
import pytest

import rapt_amd as rapt
import rapt_amd.fields as F
from rapt_amd import _native as N
from rapt_amd import constants as K
from rapt_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "rapt_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(rapt_\w+)\(", txt, re.M)))


def test_library_exports_header_symbols():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("librapt_hip.so not built")
    L = ctypes.CDLL(N.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(L, s), s


def test_struct_layouts_match_header():
    # field descriptor: 2 int32 + 2 double + 16 double
    assert ctypes.sizeof(N.FieldDesc) == 8 + 16 + 128
    assert ctypes.sizeof(N.TraceDesc) == 8 * 4 + 8 * 18 + 8 + 8 + 8 + 8
    assert ctypes.sizeof(N.Counters) == 7 * 8


def test_descriptor_traces_python_overrides():
    class MyField(F._Field):
        def B(self, tpos):
            return np.array([0.0, 0.0, 1e-7])
    assert MyField().descriptor()["model"] == F.FIELD_EXPR
    d = type("E2", (F.UniformBz,), {"E": lambda self, t: np.ones(3)})().descriptor()
    assert d["model"] == F.FIELD_EXPR and d["expr"].ne > 0

    class Branchy(F._Field):
        def B(self, tpos):
            return np.array([0.0, 0.0, 1.0 if tpos[3] > 0 else 2.0])
    with pytest.raises(NotImplementedError):
        Branchy().descriptor()
    d = F.DoubleDipole().descriptor()
    assert d["model"] == 2 and d["gstep"] == K.Re / 1000


@pytest.mark.parametrize("name,make", [
    ("EarthDipole", F.EarthDipole), ("DoubleDipole", F.DoubleDipole),
    ("VarEarthDipole", lambda: F.VarEarthDipole(amp=0.2, period=7.0)),
    ("Parabolic", lambda: F.Parabolic(B0=10.0, Bn=1.0, d=0.2)),
    ("DipolePlusUniform", F.DipolePlusUniform)])
def test_host_field_api_matches_reference(golden, name, make):
    g = golden("fields")
    f = make()
    for i, tp in enumerate(g[name + "/tpos"]):
        for w in ("B", "magB", "gradB", "curlb", "jacobianB", "lengthscale"):
            got = np.asarray(getattr(f, w)(tp), dtype=float).reshape(-1)
            np.testing.assert_array_equal(got, g[name + "/" + w][i], err_msg=w)


def test_halfbouncepath_matches_reference(golden):
    rows = golden("bounce")["halfbouncepath"]
    for fid, L, pa, ph, Bm, sb in rows[::3]:
        f = F.DoubleDipole() if fid == 0 else F.EarthDipole()
        tp = [0.0, L * K.Re * np.cos(ph), L * K.Re * np.sin(ph), 0.0]
        assert rapt.flutils.halfbouncepath(tp, f, Bm) == pytest.approx(sb, rel=1e-12)


def test_host_rkf_matches_reference(golden):
    R = golden("rkf")
    fn = lambda x, t: np.array([x[1], -x[0]])
    for j in range(4):
        tol, hmax, hmin, raised = R["osc%d/cfg" % j]
        if raised:
            with pytest.raises(RuntimeError):
                rapt.rkf.rkf(fn, 0.0, 3.0, [1.0, 0.0], tol, hmax, hmin)
            continue
        T, X = rapt.rkf.rkf(fn, 0.0, 3.0, [1.0, 0.0], tol, hmax, hmin)
        np.testing.assert_array_equal(T, R["osc%d/T" % j])
        np.testing.assert_array_equal(X, R["osc%d/X" % j])


def test_host_utils_match_reference(golden):
    U = golden("utils")
    f = F.DoubleDipole()
    for i in range(len(U["tpos"])):
        t, pos = U["tpos"][i, 0], U["tpos"][i, 1:]
        gc, vp, spd = rapt.utils.guidingcenter(t, pos, U["vel"][i], f, K.m_el, -K.e)
        np.testing.assert_array_equal(np.r_[gc, vp, spd], U["guidingcenter"][i])
    for k, m_el, m_pr in U["speedfromKE"]:
        assert rapt.utils.speedfromKE(k, K.m_el) == m_el
        assert rapt.utils.speedfromKE(k, K.m_pr) == m_pr


def test_workloads_are_prefix_stable():
    a = W.dipole_protons_1mev(n=16)
    b = W.dipole_protons_1mev(n=1024)
    np.testing.assert_array_equal(a["pos"], b["pos"][:16])
    np.testing.assert_array_equal(a["vel"], b["vel"][:16])
    r = np.linalg.norm(b["pos"], axis=1) / K.Re
    assert r.min() >= 4 and r.max() <= 7
    v = np.linalg.norm(b["vel"], axis=1)
    np.testing.assert_allclose(v, rapt.utils.speedfromKE(1e6, K.m_pr), rtol=1e-14)
    # pitch angle to +z between 10 and 90 degrees
    pa = np.degrees(np.arccos(b["vel"][:, 2] / v))
    assert pa.min() >= 10 - 1e-9 and pa.max() <= 90 + 1e-9
    c2 = W.doubledipole_electrons_gc(n=4096)
    assert c2["L"].min() >= 3 and c2["L"].max() <= 8
    assert c2["logK"].min() >= 4 and c2["logK"].max() <= 7


def test_gpu_path_fails_loudly_without_device():
    """No CPU fallback: without a HIP device the tracer raises."""
    try:
        N.context()
    except RuntimeError:
        pass
    else:
        pytest.skip("a HIP device is visible")
    p = rapt.Particle([5 * K.Re, 0, 0], [0, 1e6, 1e6], 0, K.m_pr, K.e, F.EarthDipole())
    with pytest.raises(RuntimeError):
        p.advance(0.1)
