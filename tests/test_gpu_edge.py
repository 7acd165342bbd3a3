"""Edge cases of the batched entry points: empty batches, one particle per
layout (a mixed launch of one particle has no pair blocks), and the
diagnostics on empty inputs."""
import numpy as np
import pytest

import rapt_amd as rapt
from rapt_amd import _native as N
from rapt_amd import batch
from rapt_amd import constants as K
from rapt_amd import fields as F

pytestmark = pytest.mark.gpu


def test_empty_batches_return_empty_results():
    f = F.EarthDipole()
    p = batch.advance_particles(f, np.zeros((0, 7)), K.m_pr, K.e, 1.0)
    assert len(p.rows) == 0 and p.status.shape == (0,)
    g = batch.advance_gcs(f, np.zeros((0, 5)), K.m_el, -K.e, np.zeros(0), np.zeros(0), 0.1, 1.0)
    assert len(g.rows) == 0 and g.status.shape == (0,)
    ctx = N.context()
    out, st = ctx.particle_gc(f.descriptor(), np.zeros((0, 7)), K.m_pr, K.e)
    assert out.shape == (0, 6) and st.shape == (0,)
    I, st = ctx.eye(f.descriptor(), np.zeros((0, 4)), np.zeros(0), 50)
    assert I.shape == (0,) and st.shape == (0,)
    B = ctx.field_eval(f.descriptor(), "B", np.zeros((0, 4)))
    assert B.shape == (0, 3)


@pytest.mark.parametrize("lpp", [1, 2, 3, 4, 5, 8])
def test_single_guiding_centre_every_layout(lpp):
    f = F.DoubleDipole()
    v = rapt.utils.speedfromKE(1e5, K.m_el)
    gc = rapt.GuidingCenter([5 * K.Re, 0, 0], v, 40.0, None, 0, K.m_el, -K.e, f)
    last = gc.trajectory[-1:].copy()
    prm = dict(rapt.params, solvertolerances=(1e-8, 1e-8))
    ref = batch.advance_gcs(f, last, K.m_el, -K.e, gc.mu, v, 0.05, 0.5, params=prm,
                            lanes_per_particle=1)
    got = batch.advance_gcs(f, last, K.m_el, -K.e, gc.mu, v, 0.05, 0.5, params=prm,
                            lanes_per_particle=lpp)
    np.testing.assert_array_equal(got.status, ref.status)
    np.testing.assert_array_equal(got.rows[0], ref.rows[0])
    assert ref.status[0] == 1 and len(ref.rows[0]) >= 10


class _UniformGrid(F.Grid):
    """B = (0, 0, 100 nT) sampled on a small uniform grid: straight field
    lines, zero curvature."""

    def __init__(self):
        F.Grid.__init__(self, ["0"])

    def parsefile(self, filename):
        x = np.linspace(-2.0, 2.0, 5) * K.Re
        sh = (5, 5, 5)
        return dict(time=0.0, x=x, y=x.copy(), z=x.copy(), Bx=np.zeros(sh), By=np.zeros(sh),
                    Bz=np.full(sh, 1e-7), Ex=np.zeros(sh), Ey=np.zeros(sh), Ez=np.zeros(sh))


def test_field_line_integrals_refuse_untraceable_starts():
    """A start with no traceable field line returns status -23 (and NaN) at once
    instead of tracing: the dipole's centre, where |B| = 0 and the curvature is
    NaN, and a uniform field, whose straight line has zero curvature (an
    infinite chunk length).  Non-finite inputs are refused on the host before
    any launch (flutils.halfbouncepath and eye, rapt/flutils.py:65-151,
    254-316, have no field line to trace there either)."""
    ctx = N.context()
    f = F.EarthDipole()
    tpos = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 5 * K.Re, 0.0, 0.0]])
    Bm = np.array([1e-5, 2 * f.magB(tpos[1])])
    sb, st = ctx.halfbounce(f.descriptor(), tpos, Bm, 50)
    assert st[0] == -23 and np.isnan(sb[0])
    assert st[1] == 1 and np.isfinite(sb[1])
    I, st = ctx.eye(f.descriptor(), tpos, Bm, 50)
    assert st[0] == -23 and st[1] == 1
    g = _UniformGrid()
    sb, st = ctx.halfbounce(g.descriptor(), [[0.0, 0.0, 0.0, 0.0]], [2e-7], 50)
    assert st[0] == -23 and np.isnan(sb[0])
    with pytest.raises(RuntimeError, match="non-finite"):
        ctx.halfbounce(f.descriptor(), [[0.0, np.nan, 0.0, 0.0]], [1e-5], 50)
    with pytest.raises(RuntimeError, match="non-finite"):
        ctx.eye(f.descriptor(), tpos[1:], [np.inf], 50)
    with pytest.raises(RuntimeError, match="non-finite"):
        ctx.bounce_center_advance(f.descriptor(), [[0.0, 5 * K.Re, 0.0, np.nan]], 1e-5, 1.0,
                                  1.0, 0.0, 1, 1e-6, 1e-6, 0.03 * K.Re, 50)
