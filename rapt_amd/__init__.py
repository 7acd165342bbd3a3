"""rapt_amd: MI355X-native batched particle tracer with RAPT's Python API.

``import rapt_amd as rapt`` gives the names of rapt (rapt/__init__.py:1-42):
constants, the ``params`` dict, the ``Adiabatic``/``NonAdiabatic``
control-flow exceptions, the field models and the tracers.  ``advance()``
runs on the GPU through the C-ABI library ``librapt_hip.so``; there is no CPU
integration path.
"""
from .constants import e, m_pr, m_el, c, B0, re, Re


class Adiabatic(Exception):
    """Raised by Particle.advance when the motion becomes adiabatic."""


class NonAdiabatic(Exception):
    """Raised by GuidingCenter.advance when the motion becomes non-adiabatic."""


# rapt/__init__.py:21-34; read at call time by the tracers
params = {
    "cyclotronresolution": 10,
    "Ptimestep": 0,
    "bounceresolution": 10,
    "GCtimestep": 0,
    "BCtimestep": 0.1,
    "solvertolerances": (1.49012e-8, 1.49012e-8),
    "fieldlineresolution": 50,
    "flsolver": "rkf",
    "eyegradientstep": 0.03 * Re,
    "epss": 5e-2,
    "epst": 5e-2,
    "enforce equatorial": False,
}

from . import fields, utils, rkf, fieldline, flutils  # noqa: E402
from .Particle import Particle  # noqa: E402
from .GuidingCenter import GuidingCenter  # noqa: E402
from .Adaptive import Adaptive  # noqa: E402
from .BounceCenter import BounceCenter  # noqa: E402
from .fieldline import Fieldline  # noqa: E402
from .flutils import eye, gradI, halfbouncepath  # noqa: E402
from . import batch  # noqa: E402
