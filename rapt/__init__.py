"""``import rapt`` for code written against mkozturk/rapt.

This package is the same objects as ``rapt_amd`` under the reference's module
names (rapt/__init__.py:1-42): the constants, the ``params`` dict (one dict,
read by the GPU tracers), ``Adiabatic``/``NonAdiabatic``, the tracer classes,
and the submodules ``rapt.fields``, ``rapt.utils``, ``rapt.rkf``,
``rapt.fieldline``, ``rapt.flutils``, ``rapt.Particle``, ``rapt.GuidingCenter``,
``rapt.Adaptive`` and ``rapt.BounceCenter``, so ``from rapt import fields``,
``import rapt.fields as rf`` and ``from rapt.Particle import Particle`` work
unchanged.
"""
import importlib
import sys

import rapt_amd as _impl
from rapt_amd import (Adaptive, Adiabatic, B0, BounceCenter, Fieldline, GuidingCenter,  # noqa: F401
                      NonAdiabatic, Particle, Re, batch, c, e, eye, fieldline, fields, flutils,
                      gradI, halfbouncepath, m_el, m_pr, params, re, rkf, utils)

for _name in ("fields", "utils", "rkf", "fieldline", "flutils", "Particle", "GuidingCenter",
              "Adaptive", "BounceCenter", "batch", "constants", "surrogate"):
    sys.modules[__name__ + "." + _name] = importlib.import_module("rapt_amd." + _name)
del _name
