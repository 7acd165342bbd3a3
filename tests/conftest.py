import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and librapt_hip.so")


GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))
    return load


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def user_fields():
    """User-written field classes (tests/golden/userfields.py) built on
    rapt_amd.fields, the same class bodies the golden fixtures ran under the
    reference."""
    import importlib.util

    import rapt_amd.fields as F
    spec = importlib.util.spec_from_file_location("userfields",
                                                  os.path.join(GOLDEN, "userfields.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.classes(F)
