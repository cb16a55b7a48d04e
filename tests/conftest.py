import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def rel_l2(a, b):
    """Relative l2 error |a-b| / |b| (the reference tests' comperr)."""
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.oracle_lib()


@pytest.fixture(scope="session")
def ref():
    import oracle
    return oracle.ref_lib()


@pytest.fixture(scope="session")
def eng():
    import fmm2d_c_translation_amd as F
    return F.lib()
