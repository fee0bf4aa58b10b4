import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtpgpu.so)")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.fixture
def oracle_libm():
    """Oracle with libm sin/cos, as the reference binary computes."""
    from oracle import tpref
    tpref.set_math(True)
    yield tpref
    tpref.set_math(True)


@pytest.fixture
def oracle_tp():
    """Oracle with the shared tp_sincos, as the GPU kernels compute."""
    from oracle import tpref
    tpref.set_math(False)
    yield tpref
    tpref.set_math(True)
