import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def allclose(a, b, rtol, atol=1e-8):
    """unit_test/Test_Utils.hpp:19-45: every |a-b| <= atol + rtol*|b|."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    return bool(np.all(np.abs(a - b) <= atol + rtol * np.abs(b)))


def rel_err(a, b):
    """max|a-b| / max|b| (a relative error that does not blow up at zeros)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O
