import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "eol-cloth_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (gfx950) GPU")


GRAV = np.array([0.0, 0.0, -9.8])
H = 5e-3


@pytest.fixture(scope="session")
def ctx():
    import eolc
    c = eolc.Context(0)
    yield c
    c.close()


def rel_err(a, b):
    """max |a - b| / max |b| (the scale is the whole array, not per entry)."""
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))
