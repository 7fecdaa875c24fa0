"""Test configuration: markers, import paths and shared helpers."""
import math
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "solstice-solver_amd", "python"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def within(a, b, k=3.0, floor=0.0):
    """|E_a - E_b| <= k * sqrt(SE_a^2 + SE_b^2) + floor for (E, V, SE) triples."""
    tol = k * math.sqrt(a[2] ** 2 + b[2] ** 2) + floor
    return abs(a[0] - b[0]) <= tol, tol


@pytest.fixture(scope="session")
def gpu():
    """Skips when no GPU is visible; the product library is loaded either way."""
    import ssol_ctypes
    lib = ssol_ctypes.lib_hip()
    lib.ssolh_device_count.restype = __import__("ctypes").c_int
    if lib.ssolh_device_count() < 1:
        pytest.skip("no GPU visible")
    return lib
