import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: longer CPU run")


def assert_kat(actual, expected, tol, what=""):
    """The reference's ``Approx(expected).epsilon(tol)`` (Catch v1.12, scale 1):
    |a - e| < tol * (1 + max(|a|, |e|)), elementwise."""
    a = np.asarray(actual, dtype=np.float64)
    e = np.broadcast_to(np.asarray(expected, dtype=np.float64), a.shape)
    ok = np.abs(a - e) < tol * (1.0 + np.maximum(np.abs(a), np.abs(e)))
    if not np.all(ok):
        raise AssertionError("%s: KAT mismatch at %s\nactual   %s\nexpected %s" % (
            what, np.argwhere(~ok).tolist(), a[~ok], e[~ok]))


@pytest.fixture(scope="session")
def kats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)
