"""Shared test set-up: the gpu marker and paths to the package, the oracle
and the golden fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = os.path.join(ROOT, "cases")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case_dir(name):
    return os.path.join(CASES, name)


def rel_l2(a, b):
    """Relative L2 error of a against b, per variable group (rho, momentum, energy)."""
    out = {}
    for key, sl in (("rho", slice(0, 1)), ("mom", slice(1, 4)), ("energy", slice(4, 5))):
        num = np.sqrt(np.sum((a[sl] - b[sl]) ** 2))
        den = np.sqrt(np.sum(b[sl] ** 2))
        out[key] = float(num / den) if den > 0 else float(num)
    return out


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.library()
    return oracle
