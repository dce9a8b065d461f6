import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_golden(name):
    """Fixture written by tests/golden/make_golden.py from the reference's own routines."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return {k: z[k] for k in z.files}


def grid_coef(z):
    """26-double coefficient block of x2g_grid_create from a golden fixture."""
    s = z["scal"]  # hni, hmu, diag, ovforb, ovfcoul, ovfexch, r, rinf
    return np.concatenate([z["coef"], [s[2], s[0], s[1], s[6]]])


@pytest.fixture(params=["h2_61x91", "n2_61x91"])
def golden(request):
    return load_golden(request.param)
