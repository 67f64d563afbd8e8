import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# cahn_hilliard_options_conserve 2: the golden vectors and logs come from the
# reference run on one thread, so the tests take the library's site-order
# sum (bit-identical with that run); the default compensated tree sum has
# tests of its own (test_gpu_parity.py), which switch it on explicitly.
os.environ.setdefault("LDW_CONSERVE2_SITE_ORDER", "1")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


GOLDEN = os.path.join(REPO, "tests", "golden")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))
    return load
