"""The smoke entry point (__graft_entry__.smoke) as a GPU test: a 16^3
binary fluid, 5 steps through the C ABI, bit-exact with the CPU oracle.
Run here too so that a change on either side of the comparison (the GPU
path, the oracle, or their parameters) fails in the test suite."""

import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_smoke_entry_point():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import __graft_entry__ as entry
    entry.smoke()
