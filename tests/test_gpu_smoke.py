"""The round-end smoke entry point runs as a GPU test too, so that a
change that breaks it shows up with the other GPU tests."""
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_smoke_entry_point():
    sys.path.insert(0, ROOT)
    import __graft_entry__
    print("smoke", flush=True)
    __graft_entry__.smoke()
