import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _announce(item):
    # name the case before any GPU work so a crash can be attributed
    print("\n[case] %s" % item.nodeid, flush=True)


def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu"):
        _announce(item)
