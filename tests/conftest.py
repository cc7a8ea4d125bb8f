import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkcore on HIP)")


def _announce(item):
    # name the case on stdout before any GPU work so a lost box still says
    # which test was running
    print("\n[case] %s" % item.nodeid, flush=True)


@pytest.hookimpl(tryfirst=True)
def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu"):
        _announce(item)
