import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


def _print_case(item):
    # Name the case before any device work so a crash still identifies it.
    print(f"\n[case] {item.nodeid}", flush=True)


def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu"):
        _print_case(item)
