"""Test configuration: import paths and the ``gpu`` marker.

GPU tests (marked ``gpu``) call the HIP library through its C ABI and compare
against the CPU float64 oracle in ``oracle/``.  CPU tests pin the oracle and
check host logic and the library's exported symbols.
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "torchff-lib_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: large CPU oracle case")


def pytest_collection_modifyitems(config, items):
    import torch
    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if item.get_closest_marker("gpu"):
            item.add_marker(skip)
