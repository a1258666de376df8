"""pytest configuration: the `gpu` marker and import paths.

CPU-only tests (-m "not gpu") cover the oracle against the compiled reference
and its golden fixtures, host-side code, and that libcylmhd.so loads and
exports every symbol of include/cylmhd.h.  GPU tests (-m gpu) are the parity
tests proper: they call the HIP library through its C ABI and compare with
the oracle on the same inputs.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "171026-1-3dhlld-cyli-gal_amd")
for p in (PKG, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU in this environment")
    return True
