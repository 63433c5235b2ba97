"""Test configuration: the `gpu` marker, import paths, shared fixtures.

CPU tests (`-m "not gpu"`) check the oracle against the reference's own
known answers and closed forms, the host-side logic, and that the C-ABI
library loads and exports every declared symbol.  GPU tests (`-m gpu`) are
the parity tests proper: every result comes from libbvht_hip.so and is
compared with the oracle on identical inputs.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (built with its Makefile if the .so is missing)."""
    lib_path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib_path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    """A bvht context on device 0 (GPU tests only)."""
    import bvhtracer_amd as B

    c = B.Context(0)
    yield c
    c.close()
