"""Shared test fixtures.

GPU tests are marked ``gpu`` and run only on a machine with an MI355X; the
rest (oracle pinning, host logic, library exports) run on the CPU.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmgad.so")


@pytest.fixture(scope="session")
def hipdev():
    """The HIP device must exist: a GPU test never passes without running the HIP path."""
    from multigpu_advectiondiffusion_amd import hip

    n = hip.device_count()
    assert n > 0, "no HIP device visible: GPU tests must run on the MI355X"
    return hip


def pytest_runtest_setup(item):
    # name the case before any GPU step so a crash log says which one ran
    if item.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-case] {item.nodeid}", flush=True)
